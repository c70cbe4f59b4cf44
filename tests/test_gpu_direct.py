"""GPU: launches of a single field sweep full-wave lines (512, 1024 and 2048
cells: 8, 16 and 32 cells per lane) with each lane moving its own cells
between HBM and registers; batches and ragged launches stage the same lines
through coalesced LDS tiles.  Same per-line arithmetic: bitwise equal, for
the plain sweep, the fine propagator and the Parareal chain, against the
oracle within the parity bound as well."""
import numpy as np
import pytest

from ccfv import types as T

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("make", [
    lambda: (T.rotation_problem(512), T.rotation_ic()),
    lambda: T.constant_gaussians(1024, 64, 0.7, -0.4, 2e-3,
                                 [(0.3, 0.6, 0.08, 1.0), (0.7, 0.2, 0.05, 0.6)]),
    lambda: T.constant_gaussians(64, 2048, -1.3, 0.25, 1e-4,
                                 [(0.5, 0.5, 0.1, 0.8)]),
    lambda: (T.rotation_problem(2048), T.rotation_ic()),
])
def test_single_field_equals_batch(gpu, make):
    p, ic = make()
    nt = 4
    with gpu.Context(p) as c:
        c.fill_exact(ic, 0.0)
        u0 = c.get_state()
        c.plain_solve(0.0, 1.0 / 256, nt)          # one field: direct sweeps
        solo = c.get_state()
    with gpu.Batch([p, p]) as b:                     # ragged launch: LDS tiles
        for i in range(2):
            b[i].set_state(u0)
        b.plain([0.0, 0.0], [1.0 / 256, 1.0 / 256], [nt, nt])
        got = [b[i].get_state() for i in range(2)]
    assert np.array_equal(got[0], solo)
    assert np.array_equal(got[1], solo)


def test_direct_2048_matches_oracle(gpu, oracle):
    """Rotation at 2048^2: 32 cells per lane, Thomas factors read from the
    table in memory; two steps against the oracle."""
    p, ic = T.rotation_problem(2048), T.rotation_ic()
    u0 = oracle.exact(p, ic, 0.0)
    with gpu.Context(p) as c:
        c.set_state(u0)
        c.plain_solve(0.0, 2.0 / 4096, 2)
        g = c.get_state()
    o = oracle.plain(p, u0, 0.0, 2.0 / 4096, 2)
    assert np.linalg.norm(g - o) / np.linalg.norm(o) <= 1e-13
