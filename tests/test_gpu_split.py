"""GPU: split lines — uniform-offset sweeps of 1024-cell lines run as two
waves of 8-cell lanes per line, meeting in LDS (compact-operator neighbours,
the chunk maps at the segment ends, the Sherman-Morrison end values).

Checked against the oracle (T2S4_Parareal.cu:785-968 restated) on the
one-field launches (lanes move their own cells), the tile launches of
batches and the coarse chain's epilogue (Parareal), with step sizes whose
lane scans need one, several and all six doubling steps; and the one-field
launches bitwise against the tile launches of a batch."""
import numpy as np
import pytest

from ccfv import types as T

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


CASES = {
    "rotation1024": lambda: (T.rotation_problem(1024), T.rotation_ic()),
    "x1024": lambda: T.constant_gaussians(1024, 64, 0.7, -0.4, 2e-3,
                                          [(0.3, 0.6, 0.08, 1.0), (0.7, 0.2, 0.05, 0.6)]),
    "y1024": lambda: T.constant_gaussians(128, 1024, -0.6, 0.9, 1e-4,
                                          [(0.5, 0.4, 0.06, 0.9)]),
}


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("h", [1.0 / 1024, 1.0 / 64])
def test_split_steps_match_oracle(gpu, oracle, name, h):
    """Two steps: a fine step (one scan step) and a coarse one (several)."""
    p, ic = CASES[name]()
    u0 = oracle.exact(p, ic, 0.0)
    with gpu.Context(p) as c:
        c.set_state(u0)
        c.plain_solve(0.0, 2 * h, 2)
        g = c.get_state()
    o = oracle.plain(p, u0, 0.0, 2 * h, 2)
    assert rel(g, o) <= 1e-13


def test_split_slow_decay_matches_oracle(gpu, oracle):
    """Strong diffusion and a long step: the chunk products decay slowly,
    so the lane scans run all six doubling steps and the segment carries
    fold exactly."""
    p, ic = T.constant_gaussians(1024, 64, 0.3, 0.2, 2e-2,
                                 [(0.5, 0.5, 0.1, 1.0)])
    u0 = oracle.exact(p, ic, 0.0)
    with gpu.Context(p) as c:
        c.set_state(u0)
        c.plain_solve(0.0, 0.05, 1)
        g = c.get_state()
    o = oracle.plain(p, u0, 0.0, 0.05, 1)
    assert rel(g, o) <= 1e-13


@pytest.mark.parametrize("name", ["rotation1024", "y1024"])
def test_split_single_equals_batch(gpu, name):
    """One-field launches (per-lane HBM access) against a ragged batch of
    three (LDS tiles): the same per-line arithmetic, bitwise."""
    p, ic = CASES[name]()
    nt = 3
    with gpu.Context(p) as c:
        c.fill_exact(ic, 0.0)
        u0 = c.get_state()
        c.plain_solve(0.0, 3.0 / 1024, nt)
        solo = c.get_state()
    with gpu.Batch([p, p, p]) as b:
        for i in range(3):
            b[i].set_state(u0)
        b.plain([0.0] * 3, [3.0 / 1024] * 3, [nt] * 3)
        got = [b[i].get_state() for i in range(3)]
    for g in got:
        assert np.array_equal(g, solo)


def test_split_parareal_matches_oracle(gpu, oracle):
    """1024^2 rotation, 8 slices of 4 fine steps, 3 iterations: batched fine
    propagators (panel layout), the coarse chain's fused correction
    epilogue; increments and end state against the oracle's literal loop."""
    p, ic = T.rotation_problem(1024), T.rotation_ic()
    u0 = oracle.exact(p, ic, 0.0)
    t1, nt, ns, kmax = 32.0 / 1024, 32, 8, 3
    with gpu.Context(p) as c:
        c.set_state(u0)
        rep = c.parareal_solve(0.0, t1, nt, ns, kmax, 0.0)
        g = c.get_state()
    o = oracle.parareal(p, u0, 0.0, t1, nt, ns, kmax)
    assert rel(g, o["u"]) <= 1e-12
    inc = np.asarray(rep.increments[:kmax])
    oinc = np.asarray(o["increments"][:kmax])
    assert np.allclose(inc, oinc, rtol=1e-9, atol=0.0)
