"""GPU: the allocation-free plain solve (ccfv_plain_solve_noalloc) against
ccfv_plain_solve, through the C ABI: bitwise the same field for every sweep
path (uniform-offset periodic rotation, the general sweep of the swirl, the
closed reference box with its quirks, a non-square constant-velocity
problem, 1024-cell split lines), on two calls in a row with the second on a smaller grid; the
largest grid (2048 x 2048) fits its static scratch; bad arguments are
refused before any launch; crossed feet are reported as by
ccfv_plain_solve, and the next call starts from a cleared flag.  The
caller's fields come from contexts (ccfv_create allocates them; the entry
point under test allocates nothing).  Named to be collected last.
"""
import numpy as np
import pytest

import ccfv
from ccfv import types as T

pytestmark = pytest.mark.gpu


def cases():
    g, gic = T.constant_gaussians(96, 40, -0.7, 0.4, 3e-3,
                                  [(0.3, 0.6, 0.05, 1.0), (0.7, 0.2, 0.08, 0.6)])
    return [
        ("rotation-512", T.rotation_problem(512), T.rotation_ic(), 0.0, 0.25, 256),
        ("swirl-128", T.swirl_problem(128, T=1.0), T.cosine_bell_ic(), 0.0, 0.5, 64),
        ("reference-32", T.reference_problem(32, 0.01), T.reference_ic(), 0.0, 0.3, 24),
        ("gauss-96x40", g, gic, 0.1, 0.4, 30),
        # x lines of 1024 cells: split lines, whose chunk products the
        # device builds in the static scratch with the rest of the table
        ("gauss-1024x64", *T.constant_gaussians(1024, 64, 0.7, -0.4, 2e-3,
                                                [(0.3, 0.6, 0.08, 1.0)]), 0.0, 0.05, 20),
    ]


def reference(gpu, p, ic, t0, t1, nt):
    with gpu.Context(p) as c:
        c.fill_exact(ic, t0)
        u0 = c.get_state()
        c.plain_solve(t0, t1, nt)
        return u0, c.get_state()


def noalloc(gpu, p, u0, t0, t1, nt):
    with gpu.Context(p) as c:
        c.set_state(u0)
        ccfv.plain_solve_noalloc(p, c.state_ptr(), t0, t1, nt)
        return c.get_state()


@pytest.mark.parametrize("first,second", [(0, 2), (0, 1), (1, 2), (3, 2), (4, 0)])
def test_noalloc_equals_plain_solve_twice(gpu, first, second):
    """Two calls in a row on one GPU (its static scratch keeps the first
    call's tables), the second on a smaller grid."""
    cs = cases()
    for i in (first, second):
        name, p, ic, t0, t1, nt = cs[i]
        print(f"noalloc {name} nt={nt}", flush=True)
        u0, want = reference(gpu, p, ic, t0, t1, nt)
        got = noalloc(gpu, p, u0, t0, t1, nt)
        assert np.array_equal(got, want), name


def test_noalloc_refuses_and_reports(gpu):
    # 2048 x 2048 fits; the solve itself is checked on a short run
    p, ic = T.rotation_problem(2048), T.rotation_ic()
    u0, want = reference(gpu, p, ic, 0.0, 1.0 / 512, 2)
    assert np.array_equal(noalloc(gpu, p, u0, 0.0, 1.0 / 512, 2), want)
    # bad arguments fail before any launch
    with gpu.Context(p) as c:
        with pytest.raises(ccfv.CcfvError) as ei:
            ccfv.plain_solve_noalloc(p, c.state_ptr(), 0.0, 1.0, 0)
        assert ei.value.code == T.CCFV_EINVAL
        with pytest.raises(ccfv.CcfvError) as ei:
            ccfv.plain_solve_noalloc(p, 0, 0.0, 1.0, 4)
        assert ei.value.code == T.CCFV_EINVAL
    # crossed feet (swirl, one step of 1 at 64^2) as ccfv_plain_solve
    p, ic = T.swirl_problem(64, T=1.0), T.cosine_bell_ic()
    with gpu.Context(p) as c:
        c.fill_exact(ic, 0.0)
        with pytest.raises(ccfv.CcfvError) as ei:
            ccfv.plain_solve_noalloc(p, c.state_ptr(), 0.0, 1.0, 1)
        assert ei.value.code == T.CCFV_ECFL
    # the flag was cleared: a valid solve afterwards succeeds
    u0, want = reference(gpu, p, ic, 0.0, 0.01, 1)
    assert np.array_equal(noalloc(gpu, p, u0, 0.0, 0.01, 1), want)
