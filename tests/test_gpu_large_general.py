"""GPU: the general sweep (velocity varying along the line: the swirl) at
large grids, where its lanes hold 16 and 32 cells and keep the RHS in
registers, against the CPU oracle on the same inputs.  Tolerance: relative
L2 <= 1e-13 after a few steps and <= 1e-12 for the Parareal result (both
sides FP64, rounding only)."""
import numpy as np
import pytest

from ccfv import types as T

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("n", [1024, 2048])
def test_swirl_plain_large_matches_oracle(gpu, oracle, n):
    p, ic = T.swirl_problem(n, T=5.0, nu=1e-4), T.cosine_bell_ic()
    u0 = oracle.exact(p, ic, 0.0)
    nt, t1 = 4, 4 * 5.0 / 4096
    print(f"swirl {n}^2 plain {nt} steps", flush=True)
    with gpu.Context(p) as c:
        c.set_state(u0)
        c.plain_solve(0.0, t1, nt)
        g = c.get_state()
    o = oracle.plain(p, u0, 0.0, t1, nt)
    assert rel_l2(g, o) <= 1e-13


def test_swirl_parareal_1024_matches_oracle(gpu, oracle):
    """Batched fine propagators of 16-cell lanes (4 slices of 2 steps, two
    iterations) against the oracle's restatement of the reference loop."""
    p, ic = T.swirl_problem(1024, T=5.0, nu=1e-4), T.cosine_bell_ic()
    u0 = oracle.exact(p, ic, 0.0)
    ns, nt, kmax, t1 = 4, 8, 2, 8 * 5.0 / 4096
    r = oracle.parareal(p, u0, 0.0, t1, nt, ns, kmax)
    with gpu.Context(p) as c:
        c.set_state(u0)
        rep = c.parareal_solve(0.0, t1, nt, ns, kmax)
        g = c.get_state()
    assert rep.iterations == kmax
    assert rel_l2(g, r["u"]) <= 1e-12
    for a, b in zip(rep.increments, r["increments"]):
        assert abs(a - b) <= 1e-9 * max(abs(b), 1e-30) + 1e-15
