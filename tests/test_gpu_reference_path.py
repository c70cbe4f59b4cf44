"""GPU checks of the reference path at full size and of the exported
correction, through the C ABI:

* the Parareal correction U = G_new + (F - G_old) (T2S4_Parareal.cu:771-782)
  against its elementwise closed form, bitwise, for whole fields, a length
  that is not a multiple of the 256-wide blocks and a length of zero;
* the 512^2 rotation Parareal (64 slices, 16 fine steps per slice, 8
  iterations, tol = 0: the reference's fixed-iteration loop,
  T2S4_Parareal.cu:1250-1355) against the record the CPU oracle wrote for it
  (tests/golden/make_rotation512.py): the eight increments to 1e-9 relative,
  the distance of U^8(T) to the serial fine solution, and the 16x16 block sums
  of both fields to 1e-12 relative;
* the plain method's spatial order at one fixed small step on 32^2..256^2
  (fourth order: 3.8 < order < 4.2 for every doubling), with the GPU state
  equal to the oracle's at 32^2 within the rounding of 40960 steps.
"""
import json
import math
import os

import numpy as np
import pytest

from ccfv import types as T

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("n", ["all", "ragged", "zero"])
def test_parareal_correction_closed_form(gpu, n):
    p = T.rotation_problem(64)
    cells = p.nx * p.ny
    count = {"all": cells, "ragged": cells - 37, "zero": 0}[n]
    rng = np.random.default_rng(11)
    gn, go, f, u = (rng.normal(size=cells) for _ in range(4))
    print(f"correction n={count}", flush=True)
    ctx = [gpu.Context(p) for _ in range(4)]
    try:
        for c, a in zip(ctx, (gn, go, f, u)):
            c.set_state(a)
        ptr = [c.state_ptr() for c in ctx]
        ctx[0].parareal_correction(ptr[0], ptr[1], ptr[2], ptr[3], count)
        ctx[0].sync()
        got = ctx[3].get_state()
        # the other fields are inputs only
        assert np.array_equal(ctx[0].get_state(), gn)
        assert np.array_equal(ctx[1].get_state(), go)
        assert np.array_equal(ctx[2].get_state(), f)
    finally:
        for c in ctx:
            c.close()
    want = u.copy()
    want[:count] = gn[:count] + (f[:count] - go[:count])
    assert np.array_equal(got, want)


def block_sums(u, n, b):
    return u.reshape(n // b, b, n // b, b).sum(axis=(1, 3)).ravel()


def test_full_size_rotation_parareal_matches_oracle_record(gpu):
    with open(os.path.join(GOLDEN, "rotation512_parareal.json")) as f:
        rec = json.load(f)
    with np.load(os.path.join(GOLDEN, "rotation512_parareal.npz")) as z:
        pb, fb = z["parareal_blocks"].copy(), z["serial_fine_blocks"].copy()
    n, nt, ns, kmax, b = rec["n"], rec["nt_fine"], rec["slices"], rec["kmax"], rec["block"]
    p, ic = T.rotation_problem(n), T.rotation_ic()
    print(f"rotation {n}^2 parareal slices={ns} kmax={kmax}", flush=True)
    with gpu.Context(p) as c:
        c.fill_exact(ic, 0.0)
        u0 = c.get_state()
        c.plain_solve(0.0, 1.0, nt)
        fine = c.get_state()
        c.set_state(u0)
        rep = c.parareal_solve(0.0, 1.0, nt, ns, kmax, rec["tol"])
        u = c.get_state()
    assert rep.iterations == rec["iterations"] == kmax
    for a, want in zip(rep.increments, rec["increments"]):
        assert abs(a - want) <= 1e-9 * abs(want), (a, want)
    assert rel_l2(u, fine) == pytest.approx(rec["rel_l2_vs_serial_fine"], rel=1e-6)
    assert float(np.abs(u - fine).max()) == pytest.approx(rec["max_abs_vs_serial_fine"],
                                                          rel=1e-6)
    assert float(np.linalg.norm(u)) == pytest.approx(rec["norm_parareal"], rel=1e-12)
    assert float(np.linalg.norm(fine)) == pytest.approx(rec["norm_serial_fine"], rel=1e-12)
    assert rel_l2(block_sums(u, n, b), pb) <= 1e-12
    assert rel_l2(block_sums(fine, n, b), fb) <= 1e-12
    # the reference's loop does not reach 1e-10 on this workload: the record
    # and the GPU agree that it stalls near 2e-7 and then grows
    assert not rep.converged and min(rep.increments) > 1e-8


def test_fixed_step_fourth_order_32_to_256(gpu, oracle):
    """Periodic sin*sin, velocity (1, 0.5), nu = 0.01, T = 1, 40960 steps of
    one size on every grid: the error falls 16x per halving of the cell
    width (the time-step part stays below the spatial part up to 256^2)."""
    nt = 40960

    def solve(n):
        p, ic = T.advection_sinsin_problem(n), T.advection_sinsin_ic()
        with gpu.Context(p) as c:
            c.fill_exact(ic, 0.0)
            c.plain_solve(0.0, 1.0, nt)
            return c.get_state(), oracle.exact(p, ic, 1.0)

    ns = (32, 64, 128, 256)
    print(f"fixed-step order, {nt} steps on {ns}", flush=True)
    e = []
    for n in ns:
        u, ex = solve(n)
        e.append(math.sqrt(float(np.mean((u - ex) ** 2))))
    orders = [math.log2(e[i] / e[i + 1]) for i in range(3)]
    assert all(3.8 < o < 4.2 for o in orders), orders
    # the GPU state against the oracle's at 32^2: rounding accumulates over
    # the 40960 steps, so the bound is the 1e-12 of a 1024-step solve scaled
    # by the step count
    p, ic = T.advection_sinsin_problem(32), T.advection_sinsin_ic()
    o = oracle.plain(p, oracle.exact(p, ic, 0.0), 0.0, 1.0, nt)
    assert rel_l2(solve(32)[0], o) <= 1e-12 * nt / 1024
