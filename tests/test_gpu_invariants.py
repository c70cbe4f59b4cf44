"""GPU: size-independent properties of the scheme at the benchmark sizes,
through the C ABI.

Every operation of a step is linear in the field — the remap, the foot
derivative, the compact operator and the tridiagonal solve; the feet depend
on the velocity only — so
* scaling the initial field by a power of two scales every result by it,
  bit for bit, for the plain solve, the Parareal solve and a ragged batch,
  as long as nothing underflows: scaling changes no rounding of normal
  numbers.  The swirl's cosine bell is exactly zero outside its support and
  the implicit solves spread it into values below the normal range there,
  so for it the scaled solve may differ only by amounts below 1e-280;
* a solve of a u + b w equals a solve(u) + b solve(w) up to rounding;
* on periodic problems the field's sum is conserved (the remap is
  conservative, the compact operator's and the implicit matrix's rows sum
  to one, the diffusive fluxes telescope).
Sizes: the 512^2 rotation of the headline (1024 steps; Parareal with 64
slices of 16 steps, 8 iterations), the 2048^2 swirl (32 steps) and the
seeded 256-problem ragged batch (plain, first 5% of the horizon).
"""
import numpy as np
import pytest

from ccfv import types as T
from ccfv import workloads

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def noise(n, seed):
    return np.random.default_rng(seed).uniform(-1.0, 1.0, n)


def plain(gpu, p, u0, t1, nt):
    with gpu.Context(p) as c:
        c.set_state(u0)
        c.plain_solve(0.0, t1, nt)
        return c.get_state()


@pytest.mark.parametrize("case", ["rotation-512", "swirl-2048"])
def test_plain_solve_is_linear_and_conservative(gpu, case):
    if case == "rotation-512":
        p, ic, t1, nt = T.rotation_problem(512), T.rotation_ic(), 1.0, 1024
    else:
        p, ic, t1, nt = T.swirl_problem(2048, T=5.0, nu=1e-4), T.cosine_bell_ic(), 32 * 5.0 / 4096, 32
    print(f"{case}: {nt} steps", flush=True)
    with gpu.Context(p) as c:
        u = c.exact(ic, 0.0)
    w = noise(u.size, 7)
    su, sw = plain(gpu, p, u, t1, nt), plain(gpu, p, w, t1, nt)
    s4 = plain(gpu, p, 4.0 * u, t1, nt)
    if case == "rotation-512":
        assert np.array_equal(s4, 4.0 * su)
    else:
        assert float(np.max(np.abs(s4 - 4.0 * su))) <= 1e-280
        assert np.array_equal(s4[np.abs(su) > 1e-250], 4.0 * su[np.abs(su) > 1e-250])
    assert np.array_equal(plain(gpu, p, -0.5 * w, t1, nt), -0.5 * sw)
    mix = plain(gpu, p, 3.0 * u - 2.0 * w, t1, nt)
    assert rel_l2(mix, 3.0 * su - 2.0 * sw) <= 1e-12
    for a, s in ((u, su), (w, sw)):
        scale = float(np.sum(np.abs(a)))
        assert abs(float(np.sum(s)) - float(np.sum(a))) <= 1e-12 * scale


def test_parareal_is_linear(gpu):
    p, ic = T.rotation_problem(512), T.rotation_ic()
    print("rotation 512^2 parareal 64 x 16, 8 iterations", flush=True)

    def solve(u0):
        with gpu.Context(p) as c:
            c.set_state(u0)
            rep = c.parareal_solve(0.0, 1.0, 1024, 64, 8, 0.0)
            return c.get_state(), rep

    with gpu.Context(p) as c:
        u = c.exact(ic, 0.0)
    w = noise(u.size, 11)
    pu, ru = solve(u)
    pw, _ = solve(w)
    p2, r2 = solve(2.0 * u)
    assert np.array_equal(p2, 2.0 * pu)
    # the relative increments do not change with the scale
    assert np.array_equal(np.asarray(r2.increments), np.asarray(ru.increments))
    mix, _ = solve(u + 0.25 * w)
    assert rel_l2(mix, pu + 0.25 * pw) <= 1e-11


def test_ragged_batch_scales_bitwise(gpu):
    sol = workloads.ragged_batch(256)
    print("ragged 256 plain, scaled by 2", flush=True)
    t0 = [s.t0 for s in sol]
    t1 = [s.t0 + 0.05 * (s.t1 - s.t0) for s in sol]
    nt = [max(1, s.nt_fine // 20) for s in sol]
    with gpu.Batch([s.problem for s in sol]) as b:
        for i, s in enumerate(sol):
            b[i].fill_exact(s.ic, s.t0)
        u0 = [b[i].get_state() for i in range(len(sol))]
        b.plain(t0, t1, nt)
        one = [b[i].get_state() for i in range(len(sol))]
        for i in range(len(sol)):
            b[i].set_state(2.0 * u0[i])
        b.plain(t0, t1, nt)
        two = [b[i].get_state() for i in range(len(sol))]
    for i in range(len(sol)):
        assert np.array_equal(two[i], 2.0 * one[i]), i
        scale = float(np.sum(np.abs(u0[i])))
        assert abs(float(np.sum(one[i])) - float(np.sum(u0[i]))) <= 1e-12 * scale, i
