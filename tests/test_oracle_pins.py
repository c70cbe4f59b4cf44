"""Pins of the oracle's velocity and initial-data branches that the
reference's own problem (solid-body rotation, one Gaussian) does not reach:
the deformational swirl velocity, the sum-of-Gaussians data that diffuses
exactly, and the cosine bell.

Each check is chosen so that a plausible misreading of the workload
definitions fails it:
  swirl  u =  sin^2(pi x) sin(2 pi y) cos(pi t / T),
         v = -sin^2(pi y) sin(2 pi x) cos(pi t / T):
    - the small-step limit of the traced feet is the velocity itself, at
      points and times where cos(pi t/T), cos(2 pi t/T) and the sign differ;
    - the flow reverses: ∫_0^T cos(pi t/T) dt = 0, so with nu = 0 the solution
      returns to the initial data at t = T (error falling with the grid), and
      it is furthest from it at t = T/2 (a cos(2 pi t/T) factor would return
      there instead);
  Gaussians  a exp(-r^2 / (2 w^2)) translated by (u, v) and diffused:
    - cell averages at t = 0 against two-dimensional quadrature (w is the
      standard deviation);
    - the heat kernel: mass a 2 pi w^2 is conserved, the centroid moves with
      (u, v) t and the variance per direction is w^2 + 2 nu t (plus dx^2/12
      for cell averages) — a 4 nu t or a w^2 / (w^2 + 4 nu t) amplitude fails;
    - the sum of three Gaussians is the sum of the single ones;
  cosine bell  0.5 (1 + cos(pi r / R)) for r < R:
    - cell averages against quadrature inside, zero outside;
    - the mass R^2 (pi/2 - 2/pi) (closed form of the disk integral).
"""
import math

import numpy as np
import pytest
from scipy import integrate

from ccfv import types as T


# ---------------------------------------------------------------------------
# swirl velocity
# ---------------------------------------------------------------------------
def swirl_u(x, y, t, TT):
    return np.sin(np.pi * x) ** 2 * np.sin(2 * np.pi * y) * np.cos(np.pi * t / TT)


def swirl_v(x, y, t, TT):
    return -np.sin(np.pi * y) ** 2 * np.sin(2 * np.pi * x) * np.cos(np.pi * t / TT)


@pytest.mark.parametrize("t", [0.0, 0.7, 1.9, 3.3])
def test_swirl_feet_small_step_limit_is_the_velocity(oracle, t):
    """(s - foot)/tau -> mean over the face's two corners of the velocity
    component along the line, to O(tau)."""
    TT, n, tau = 5.0, 32, 1e-7
    p = T.swirl_problem(n, T=TT)
    d = 1.0 / n
    s = np.arange(n + 1) * d
    for line in (3, 11, 20):
        fx = oracle.feet_x(p, line, tau, t, t)
        want = 0.5 * (swirl_u(s, line * d, t, TT) + swirl_u(s, (line + 1) * d, t, TT))
        np.testing.assert_allclose((s - fx) / tau, want, atol=1e-6)
        fy = oracle.feet_y(p, line, tau, t, t)
        want = 0.5 * (swirl_v(line * d, s, t, TT) + swirl_v((line + 1) * d, s, t, TT))
        np.testing.assert_allclose((s - fy) / tau, want, atol=1e-6)
    # the check is not vacuous: a doubled time frequency differs by O(1) here
    if t > 0:
        alt = np.cos(2 * np.pi * t / TT) - np.cos(np.pi * t / TT)
        assert abs(alt) > 0.05


def test_swirl_reverses_at_the_period(oracle):
    """nu = 0: u(T) returns to u0 (error falls with the grid); at T/2 the
    bell is strongly deformed."""
    TT = 1.0
    ret, half = [], []
    for n in (64, 128):
        p = T.swirl_problem(n, T=TT, nu=0.0)
        ic = T.cosine_bell_ic()
        u0 = oracle.exact(p, ic, 0.0)
        nt = n  # CFL <= 1 with |u| <= 1
        uh = oracle.plain(p, u0, 0.0, 0.5 * TT, nt // 2)
        u1 = oracle.plain(p, uh, 0.5 * TT, TT, nt // 2)
        half.append(np.linalg.norm(uh - u0) / np.linalg.norm(u0))
        ret.append(np.linalg.norm(u1 - u0) / np.linalg.norm(u0))
    assert ret[1] < ret[0] / 2.5, ret
    assert ret[1] < 0.02, ret
    assert min(half) > 0.3, half


# ---------------------------------------------------------------------------
# sum of Gaussians, translated and diffused (CCFV_IC_GAUSSIANS)
# ---------------------------------------------------------------------------
def _gauss_problem(n, u, v, nu, bumps):
    return T.constant_gaussians(n, n, u, v, nu, bumps)


def test_gaussian_cell_averages_match_quadrature(oracle):
    n = 32
    x0, y0, w, a = 0.45, 0.55, 0.06, 0.8
    p, ic = _gauss_problem(n, 0.3, -0.2, 1e-3, [(x0, y0, w, a)])
    u = oracle.exact(p, ic, 0.0).reshape(n, n)
    f = lambda y, x: a * np.exp(-((x - x0) ** 2 + (y - y0) ** 2) / (2 * w * w))  # noqa: E731
    for i, j in ((14, 17), (15, 16), (12, 20), (18, 14)):
        v, _ = integrate.dblquad(f, i / n, (i + 1) / n, j / n, (j + 1) / n,
                                 epsabs=1e-13, epsrel=1e-11)
        # Simpson 3x3 is exact for bicubics; the Gaussian differs at O(d^4)
        assert u[i, j] == pytest.approx(v * n * n, rel=3e-3)


def _moments(u, n):
    """mass, centroid and variances of cell averages on the unit torus
    (distances measured from the grid centre, the data kept away from the
    edges)."""
    d = 1.0 / n
    xc = (np.arange(n) + 0.5) * d
    U = u.reshape(n, n)
    m = U.sum() * d * d
    mx = (U.sum(axis=1) * xc).sum() * d * d / m
    my = (U.sum(axis=0) * xc).sum() * d * d / m
    vx = (U.sum(axis=1) * (xc - mx) ** 2).sum() * d * d / m
    vy = (U.sum(axis=0) * (xc - my) ** 2).sum() * d * d / m
    return m, mx, my, vx, vy


@pytest.mark.parametrize("t", [0.0, 0.25])
def test_gaussian_is_the_heat_kernel(oracle, t):
    n, w, a = 256, 0.05, 0.9
    u_, v_, nu = 0.4, -0.3, 1e-2
    x0, y0 = 0.4, 0.6
    p, ic = _gauss_problem(n, u_, v_, nu, [(x0, y0, w, a)])
    m, mx, my, vx, vy = _moments(oracle.exact(p, ic, t), n)
    assert m == pytest.approx(a * 2 * math.pi * w * w, rel=1e-6)
    assert mx == pytest.approx(x0 + u_ * t, abs=1e-6)
    assert my == pytest.approx(y0 + v_ * t, abs=1e-6)
    var = w * w + 2 * nu * t + (1.0 / n) ** 2 / 12
    assert vx == pytest.approx(var, rel=1e-5)
    assert vy == pytest.approx(var, rel=1e-5)
    # the peak decays as w^2 / (w^2 + 2 nu t): the centre cell of a run with
    # no translation, against the t = 0 centre cell
    q, icq = _gauss_problem(n, 0.0, 0.0, nu, [(0.5, 0.5, w, a)])
    c0 = oracle.exact(q, icq, 0.0).reshape(n, n)[n // 2 - 1: n // 2 + 1, n // 2 - 1: n // 2 + 1]
    ct = oracle.exact(q, icq, t).reshape(n, n)[n // 2 - 1: n // 2 + 1, n // 2 - 1: n // 2 + 1]
    assert ct.mean() / c0.mean() == pytest.approx(w * w / (w * w + 2 * nu * t), rel=2e-3)


def test_sum_of_gaussians_is_the_sum(oracle):
    n = 64
    bumps = [(0.3, 0.6, 0.05, 1.0), (0.7, 0.2, 0.08, 0.6), (0.5, 0.5, 0.03, 0.5)]
    p, ic = _gauss_problem(n, -0.7, 0.4, 3e-3, bumps)
    for t in (0.0, 0.4):
        whole = oracle.exact(p, ic, t)
        parts = sum(oracle.exact(*_gauss_problem(n, -0.7, 0.4, 3e-3, [b]), t) for b in bumps)
        np.testing.assert_allclose(whole, parts, rtol=0, atol=1e-14)


# ---------------------------------------------------------------------------
# cosine bell (CCFV_IC_COSINE_BELL)
# ---------------------------------------------------------------------------
def test_cosine_bell_cell_averages_match_quadrature(oracle):
    n = 64
    x0, y0, R = 0.5, 0.75, 0.15
    p, ic = T.swirl_problem(n), T.cosine_bell_ic(x0, y0, R)
    u = oracle.exact(p, ic, 0.0).reshape(n, n)

    def f(y, x):
        r = math.hypot(x - x0, y - y0)
        return 0.5 * (1 + math.cos(math.pi * r / R)) if r < R else 0.0
    # cells well inside the bell (smooth there) and one outside
    for i, j in ((32, 48), (30, 50), (34, 45), (28, 52)):
        v, _ = integrate.dblquad(f, i / n, (i + 1) / n, j / n, (j + 1) / n,
                                 epsabs=1e-13, epsrel=1e-11)
        assert u[i, j] == pytest.approx(v * n * n, rel=1e-4)
    assert u[5, 5] == 0.0 and u[32, 20] == 0.0


def test_cosine_bell_mass(oracle):
    R = 0.15
    exact = R * R * (math.pi / 2 - 2 / math.pi)
    errs = []
    for n in (64, 128, 256):
        p, ic = T.swirl_problem(n), T.cosine_bell_ic(0.5, 0.75, R)
        m = oracle.exact(p, ic, 0.0).sum() / (n * n)
        errs.append(abs(m - exact) / exact)
    assert errs[2] < 1e-4, errs
    assert errs[2] < errs[0], errs


# ---------------------------------------------------------------------------
# rotating Gaussian (CCFV_IC_ROT_GAUSSIAN, C_Exact of T2S4_Parareal.cu:51-56)
# at t > 0: rotated about the centre, diffused
# ---------------------------------------------------------------------------
def test_rotating_gaussian_moments(oracle):
    n, nu, t = 256, 2e-3, 0.1
    p, ic = T.rotation_problem(n, nu=nu), T.rotation_ic(0.5, 0.75, 0.05)
    m, mx, my, vx, vy = _moments(oracle.exact(p, ic, t), n)
    om = 2 * math.pi
    # (0.5, 0.75) rotated counter-clockwise by om*t about (0.5, 0.5)
    cx = 0.5 - 0.25 * math.sin(om * t)
    cy = 0.5 + 0.25 * math.cos(om * t)
    assert m == pytest.approx(2 * math.pi * 0.05 ** 2, rel=1e-6)
    assert mx == pytest.approx(cx, abs=1e-6)
    assert my == pytest.approx(cy, abs=1e-6)
    var = 0.05 ** 2 + 2 * nu * t + (1.0 / n) ** 2 / 12
    assert vx == pytest.approx(var, rel=1e-5)
    assert vy == pytest.approx(var, rel=1e-5)
