"""Property-based GPU parity: random problems through the C ABI against the
CPU oracle (hypothesis, derandomized so every run draws the same cases).

A case draws a grid (nx, ny, each a supported size from 8 to 256), a
boundary mode (periodic or the closed box without the reference's quirks),
a velocity (constant, solid-body rotation about a random centre, or the
time-dependent swirl), a diffusion coefficient, a step size, a start time
and a sum of one to three Gaussians.  One Strang step on the GPU must match
the oracle's step to 1e-12 relative L2 (rounding only: both sides are FP64
and evaluate the feet with the same contraction-free expressions).  A step
the library refuses because characteristic feet cross is a refusal of the
oracle too (both report CCFV_ECFL); such cases are discarded.  On periodic
problems the GPU step also conserves the field's sum to 1e-13 of the sum of
its magnitudes (the remap is conservative, the compact operator's and the
implicit matrix's rows sum to one, and the diffusive fluxes telescope).
"""
import math

import numpy as np
import pytest

hypothesis = pytest.importorskip("hypothesis")
from hypothesis import HealthCheck, assume, given, settings  # noqa: E402
from hypothesis import strategies as st  # noqa: E402

import ccfv  # noqa: E402
from ccfv import types as T  # noqa: E402

pytestmark = pytest.mark.gpu

# sizes the library accepts (even, a multiple of ceil-pow2(n/64))
SIZES = [8, 16, 24, 32, 40, 48, 64, 72, 96, 128, 136, 160, 256]


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@st.composite
def cases(draw):
    nx = draw(st.sampled_from(SIZES))
    ny = draw(st.sampled_from(SIZES))
    bc = draw(st.sampled_from([T.BC_PERIODIC, T.BC_REFERENCE]))
    kind = draw(st.sampled_from([T.VEL_CONSTANT, T.VEL_SOLID_BODY, T.VEL_SWIRL]))
    if kind == T.VEL_CONSTANT:
        vel = (draw(st.floats(-2.0, 2.0)), draw(st.floats(-2.0, 2.0)))
    elif kind == T.VEL_SOLID_BODY:
        vel = (draw(st.floats(-7.0, 7.0)), draw(st.floats(0.2, 0.8)), draw(st.floats(0.2, 0.8)))
    else:
        vel = (draw(st.floats(0.5, 5.0)),)
    nu = draw(st.sampled_from([0.0, 1e-4, 1e-3, 1e-2]))
    h = draw(st.floats(1e-4, 0.05))
    t0 = draw(st.floats(0.0, 1.0))
    nb = draw(st.integers(1, 3))
    bumps = [(draw(st.floats(0.2, 0.8)), draw(st.floats(0.2, 0.8)),
              draw(st.floats(0.04, 0.15)), draw(st.floats(0.2, 1.5))) for _ in range(nb)]
    p = T.make_problem(nx, ny, bc=bc, vel_kind=kind, vel=vel, nu=nu)
    params = [0.0, 0.0, 0.0]
    for b in bumps:
        params.extend(b)
    ic = T.make_ic(T.IC_GAUSSIANS, params, n=nb)
    return p, ic, h, t0


@settings(max_examples=200, deadline=None, derandomize=True, database=None,
          suppress_health_check=[HealthCheck.too_slow])
@given(case=cases())
def test_random_step_matches_oracle(gpu, oracle, case):
    p, ic, h, t0 = case
    print(f"case nx={p.nx} ny={p.ny} bc={p.bc} vel={p.vel_kind}:{list(p.vel)[:3]} "
          f"nu={p.nu} h={h} t0={t0}", flush=True)
    u0 = oracle.exact(p, ic, 0.0)
    try:
        want = oracle.step(p, u0, t0, h)
    except RuntimeError as e:
        assert "ECFL" in str(e) or str(T.CCFV_ECFL) in str(e), e
        want = None
    with gpu.Context(p) as c:
        c.set_state(u0)
        try:
            c.step(t0, h)
            got = c.get_state()
        except ccfv.CcfvError as e:
            assert e.code == T.CCFV_ECFL, e
            got = None
    # both refuse, or both step
    assert (got is None) == (want is None)
    assume(got is not None)
    assert rel_l2(got, want) <= 1e-12
    if p.bc == T.BC_PERIODIC:
        s0 = float(np.sum(u0))
        scale = max(float(np.sum(np.abs(u0))), 1e-300)
        assert math.isfinite(float(np.sum(got)))
        assert abs(float(np.sum(got)) - s0) <= 1e-13 * scale
