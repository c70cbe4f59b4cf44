"""GPU parity: the HIP path (through the C ABI) against the CPU oracle on the
same inputs.  Tolerances: relative L2 <= 1e-13 for one step (rounding only),
<= 1e-12 after hundreds of steps (the bound stated for the serial fine
solution); both sides are FP64."""
import math

import numpy as np
import pytest

from ccfv import types as T

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def problems():
    out = []
    for n in (32, 64):
        out.append(("reference", T.reference_problem(n, 0.01), T.reference_ic()))
    for n in (32, 64, 128):
        out.append(("rotation", T.rotation_problem(n), T.rotation_ic()))
    out.append(("sinsin", T.advection_sinsin_problem(64), T.advection_sinsin_ic()))
    out.append(("swirl", T.swirl_problem(64), T.cosine_bell_ic()))
    p, ic = T.constant_gaussians(64, 128, -0.7, 0.4, 3e-3,
                                 [(0.3, 0.6, 0.05, 1.0), (0.7, 0.2, 0.08, 0.6)])
    out.append(("gauss-nonsquare", p, ic))
    return out


@pytest.mark.parametrize("name,p,ic", problems(), ids=lambda v: v if isinstance(v, str) else None)
def test_fill_exact_matches_oracle(gpu, oracle, name, p, ic):
    with gpu.Context(p) as c:
        for t in (0.0, 0.3):
            g = c.exact(ic, t)
            o = oracle.exact(p, ic, t)
            assert rel_l2(g, o) <= 1e-14, (name, t)


@pytest.mark.parametrize("name,p,ic", problems(), ids=lambda v: v if isinstance(v, str) else None)
@pytest.mark.parametrize("h", [1.0 / 256, 1.0 / 16])
def test_one_step_matches_oracle(gpu, oracle, name, p, ic, h):
    u0 = oracle.exact(p, ic, 0.0)
    with gpu.Context(p) as c:
        c.set_state(u0)
        c.step(0.25, h)
        g = c.get_state()
    o = oracle.step(p, u0, 0.25, h)
    assert rel_l2(g, o) <= 1e-13, name


@pytest.mark.parametrize("name,p,ic", problems()[::2], ids=lambda v: v if isinstance(v, str) else None)
def test_plain_solve_matches_oracle(gpu, oracle, name, p, ic):
    u0 = oracle.exact(p, ic, 0.0)
    nt = 128
    with gpu.Context(p) as c:
        c.set_state(u0)
        c.plain_solve(0.0, 0.5, nt)
        g = c.get_state()
    o = oracle.plain(p, u0, 0.0, 0.5, nt)
    assert rel_l2(g, o) <= 1e-12, name


def test_fine_equals_steps(gpu, oracle):
    p, ic = T.rotation_problem(64), T.rotation_ic()
    u0 = oracle.exact(p, ic, 0.0)
    with gpu.Context(p) as c:
        c.set_state(u0)
        c.fine(0.125, 1.0 / 16, 8)
        g = c.get_state()
    o = u0.copy()
    for j in range(8):
        o = oracle.step(p, o, 0.125 + j * (1.0 / 128), 1.0 / 128)
    assert rel_l2(g, o) <= 1e-13


@pytest.mark.parametrize("name,p,ic", [
    ("rotation", T.rotation_problem(64), T.rotation_ic()),
    ("reference", T.reference_problem(32, 0.01), T.reference_ic()),
    ("swirl", T.swirl_problem(32, T=1.0), T.cosine_bell_ic()),
], ids=lambda v: v if isinstance(v, str) else None)
def test_parareal_matches_reference_schedule(gpu, oracle, name, p, ic):
    """The batched schedule (cached G, converged-slice skip, fused correction)
    against the oracle's literal restatement of the reference's serial loop."""
    u0 = oracle.exact(p, ic, 0.0)
    ns, nt, kmax = 8, 64, 3
    r = oracle.parareal(p, u0, 0.0, 0.5, nt, ns, kmax)
    with gpu.Context(p) as c:
        c.set_state(u0)
        rep = c.parareal_solve(0.0, 0.5, nt, ns, kmax)
        g = c.get_state()
    assert rep.iterations == kmax
    assert rel_l2(g, r["u"]) <= 1e-12, name
    for a, b in zip(rep.increments, r["increments"]):
        assert abs(a - b) <= 1e-9 * max(abs(b), 1e-30) + 1e-15


def test_parareal_kmax_ns_is_serial_fine(gpu, oracle):
    p, ic = T.rotation_problem(64), T.rotation_ic()
    u0 = oracle.exact(p, ic, 0.0)
    ns, nt = 8, 64
    serial = oracle.plain(p, u0, 0.0, 0.5, nt)
    with gpu.Context(p) as c:
        c.set_state(u0)
        c.parareal_solve(0.0, 0.5, nt, ns, ns)
        g = c.get_state()
    assert rel_l2(g, serial) <= 1e-12


def test_full_size_rotation_plain_matches_oracle(gpu, oracle):
    """512^2 solid-body rotation, 1024 steps of 1/1024: the serial fine
    solution of the Parareal workload, GPU against oracle."""
    p, ic = T.rotation_problem(512), T.rotation_ic()
    u0 = oracle.exact(p, ic, 0.0)
    with gpu.Context(p) as c:
        c.set_state(u0)
        c.plain_solve(0.0, 1.0, 1024)
        g = c.get_state()
    o = oracle.plain(p, u0, 0.0, 1.0, 1024)
    assert rel_l2(g, o) <= 1e-12


def test_plain_method_convergence_on_gpu(gpu, oracle):
    """The plain-method workload (periodic sin*sin, velocity (1, 0.5),
    nu = 0.01, CFL 0.8, T = 1) on 32^2..256^2 against the exact decaying
    translated solution: third order at CFL 0.8 (the fourth-order remap
    error accumulates over O(1/dx) steps), fourth order in space at a fixed
    step count; the GPU errors equal the oracle's."""
    def err(n, nt, dev):
        p, ic = T.advection_sinsin_problem(n), T.advection_sinsin_ic()
        ex = oracle.exact(p, ic, 1.0)
        if dev:
            with gpu.Context(p) as c:
                c.fill_exact(ic, 0.0)
                c.plain_solve(0.0, 1.0, nt)
                u = c.get_state()
        else:
            u = oracle.plain(p, oracle.exact(p, ic, 0.0), 0.0, 1.0, nt)
        return math.sqrt(float(np.mean((u - ex) ** 2)))
    ns = (32, 64, 128, 256)
    e = [err(n, T.advection_sinsin_steps(n), True) for n in ns]
    orders = [math.log2(e[i] / e[i + 1]) for i in range(3)]
    assert all(2.8 < o < 3.4 for o in orders), orders
    for n, eg in zip(ns[:2], e[:2]):
        eo = err(n, T.advection_sinsin_steps(n), False)
        assert abs(eg - eo) <= 1e-9 * eo, (n, eg, eo)
    es = [err(n, 2560, True) for n in (32, 64)]
    assert 3.8 < math.log2(es[0] / es[1]) < 4.2


def test_crossed_feet_are_an_error_and_recoverable(gpu, oracle):
    """A step so long that characteristic feet cross (swirl, h = 1 at 64^2:
    the foot map is no longer monotone) reports CCFV_ECFL; the context stays
    usable and the next valid step matches the oracle."""
    p, ic = T.swirl_problem(64, T=1.0), T.cosine_bell_ic()
    u0 = oracle.exact(p, ic, 0.0)
    with gpu.Context(p) as c:
        c.set_state(u0)
        with pytest.raises(gpu.CcfvError) as ei:
            c.step(0.0, 1.0)
        assert ei.value.code == T.CCFV_ECFL
        c.set_state(u0)
        c.step(0.0, 0.01)
        g = c.get_state()
    assert rel_l2(g, oracle.step(p, u0, 0.0, 0.01)) <= 1e-13


@pytest.mark.parametrize("n", [256, 512, 1024])
def test_reference_preset_large_matches_oracle(gpu, oracle, n):
    """The reference program's own problem (closed box [-1,1]^2, V = (-4y,
    4x), K = 0.01, Gaussian at (-0.35, 0), both quirks on) at the grid sizes
    its config file is run with, where the general sweep's lanes hold 4 and 8
    cells and stage the right-hand side in LDS: 16 serial fine steps, and a
    Parareal solve of 4 slices x 4 steps with 2 iterations, against the
    oracle's restatement of the reference loops."""
    p, ic = T.reference_problem(n, 0.01), T.reference_ic()
    u0 = oracle.exact(p, ic, 0.0)
    t1, nt = 16.0 / 512, 16
    print(f"reference preset {n}^2: plain {nt} steps, Parareal 4 x 4", flush=True)
    with gpu.Context(p) as c:
        c.set_state(u0)
        c.plain_solve(0.0, t1, nt)
        g = c.get_state()
        c.set_state(u0)
        rep = c.parareal_solve(0.0, t1, nt, 4, 2)
        gp = c.get_state()
    assert rel_l2(g, oracle.plain(p, u0, 0.0, t1, nt)) <= 1e-12
    r = oracle.parareal(p, u0, 0.0, t1, nt, 4, 2)
    assert rep.iterations == 2
    assert rel_l2(gp, r["u"]) <= 1e-12
    for a, b in zip(rep.increments, r["increments"]):
        assert abs(a - b) <= 1e-9 * abs(b)


def test_reference_preset_2048_matches_oracle(gpu, oracle):
    """The reference program's problem at 2048^2: the general sweep's
    32-cell lanes with the right-hand side and the third-sub-step line in
    LDS (two auxiliary lines per wave), 8 serial fine steps."""
    p, ic = T.reference_problem(2048, 0.01), T.reference_ic()
    u0 = oracle.exact(p, ic, 0.0)
    t1, nt = 8.0 / 2048, 8
    print("reference preset 2048^2: plain 8 steps", flush=True)
    with gpu.Context(p) as c:
        c.set_state(u0)
        c.plain_solve(0.0, t1, nt)
        g = c.get_state()
    assert rel_l2(g, oracle.plain(p, u0, 0.0, t1, nt)) <= 1e-12
