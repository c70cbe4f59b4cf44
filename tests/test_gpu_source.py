"""GPU: problems with a source term (T2S4_Parareal.cu:58-61 f, :142-163 its
Simpson cell average, added as theta h (fbar0 + f1) in every sub-step,
:574 x and :606 y) against the CPU oracle, and the batched and
serving paths against the one-problem path.  Tolerances as for the
source-free problems: relative L2 <= 1e-13 for one step, <= 1e-12 after
hundreds of steps; batch and serving results are bitwise equal to the
one-problem results."""
import math

import numpy as np
import pytest

import ccfv
from ccfv import types as T
from ccfv import workloads

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def sourced(p, kind, *params):
    return p.copy(src_kind=kind, src=params)


def problems():
    man, man_ic, _ = T.manufactured_sinsin(64)
    g, gic = T.constant_gaussians(64, 128, -0.7, 0.4, 3e-3,
                                  [(0.3, 0.6, 0.05, 1.0), (0.7, 0.2, 0.08, 0.6)])
    return [
        # periodic constant velocity: the path without a source would be the
        # uniform-displacement kernel; with one it is the general kernel
        ("manufactured", man, man_ic),
        ("rotation-const", sourced(T.rotation_problem(64), T.SRC_CONSTANT, 0.3),
         T.rotation_ic()),
        # closed box, post-remap last x sub-step and the edge-derivative quirk
        ("reference-sinsin", sourced(T.reference_problem(32, 0.01), T.SRC_SINSIN,
                                     5.0, 0.4, -0.2, 0.3, 0.1), T.reference_ic()),
        ("swirl-sinsin", sourced(T.swirl_problem(48, T=1.0), T.SRC_SINSIN,
                                 2.0, 0.0, 0.5, -0.25, 0.25), T.cosine_bell_ic()),
        ("gauss-nonsquare-const", sourced(g, T.SRC_CONSTANT, -0.6), gic),
    ]


IDS = lambda v: v if isinstance(v, str) else None  # noqa: E731


@pytest.mark.parametrize("name,p,ic", problems(), ids=IDS)
@pytest.mark.parametrize("h", [1.0 / 256, 1.0 / 16])
def test_one_step_with_source_matches_oracle(gpu, oracle, name, p, ic, h):
    u0 = oracle.exact(p, ic, 0.0)
    with gpu.Context(p) as c:
        c.set_state(u0)
        c.step(0.25, h)
        g = c.get_state()
    o = oracle.step(p, u0, 0.25, h)
    assert rel_l2(g, o) <= 1e-13, name
    # the source changed the step
    q = p.copy(src_kind=T.SRC_NONE)
    assert rel_l2(oracle.step(q, u0, 0.25, h), o) > 1e-6


@pytest.mark.parametrize("name,p,ic", problems(), ids=IDS)
def test_plain_solve_with_source_matches_oracle(gpu, oracle, name, p, ic):
    u0 = oracle.exact(p, ic, 0.0)
    nt = 128
    with gpu.Context(p) as c:
        c.set_state(u0)
        c.plain_solve(0.0, 0.5, nt)
        g = c.get_state()
        # the fused x launch (last x of step i with first x of step i+1)
        # against separate launches: bitwise equal
        c.set_option(T.OPT_FUSE, 0)
        c.set_state(u0)
        c.plain_solve(0.0, 0.5, nt)
        g2 = c.get_state()
    o = oracle.plain(p, u0, 0.0, 0.5, nt)
    assert rel_l2(g, o) <= 1e-12, name
    assert np.array_equal(g, g2), name


@pytest.mark.parametrize("name,p,ic", problems()[1:4], ids=IDS)
def test_parareal_with_source_matches_reference_schedule(gpu, oracle, name, p, ic):
    u0 = oracle.exact(p, ic, 0.0)
    ns, nt, kmax = 8, 64, 3
    r = oracle.parareal(p, u0, 0.0, 0.5, nt, ns, kmax)
    res = []
    for mode in (0, 1, 2):
        with gpu.Context(p) as c:
            c.set_state(u0)
            c.set_option(T.OPT_PIPELINE, mode)
            rep = c.parareal_solve(0.0, 0.5, nt, ns, kmax)
            res.append((c.get_state(), rep))
    g, rep = res[0]
    assert rep.iterations == kmax
    assert rel_l2(g, r["u"]) <= 1e-12, name
    for a, b in zip(rep.increments, r["increments"]):
        assert abs(a - b) <= 1e-9 * max(abs(b), 1e-30) + 1e-15
    for gm, rm in res[1:]:
        assert np.array_equal(g, gm), name
        assert list(rm.increments) == list(rep.increments)


def test_parareal_with_source_kmax_ns_is_serial_fine(gpu, oracle):
    _, p, ic = problems()[0]
    u0 = oracle.exact(p, ic, 0.0)
    ns, nt = 8, 64
    serial = oracle.plain(p, u0, 0.0, 0.5, nt)
    with gpu.Context(p) as c:
        c.set_state(u0)
        rep = c.parareal_solve(0.0, 0.5, nt, ns, ns)
        g = c.get_state()
    assert rep.converged == 1
    assert rel_l2(g, serial) <= 1e-12


def test_manufactured_solution_on_gpu(gpu):
    """f = g / 2 (the scheme adds 2 h f per step): first order in h, the
    order of the sub-step source times."""
    n, t1 = 64, 0.25
    d = 1.0 / n
    x = (np.arange(n) + 0.5) * d
    a = (np.cos(2 * np.pi * (x - 0.5 * d)) - np.cos(2 * np.pi * (x + 0.5 * d))) / (2 * np.pi * d)
    err = []
    p, ic, amp = T.manufactured_sinsin(n)
    with gpu.Context(p) as c:
        exact = c.exact(ic, t1) + amp(t1) * np.outer(a, a).reshape(-1)
        for nt in (64, 128, 256):
            c.fill_exact(ic, 0.0)
            c.plain_solve(0.0, t1, nt)
            err.append(math.sqrt(float(np.mean((c.get_state() - exact) ** 2))))
    for e0, e1 in zip(err, err[1:]):
        assert 1.9 <= e0 / e1 <= 2.1, err


def _mix():
    out = [(p, ic) for _, p, ic in problems()]
    # source-free problems in the same launches
    out.append((T.rotation_problem(64), T.rotation_ic()))
    out.append((T.reference_problem(32, 0.01), T.reference_ic()))
    return out


def test_batch_with_sources_matches_solo(gpu):
    m = _mix()
    t0, t1, nt = 0.1, 0.35, 20
    with gpu.Batch([p for p, _ in m]) as b:
        for i, (_, ic) in enumerate(m):
            b[i].fill_exact(ic, t0)
        b.plain(t0, t1, nt)
        got = [b[i].get_state() for i in range(len(m))]
    for i, (p, ic) in enumerate(m):
        with gpu.Context(p) as c:
            c.fill_exact(ic, t0)
            c.plain_solve(t0, t1, nt)
            assert np.array_equal(got[i], c.get_state()), i
    with gpu.Batch([p for p, _ in m]) as b:
        for i, (_, ic) in enumerate(m):
            b[i].fill_exact(ic, 0.0)
        reps = b.parareal(0.0, 0.5, 64, 8, 3)
        got = [b[i].get_state() for i in range(len(m))]
    for i, (p, ic) in enumerate(m):
        with gpu.Context(p) as c:
            c.fill_exact(ic, 0.0)
            rep = c.parareal_solve(0.0, 0.5, 64, 8, 3)
            assert np.array_equal(got[i], c.get_state()), i
            assert list(rep.increments) == list(reps[i].increments), i


def test_serving_with_sources_matches_context(gpu):
    m = _mix()
    solves = [workloads.Solve(p, ic, 0.0, 0.5, 64, 8, kind="source") for p, ic in m]
    # a plain request (n_slices 0) with a source as well
    solves.append(workloads.Solve(m[2][0], m[2][1], 0.0, 0.5, 64, 0, kind="source"))
    got = {}
    with ccfv.Server(mem_budget=1 << 30, max_active=16) as srv:
        ids = {srv.submit(s.request(3, 0.0)): i for i, s in enumerate(solves)}
        for _ in range(10000):
            st = srv.step()
            while (rid := srv.next_done()) >= 0:
                got[ids[rid]] = srv.result(rid)
            if st.active == 0 and st.queued == 0:
                break
    assert len(got) == len(solves)
    for i, s in enumerate(solves):
        u, rep, status = got[i]
        assert status == T.CCFV_OK, i
        with gpu.Context(s.problem) as c:
            c.fill_exact(s.ic, 0.0)
            if s.n_slices:
                c.parareal_solve(0.0, 0.5, 64, 8, 3)
            else:
                c.plain_solve(0.0, 0.5, 64)
            assert np.array_equal(u, c.get_state()), i
