"""GPU: the serving engine.  Each request's result is the same alone and in
a mix of requests of other sizes at other iterations (bitwise), and the
same as ccfv_parareal_solve on its own context; admission keeps the
resident state within the memory budget."""
import numpy as np
import pytest

import ccfv
from ccfv import serving, workloads
from ccfv import types as T

pytestmark = pytest.mark.gpu


def small_mix():
    rng = np.random.default_rng(11)
    out = [workloads.ragged_solve(rng, grids=(64, 128), slices=(4, 8), t_end=0.25)
           for _ in range(6)]
    out.append(workloads.Solve(T.rotation_problem(64), T.rotation_ic(), 0.0, 0.25,
                               32, 8, kind="rotation"))
    out.append(workloads.Solve(T.reference_problem(32, 0.01), T.reference_ic(), 0.0,
                               0.3, 24, 6, kind="reference"))
    out.append(workloads.Solve(T.swirl_problem(48, T=1.0), T.cosine_bell_ic(), 0.0,
                               0.5, 32, 4, kind="swirl"))
    return out


def drain(srv):
    out = {}
    for _ in range(10000):
        st = srv.step()
        while True:
            rid = srv.next_done()
            if rid < 0:
                break
            out[rid] = srv.result(rid)
        if st.active == 0 and st.queued == 0:
            return out
    raise AssertionError("engine did not drain")


def alone(s, kmax, tol):
    with ccfv.Server(mem_budget=1 << 30, max_active=4) as srv:
        rid = srv.submit(s.request(kmax, tol))
        return drain(srv)[rid]


@pytest.mark.parametrize("kmax,tol", [(3, 0.0), (8, 1e-9)])
def test_same_alone_and_in_the_mix(gpu, kmax, tol):
    mix = small_mix()
    ref = [alone(s, kmax, tol) for s in mix]
    got = {}
    with ccfv.Server(mem_budget=1 << 30, max_active=64) as srv:
        ids = {}
        # staggered: requests join while earlier ones are mid-iteration
        for i, s in enumerate(mix):
            ids[srv.submit(s.request(kmax, tol))] = i
            if i % 3 == 2:
                srv.step()
                while (rid := srv.next_done()) >= 0:
                    got[ids[rid]] = srv.result(rid)
        for rid, r in drain(srv).items():
            got[ids[rid]] = r
    for i, s in enumerate(mix):
        u, rep, st = got[i]
        ua, ra, sa = ref[i]
        assert st == sa, i
        assert np.array_equal(u, ua), (i, s.kind)
        assert rep.iterations == ra.iterations and rep.converged == ra.converged
        assert list(rep.increments) == list(ra.increments), i
        # and the single-problem entry point
        with gpu.Context(s.problem) as c:
            c.fill_exact(s.ic, s.t0)
            rs = c.parareal_solve(s.t0, s.t1, s.nt_fine, s.n_slices, kmax, tol)
            assert np.array_equal(c.get_state(), u), (i, s.kind)
            assert rs.iterations == rep.iterations
            np.testing.assert_allclose(rs.increments, rep.increments, rtol=1e-12)


def test_admission_by_memory_budget(gpu):
    s = workloads.rotation_solve(128, t_end=0.125, slices=8)
    r = s.request(4, 0.0)
    b = ccfv.request_bytes(r)
    budget = 2 * b + 4096
    ref = alone(s, 4, 0.0)
    with ccfv.Server(mem_budget=budget, max_active=16) as srv:
        ids = [srv.submit(r) for _ in range(5)]
        st = srv.step()
        assert st.active == 2 and st.queued == 3
        assert st.mem_used <= st.mem_budget <= budget
        res = drain(srv)
    assert sorted(res) == sorted(ids)
    for rid in ids:
        assert np.array_equal(res[rid][0], ref[0])


def test_explicit_initial_state_equals_ic(gpu):
    s = small_mix()[0]
    with gpu.Context(s.problem) as c:
        u0 = c.exact(s.ic, s.t0)
    with ccfv.Server(mem_budget=1 << 28, max_active=4) as srv:
        a = srv.submit(s.request(2, 0.0))
        b = srv.submit(s.request(2, 0.0), u0)
        res = drain(srv)
    assert np.array_equal(res[a][0], res[b][0])


def test_request_larger_than_budget_is_refused(gpu):
    s = workloads.rotation_solve(512)
    with ccfv.Server(mem_budget=1 << 24, max_active=4) as srv:
        with pytest.raises(ccfv.CcfvError) as ei:
            srv.submit(s.request(2, 0.0))
        assert ei.value.code == T.CCFV_EBUSY


def test_open_loop_run_on_the_engine(gpu):
    mix = small_mix()
    arr = [(0.002 * i, s.request(3, 0.0), None) for i, s in enumerate(mix)]
    with ccfv.Server(mem_budget=1 << 30, max_active=8) as srv:
        res = serving.run(srv, arr)
    assert len(res.records) == len(mix)
    assert all(r.status == T.CCFV_OK for r in res.records)
    assert all(r.latency > 0 for r in res.records)
    s = res.summary()
    assert s["p50_ms"] <= s["p99_ms"] <= s["p999_ms"]


def test_failing_request_leaves_the_mix(gpu):
    """A request whose coarse step crosses feet (swirl, DT = 1 at 64^2)
    finishes with CCFV_ECFL; the requests beside it are bitwise as alone."""
    mix = small_mix()[:3]
    bad = workloads.Solve(T.swirl_problem(64, T=1.0), T.cosine_bell_ic(), 0.0, 2.0, 4, 2,
                          kind="swirl-crossing")
    ref = [alone(s, 3, 0.0) for s in mix]
    with ccfv.Server(mem_budget=1 << 30, max_active=8) as srv:
        ids = [srv.submit(s.request(3, 0.0)) for s in mix]
        bid = srv.submit(bad.request(3, 0.0))
        res = drain(srv)
    assert res[bid][2] == T.CCFV_ECFL
    for i, rid in enumerate(ids):
        assert res[rid][2] == T.CCFV_OK
        assert np.array_equal(res[rid][0], ref[i][0])


def test_small_plans_and_grids(gpu):
    """kmax = 0, one slice and an 8x8 grid through the engine, as alone."""
    reqs = [
        workloads.Solve(T.rotation_problem(8), T.rotation_ic(), 0.0, 0.25, 4, 4),
        workloads.Solve(T.rotation_problem(64), T.rotation_ic(), 0.0, 0.25, 8, 1),
        workloads.Solve(T.reference_problem(16, 0.01), T.reference_ic(), 0.0, 0.1, 4, 2),
    ]
    plans = [(0, 0.0), (1, 0.0), (2, 1e-12)]
    with ccfv.Server(mem_budget=1 << 28, max_active=4) as srv:
        ids = [srv.submit(s.request(k, tol)) for s, (k, tol) in zip(reqs, plans)]
        res = drain(srv)
    for s, (k, tol), rid in zip(reqs, plans, ids):
        u, rep, st = res[rid]
        with gpu.Context(s.problem) as c:
            c.fill_exact(s.ic, s.t0)
            rs = c.parareal_solve(s.t0, s.t1, s.nt_fine, s.n_slices, k, tol)
            assert np.array_equal(c.get_state(), u)
            assert rs.iterations == rep.iterations and rs.converged == rep.converged
        assert st in (T.CCFV_OK, T.CCFV_ENOTCONV)


def test_plain_requests_are_the_serial_solve(gpu):
    """n_slices = 0: the plain serial solve, advanced 64 steps per engine
    step beside Parareal requests, bitwise ccfv_plain_solve."""
    mix = small_mix()
    plain = [workloads.Solve(T.rotation_problem(128), T.rotation_ic(), 0.0, 0.3, 150, 0),
             workloads.Solve(T.swirl_problem(48, T=1.0), T.cosine_bell_ic(), 0.0, 0.5, 70, 0)]
    with ccfv.Server(mem_budget=1 << 30, max_active=16) as srv:
        ids = [srv.submit(s.request(3, 0.0)) for s in mix[:3]]
        pids = [srv.submit(s.request(0, 0.0)) for s in plain]
        res = drain(srv)
    for s, rid in zip(plain, pids):
        u, rep, st = res[rid]
        assert st == T.CCFV_OK and rep.n_slices == 0 and rep.converged == 1
        assert rep.fine_steps_run == s.nt_fine
        assert np.array_equal(u, solo_plain_state(gpu, s))
    for s, rid in zip(mix[:3], ids):
        assert np.array_equal(res[rid][0], alone(s, 3, 0.0)[0])


def solo_plain_state(gpu, s):
    with gpu.Context(s.problem) as c:
        c.fill_exact(s.ic, s.t0)
        c.plain_solve(s.t0, s.t1, s.nt_fine)
        return c.get_state()


def test_adaptive_policy_runs_busy_arrivals_serially(gpu):
    """adaptive_above = 2: the first two requests run as Parareal; the ones
    submitted while two are waiting run as the plain serial solve, which is
    their serial fine solution exactly."""
    mix = small_mix()[:5]
    with ccfv.Server(mem_budget=1 << 30, max_active=16, adaptive_above=2) as srv:
        ids = [srv.submit(s.request(8, 1e-9)) for s in mix]
        res = drain(srv)
    for i, (s, rid) in enumerate(zip(mix, ids)):
        u, rep, st = res[rid]
        if i < 2:
            assert rep.n_slices == s.n_slices
            assert np.array_equal(u, alone(s, 8, 1e-9)[0])
        else:
            assert rep.n_slices == 0 and st == T.CCFV_OK
            assert np.array_equal(u, solo_plain_state(gpu, s))
