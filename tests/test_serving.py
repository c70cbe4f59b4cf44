"""Serving driver and workload generators on the CPU: the open-loop loop
against a simulated engine and clock, the seeded generators' ranges."""
import math

import numpy as np
import pytest

from ccfv import serving, workloads
from ccfv import types as T


class FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, dt):
        self.t += dt


class FakeEngine:
    """Each step costs `step_s` of simulated time; a request needs `iters`
    steps once admitted; at most `slots` are resident."""

    class St:
        def __init__(self, a, q):
            self.active, self.queued = a, q

    def __init__(self, clock, step_s=1.0, iters=2, slots=2):
        self.clock, self.step_s, self.iters, self.slots = clock, step_s, iters, slots
        self.queue, self.active, self.done, self.next = [], {}, [], 1

    def submit(self, req, u0):
        rid = self.next
        self.next += 1
        self.queue.append(rid)
        return rid

    def step(self):
        while self.queue and len(self.active) < self.slots:
            self.active[self.queue.pop(0)] = 0
        self.clock.t += self.step_s
        for rid in list(self.active):
            self.active[rid] += 1
            if self.active[rid] >= self.iters:
                del self.active[rid]
                self.done.append(rid)
        return self.St(len(self.active), len(self.queue))

    def next_done(self):
        return self.done.pop(0) if self.done else -1

    def result(self, rid):
        return None, None, 0


def test_open_loop_latencies():
    clk = FakeClock()
    eng = FakeEngine(clk, step_s=1.0, iters=2, slots=2)
    arr = [(0.0, "a", None), (0.5, "b", None), (10.0, "c", None)]
    res = serving.run(eng, arr, clock=clk, idle=clk.sleep)
    lat = [r.latency for r in res.records]
    # a: steps [0,1],[1,2]; b arrives mid-step, joins at t=1: [1,2],[2,3];
    # c: loop idles to t=10: [10,11],[11,12]
    assert lat == [2.0, 2.5, 2.0]
    assert [r.submitted for r in res.records] == [0.0, 1.0, 10.0]
    assert res.steps == 5 and res.wall == 12.0
    s = res.summary()
    assert s["requests"] == 3 and math.isclose(s["solves_per_s"], 3 / 12.0)


def test_slots_make_requests_wait():
    clk = FakeClock()
    eng = FakeEngine(clk, step_s=1.0, iters=1, slots=1)
    arr = [(0.0, i, None) for i in range(3)]
    res = serving.run(eng, arr, clock=clk, idle=clk.sleep)
    assert [r.latency for r in res.records] == [1.0, 2.0, 3.0]
    assert res.max_queued == 2


def test_poisson_stream_is_seeded_and_has_the_rate():
    a = workloads.serving_stream(400, rate=200.0, seed=7)
    b = workloads.serving_stream(400, rate=200.0, seed=7)
    ta = np.array([t for t, _ in a])
    assert np.array_equal(ta, np.array([t for t, _ in b]))
    assert np.all(np.diff(ta) > 0)
    gaps = np.diff(np.concatenate([[0.0], ta]))
    assert abs(gaps.mean() - 1 / 200.0) < 0.15 / 200.0
    kinds = {s.kind for _, s in a}
    assert kinds == {"ragged", "rotation"}


def test_ragged_batch_draws():
    sol = workloads.ragged_batch(256)
    again = workloads.ragged_batch(256)
    assert [s.problem.nx for s in sol] == [s.problem.nx for s in again]
    assert {s.problem.nx for s in sol} <= set(workloads.GRIDS)
    assert {s.problem.ny for s in sol} == set(workloads.GRIDS)
    assert {s.n_slices for s in sol} == set(workloads.SLICES)
    for s in sol:
        p = s.problem
        assert p.bc == T.BC_PERIODIC and p.vel_kind == T.VEL_CONSTANT
        assert -1.0 <= p.vel[0] <= 1.0 and -1.0 <= p.vel[1] <= 1.0
        assert 1e-4 <= p.nu <= 1e-2
        assert s.ic.kind == T.IC_GAUSSIANS and s.ic.n == 3
        for b in range(3):
            x0, y0, w, a = s.ic.p[3 + 4 * b: 7 + 4 * b]
            assert 0 <= x0 <= 1 and 0 <= y0 <= 1
            assert 0.03 <= w <= 0.1 and 0.5 <= a <= 1.0
        assert s.nt_fine % s.n_slices == 0
        dt = (s.t1 - s.t0) / s.nt_fine
        assert max(abs(p.vel[0]) * p.nx, abs(p.vel[1]) * p.ny) * dt <= 0.8 + 1e-12
        assert s.nt_fine // s.n_slices >= workloads.J_MIN
    cells = np.mean([s.cells for s in sol])
    assert 0.6 * 157450 < cells < 1.4 * 157450


@pytest.mark.parametrize("n", [256, 512, 1024])
def test_rotation_solve_plan(n):
    s = workloads.rotation_solve(n)
    assert s.nt_fine == 128 and s.n_slices == 16 and s.problem.nx == n
