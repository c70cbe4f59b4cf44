"""Open-loop serving driver over the native engine (ccfv.Server).

Requests are offered at their arrival times; between arrivals the loop runs
engine steps (admission + one Parareal iteration of every resident request,
ccfv_server_step), so requests that arrive while others are on the device
are admitted at the next step boundary as memory and slots free up.
Latency is measured from a request's scheduled arrival to the end of the
step that finished it, so time spent waiting behind a running step counts.

The clock and the idle wait are injectable: the CPU tests drive the loop
with a simulated engine and a simulated clock.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Record:
    index: int
    arrival: float
    submitted: float
    finished: float
    status: int
    report: object = None
    state: np.ndarray | None = None

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


@dataclass
class Result:
    records: list = field(default_factory=list)
    wall: float = 0.0
    steps: int = 0
    max_active: int = 0
    max_queued: int = 0

    def latencies(self) -> np.ndarray:
        return np.array([r.latency for r in self.records], dtype=np.float64)

    def summary(self) -> dict:
        lat = self.latencies()
        n = len(self.records)
        pct = (lambda q: float(np.percentile(lat, q))) if n else (lambda q: float("nan"))
        return {"requests": n, "wall_s": self.wall,
                "solves_per_s": n / self.wall if self.wall > 0 else float("nan"),
                "p50_ms": 1e3 * pct(50), "p99_ms": 1e3 * pct(99),
                "p999_ms": 1e3 * pct(99.9), "steps": self.steps,
                "max_active": self.max_active, "max_queued": self.max_queued}


def run(engine, arrivals, *, clock=time.perf_counter, idle=time.sleep,
        keep_state: bool = False) -> Result:
    """arrivals: list of (t, request, u0 or None) sorted by t (seconds from
    the start).  engine: submit(request, u0) -> id, step() -> status,
    next_done() -> id or -1, result(id) -> (state, report, status)."""
    res = Result()
    t_start = clock()
    now = lambda: clock() - t_start  # noqa: E731
    pending = {}
    i = 0
    while i < len(arrivals) or pending:
        t = now()
        while i < len(arrivals) and arrivals[i][0] <= t:
            ta, req, u0 = arrivals[i]
            rid = engine.submit(req, u0)
            pending[rid] = (i, ta, t)
            i += 1
        if not pending:
            idle(max(0.0, arrivals[i][0] - now()))
            continue
        st = engine.step()
        res.steps += 1
        res.max_active = max(res.max_active, int(st.active))
        res.max_queued = max(res.max_queued, int(st.queued))
        t_fin = now()
        while True:
            rid = engine.next_done()
            if rid < 0:
                break
            idx, ta, ts = pending.pop(rid)
            u, rep, status = engine.result(rid)
            res.records.append(Record(idx, ta, ts, t_fin, status, rep,
                                      u if keep_state else None))
    res.wall = now()
    res.records.sort(key=lambda r: r.index)
    return res
