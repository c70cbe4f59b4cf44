#!/usr/bin/env python3
"""Benchmark: FP64 cell-updates/s of batched Parareal against the plain
serial-in-time sweep on one MI355X, with p50/p99 solve latency.

Workload (one "step" = one whole solve): periodic [0,1]^2, 512x512 FP64 cell
averages, solid-body rotation u=-2 pi (y-1/2), v=2 pi (x-1/2), nu=1e-3,
Gaussian hump at (0.5, 0.75) with standard deviation 0.05, T=1 in 1024 fine
steps; Parareal with 64 slices, coarse dt = 16 fine dt, at most 8 iterations,
stop at a relative L2 increment of 1e-10.  All 64 fine propagators of an
iteration run as one batched launch per sweep.

value = serial-equivalent cell-updates/s = N^2 * 1024 / (p50 Parareal solve).
The plain sweep (the serial fine solution) is timed the same way and reported
beside it.  Inputs are resident in HBM before the timed region.

Roofline: every kernel class is timed with hipEvents on the context's stream
during one extra instrumented solve on one stream (the sequential schedule,
so durations are not inflated by the concurrent coarse chain).  The
roofline line is the class that moves the most HBM bytes, the fine
propagators' fused x sweep: its algorithmic bytes (8 B in + 8 B out per
cell per field) over its mean launch time give `achieved`; every class is
listed under kernel_classes (the single-field coarse-chain kernels are
latency-bound, far below the HBM roof).
cpu_baseline: the same scheme (the CPU restatement) on the host cores,
OpenMP over lines, on a bounded number of steps of the plain sweep.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "parareal-cc-cfv-split-gpu_amd"))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E

# HBM bytes per launch over algorithmic bytes, per kernel class, from two
# rocprofv3 --pmc passes (FETCH_SIZE x2.0 + WRITE_SIZE x1.0: the scale factors
# that make copy_kernel of a known size read its true bytes), over one
# Parareal iteration of this workload on one stream
# (profiles/rotation512_fine_sweep_pmc.md)
TRAFFIC_RATIO = {"x_batch": 1.002, "y_batch": 1.002, "y_single": 1.064,
                 "x_epilog": 1.029}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30, help="timed solves")
    ap.add_argument("--warmup", type=int, default=5, help="untimed solves")
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--nt", type=int, default=1024)
    ap.add_argument("--slices", type=int, default=64)
    ap.add_argument("--kmax", type=int, default=8)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="bound on the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain-graphs", action="store_true",
                    help="skip timing the plain sweep replayed from a hipGraph "
                         "(for rocprofv3 --kernel-trace runs: the profiler's "
                         "process ends in a segmentation fault inside the "
                         "capture of the 2048-launch plain solve)")
    ap.add_argument("--ring", type=int, default=8,
                    help="--workload copy: fields of 32 MiB in the ring (8: 256 MiB, "
                         "more than the last-level cache holds; 4: 128 MiB, the "
                         "size of one batch of 64 fine propagators at 512^2)")
    ap.add_argument("--probe", action="store_true",
                    help="short run for counter collection: one 32 MiB "
                         "calibration copy and one Parareal iteration "
                         "(--workload swirl: 64 plain steps and one Parareal "
                         "iteration, one stream, no graphs)")
    ap.add_argument("--workload", default="rotation",
                    choices=("rotation", "ragged", "serve", "swirl", "copy"),
                    help="rotation: the Parareal benchmark (default); ragged: "
                         "256 independent solves in one batch; serve: Poisson "
                         "request streams through the serving engine; swirl: "
                         "2048^2 deformational swirl, 128 slices, coarse step "
                         "32 fine steps, each Parareal phase one hipGraph, "
                         "autotuned tiles; copy: the device copy rate of the "
                         "library's copy kernel (the practical HBM ceiling)")
    ap.add_argument("--count", type=int, default=256,
                    help="ragged: problems; serve: requests per rate")
    ap.add_argument("--rates", default="50,200,500,1000,2000",
                    help="serve: arrival rates (requests/s)")
    ap.add_argument("--adaptive-above", type=int, default=8,
                    help="serve: the adaptive policy's busy threshold")
    return ap.parse_args()


def pct(xs, q):
    return float(np.percentile(np.asarray(xs), q))


def timed(ctx, u0_dev_setter, fn, reps):
    out = []
    for _ in range(reps):
        u0_dev_setter()
        ctx.sync()
        t = time.perf_counter()
        fn()
        ctx.sync()
        out.append((time.perf_counter() - t) * 1e3)
    return out


def cpu_baseline(p, ic, nt, budget_s):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle  # noqa: E402  (CPU restatement: baseline leg only)
    threads = min(16, os.cpu_count() or 1)

    def run(nthreads, budget):
        pyoracle.set_threads(nthreads)
        u = pyoracle.exact(p, ic, 0.0)
        dt = 1.0 / nt
        steps = 0
        t0 = time.perf_counter()
        while steps < nt:
            u = pyoracle.step(p, u, steps * dt, dt)
            steps += 1
            if time.perf_counter() - t0 > budget:
                break
        return steps, time.perf_counter() - t0

    # whole solves, up to 5, while the budget lasts: the median rate (one
    # partial solve when a single one does not fit)
    rates, runs, spent = [], [], 0.0
    while len(rates) < 5 and spent < budget_s:
        steps, el = run(threads, budget_s - spent)
        spent += el
        rates.append(p.nx * p.ny * steps / el)
        runs.append((steps, el))
        if steps < nt:
            break
    value = float(np.median(rates))
    steps1, el1 = run(1, max(1.0, budget_s / 4))
    return {"value": value, "unit": "cell-updates/s",
            "cores": threads, "kind": "port",
            "sample": f"median of {len(runs)} runs of the plain sweep at "
                      f"{p.nx}x{p.ny} ("
                      + "; ".join(f"{st} of {nt} fine steps in {el:.1f} s" for st, el in runs)
                      + "; OpenMP over lines)",
            "single_thread_value": p.nx * p.ny * steps1 / el1,
            "single_thread_sample": f"{steps1} fine steps on 1 thread ({el1:.1f} s)"}


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def ragged(a):
    """256 independent solves of uneven size (seeded), one batch: ragged
    plain sweeps (the serial fine solutions), ragged Parareal, and the same
    plain solves one context at a time."""
    import ccfv
    from ccfv import workloads
    sol = workloads.ragged_batch(a.count)
    n = len(sol)
    t0 = [s.t0 for s in sol]
    t1 = [s.t1 for s in sol]
    nt = [s.nt_fine for s in sol]
    ns = [s.n_slices for s in sol]
    work = float(sum(s.cell_updates for s in sol))
    reps_n = max(1, min(a.steps, 5))

    def fill(b):
        for i, s in enumerate(sol):
            b[i].fill_exact(s.ic, s.t0)

    with ccfv.Batch([s.problem for s in sol]) as b:
        fill(b)
        b.plain(t0, t1, nt)  # warm-up: tables, arenas
        plain_ms = []
        for _ in range(reps_n):
            fill(b)
            t = time.perf_counter()
            b.plain(t0, t1, nt)
            plain_ms.append((time.perf_counter() - t) * 1e3)
        fine = [b[i].get_state() for i in range(n)]
        fill(b)
        b.parareal(t0, t1, nt, ns, a.kmax, a.tol)  # warm-up
        para_ms = []
        for _ in range(reps_n):
            fill(b)
            t = time.perf_counter()
            reps = b.parareal(t0, t1, nt, ns, a.kmax, a.tol)
            para_ms.append((time.perf_counter() - t) * 1e3)
        errs = [rel_l2(b[i].get_state(), fine[i]) for i in range(n)]
    # the same plain solves, one context after another
    t = time.perf_counter()
    for s in sol:
        with ccfv.Context(s.problem) as c:
            c.fill_exact(s.ic, s.t0)
            c.sync()
            c.plain_solve(s.t0, s.t1, s.nt_fine)
            c.sync()
    solo_ms = (time.perf_counter() - t) * 1e3
    pp, pa = pct(plain_ms, 50), pct(para_ms, 50)
    line = {
        "metric": "FP64 serial-equivalent cell-updates/s, ragged batch of "
                  "independent solves",
        "value": work / (pa * 1e-3), "unit": "cell-updates/s", "n_gpus": 1,
        "steps": reps_n, "ms_per_step": pa, "higher_is_better": True,
        "dtype": "f64", "data": "synthetic (seeded sums of 3 Gaussians)",
        "config": {"workload": f"ragged-{n}", "seed": workloads.RAGGED_SEED,
                   "grids": list(workloads.GRIDS), "slices": list(workloads.SLICES),
                   "T": workloads.T_END, "fine_cfl": workloads.FINE_CFL,
                   "kmax": a.kmax, "tol": a.tol,
                   "cells_total": int(sum(s.cells for s in sol)),
                   "serial_cell_updates": work},
        "parareal_p50_ms": pa, "parareal_p99_ms": pct(para_ms, 99),
        "plain_batch_p50_ms": pp, "plain_batch_value": work / (pp * 1e-3),
        "plain_one_by_one_ms": solo_ms,
        "batching_gain_plain": solo_ms / pp,
        "parareal_vs_plain_batch": pp / pa,
        "converged": int(sum(r.converged for r in reps)),
        "iterations_mean": float(np.mean([r.iterations for r in reps])),
        "rel_l2_vs_serial_fine_max": float(max(errs)),
        "rel_l2_vs_serial_fine_median": float(np.median(errs)),
    }
    print(json.dumps(line))


def serve_stream(a, rate, adaptive):
    """One Poisson stream of a.count requests at `rate` through a fresh
    engine; every result checked against its serial fine solution."""
    import ccfv
    from ccfv import serving, workloads
    stream = workloads.serving_stream(a.count, rate)
    arr = [(t, s.request(a.kmax, a.tol), None) for t, s in stream]
    with ccfv.Server(max_active=256, max_slices=64, adaptive_above=adaptive) as srv:
        # warm-up request (module load, first launches)
        srv.submit(stream[0][1].request(1, 0.0))
        while srv.step().active:
            pass
        while srv.next_done() >= 0:
            pass
        res = serving.run(srv, arr, keep_state=True)
    sol = [s for _, s in stream]
    with ccfv.Batch([s.problem for s in sol]) as b:
        for i, s in enumerate(sol):
            b[i].fill_exact(s.ic, s.t0)
        b.plain([s.t0 for s in sol], [s.t1 for s in sol], [s.nt_fine for s in sol])
        errs = [rel_l2(r.state, b[r.index].get_state()) for r in res.records]
    summ = res.summary()
    summ.update({"policy": "adaptive" if adaptive else "as requested",
                 "rate": rate, "offered_s": stream[-1][0],
                 "ran_serially": int(sum(r.report.n_slices == 0 for r in res.records)),
                 "converged": int(sum(r.report.converged for r in res.records)),
                 "status_ok": int(sum(r.status == 0 for r in res.records)),
                 "rel_l2_vs_serial_fine_max": float(max(errs)),
                 "within_1e-10": int(sum(e <= 1e-10 for e in errs)),
                 "cell_updates_per_s": float(sum(s.cell_updates for s in sol)) / res.wall})
    return summ


def serve(a):
    """Poisson request streams (ragged draws plus rotating-flow cases)
    through the serving engine, one stream per arrival rate, with requests
    run as submitted (Parareal) and with the adaptive policy (requests that
    arrive while the engine is busy run as the plain serial solve)."""
    out = {"metric": "serving: solves/s and p50/p99/p999 latency under Poisson "
                     "arrivals", "unit": "solves/s", "n_gpus": 1, "dtype": "f64",
           "data": "synthetic (seeded)", "kmax": a.kmax, "tol": a.tol,
           "adaptive_above": a.adaptive_above, "rates": []}
    for adaptive in (0, a.adaptive_above):
        for rate in [float(x) for x in a.rates.split(",")]:
            summ = serve_stream(a, rate, adaptive)
            out["rates"].append(summ)
            print(json.dumps(summ), file=sys.stderr, flush=True)
    best = max(out["rates"], key=lambda r: r["solves_per_s"])
    out["value"] = best["solves_per_s"]
    print(json.dumps(out))


def copy_rate(a):
    """Copies for the practical HBM ceiling: the library's copy kernel
    (ccfv_get_state_device) from one 2048^2 field (32 MiB) to the next, round
    a ring of --ring fields (default 8: 256 MiB), 16 B moved per double.  Each call
    synchronizes, so the rate comes from the kernel durations under
    rocprofv3 --kernel-trace, not from a host clock."""
    import ccfv
    from ccfv import types as T
    big = T.rotation_problem(2048)
    nr = max(2, a.ring)
    ring = [ccfv.Context(big) for _ in range(nr)]
    for c in ring:
        c.fill_exact(T.rotation_ic(), 0.0)
        c.sync()
    ptrs = [c.state_ptr() for c in ring]
    copies = nr * max(1, a.steps)
    for i in range(nr * max(1, a.warmup) + copies):
        ring[i % nr].get_state_device(ptrs[(i + 1) % nr])
    print(json.dumps({"workload": "copy", "ring_fields": nr,
                      "copies": copies + nr * max(1, a.warmup),
                      "bytes_per_copy": 16 * 2048 * 2048,
                      "kernel": "ccfv::copy_kernel"}))


def swirl(a):
    """Large Parareal: 2048^2 swirl u = sin^2(pi x) sin(2 pi y) cos(pi t/5),
    v = -sin^2(pi y) sin(2 pi x) cos(pi t/5), nu = 1e-4, cosine bell of
    radius 0.15 at (0.5, 0.75), T = 5 in 4096 fine steps, 128 slices of 32
    fine steps; every Parareal phase replayed as one hipGraph; launch shapes
    from the autotuner."""
    import ccfv
    from ccfv import types as T
    n = 2048 if a.n == 512 else a.n
    nt, ns = 4096, 128
    p, ic = T.swirl_problem(n, T=5.0, nu=1e-4), T.cosine_bell_ic()
    ctx = ccfv.Context(p)
    ctx.fill_exact(ic, 0.0)
    if a.probe:
        # short run for kernel traces: 64 plain steps and one Parareal
        # iteration on one stream, default launch shapes, no graphs
        ctx.set_option(T.OPT_GRAPHS, 0)
        ctx.set_option(T.OPT_PIPELINE, 0)
        t = time.perf_counter()
        ctx.plain_solve(0.0, 5.0 * 64 / nt, 64)
        ctx.sync()
        plain_ms = (time.perf_counter() - t) * 1e3
        ctx.fill_exact(ic, 0.0)
        t = time.perf_counter()
        r = ctx.parareal_solve(0.0, 5.0, nt, ns, 1, 0.0)
        ctx.sync()
        print(json.dumps({"probe": "ok", "workload": f"swirl-{n}", "plain_64_ms": plain_ms,
                          "parareal_k1_ms": (time.perf_counter() - t) * 1e3,
                          "iterations": r.iterations}))
        ctx.close()
        return
    u0 = ctx.get_state()
    t_tune = time.perf_counter()
    tiles = ctx.autotune(ns, 5.0 / nt)
    t_tune = time.perf_counter() - t_tune
    reps = max(1, min(a.steps, 3))

    def reset():
        ctx.set_state(u0)

    plain_ms = timed(ctx, reset, lambda: ctx.plain_solve(0.0, 5.0, nt), reps)
    reset()
    ctx.plain_solve(0.0, 5.0, nt)
    uref = ctx.get_state()
    ctx.set_option(T.OPT_GRAPHS, 1)
    box = {}

    def para():
        box["r"] = ctx.parareal_solve(0.0, 5.0, nt, ns, a.kmax, a.tol)

    timed(ctx, reset, para, 1)  # captures the phases
    para_ms = timed(ctx, reset, para, reps)
    rep = box["r"]
    rel = rel_l2(ctx.get_state(), uref)
    cells = n * n
    pa, pp = pct(para_ms, 50), pct(plain_ms, 50)
    line = {
        "metric": "FP64 cell-updates/s, large Parareal with graph capture",
        "value": cells * nt / (pa * 1e-3), "unit": "cell-updates/s", "n_gpus": 1,
        "steps": reps, "ms_per_step": pa, "higher_is_better": True, "dtype": "f64",
        "data": "synthetic (cosine bell)",
        "config": {"workload": f"swirl-{n}-parareal", "grid": [n, n], "T": 5.0,
                   "fine_steps": nt, "slices": ns, "J": nt // ns, "kmax": a.kmax,
                   "tol": a.tol, "nu": 1e-4, "graphs": True},
        "tiles": tiles, "autotune_s": t_tune,
        "parareal_p50_ms": pa, "plain_p50_ms": pp,
        "plain_value": cells * nt / (pp * 1e-3), "speedup_vs_plain": pp / pa,
        "parareal_iterations": rep.iterations,
        "parareal_converged": bool(rep.converged),
        "parareal_increments": [float(x) for x in rep.increments],
        "rel_l2_vs_serial_fine": rel,
    }
    print(json.dumps(line))


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU "
                          "sharding of the slice chain is designed but not "
                          "built", "n_gpus": a.gpus}))
        sys.exit(2)
    if a.workload == "ragged":
        return ragged(a)
    if a.workload == "serve":
        return serve(a)
    if a.workload == "swirl":
        return swirl(a)
    if a.workload == "copy":
        return copy_rate(a)
    import ccfv
    from ccfv import types as T

    p, ic = T.rotation_problem(a.n), T.rotation_ic()
    if a.probe:
        # calibration: copy_kernel over 2048^2 doubles (32 MiB in, 32 MiB
        # out) between two fields that are not L2 resident
        big = T.rotation_problem(2048)
        c1, c2 = ccfv.Context(big), ccfv.Context(big)
        c1.fill_exact(ic, 0.0)
        junk = ccfv.Context(big)
        junk.fill_exact(ic, 0.0)  # evict c1's field from the caches
        c1.get_state_device(c2.state_ptr())
        c1.sync()
        ctx = ccfv.Context(p)
        ctx.fill_exact(ic, 0.0)
        r = ctx.parareal_solve(0.0, 1.0, a.nt, a.slices, 1, a.tol)
        print(json.dumps({"probe": "ok", "iterations": r.iterations}))
        return
    ctx = ccfv.Context(p)
    ctx.fill_exact(ic, 0.0)
    u0 = ctx.get_state()
    cells = a.n * a.n

    def reset():
        ctx.set_state(u0)

    # serial fine solution (plain sweep) and its timing
    plain_ms = timed(ctx, reset, lambda: ctx.plain_solve(0.0, 1.0, a.nt),
                     a.warmup)
    plain_ms = timed(ctx, reset, lambda: ctx.plain_solve(0.0, 1.0, a.nt),
                     a.steps)
    # the same solve captured once as a hipGraph and replayed (identical
    # results; the graph's launches skip the per-launch submission work)
    plain_graph_ms = [float("nan")]
    if not a.no_plain_graphs:
        ctx.set_option(T.OPT_PLAIN_GRAPHS, 1)
        timed(ctx, reset, lambda: ctx.plain_solve(0.0, 1.0, a.nt), a.warmup)
        plain_graph_ms = timed(ctx, reset, lambda: ctx.plain_solve(0.0, 1.0, a.nt),
                               a.steps)
        ctx.set_option(T.OPT_PLAIN_GRAPHS, 0)
    reset()
    ctx.plain_solve(0.0, 1.0, a.nt)
    uref = ctx.get_state()

    rep_box = {}

    def para():
        rep_box["r"] = ctx.parareal_solve(0.0, 1.0, a.nt, a.slices, a.kmax,
                                          a.tol)

    timed(ctx, reset, para, a.warmup)
    para_ms = timed(ctx, reset, para, a.steps)
    rep = rep_box["r"]
    upara = ctx.get_state()
    rel = float(np.linalg.norm(upara - uref) / np.linalg.norm(uref))
    maxabs = float(np.abs(upara - uref).max())

    # instrumented solve: per-kernel-class hipEvent times on the ctx stream,
    # one stream (isolated kernel durations)
    reset()
    ctx.set_option(T.OPT_PIPELINE, 0)
    ctx.profile_enable(True)
    para()
    stats = ctx.profile_read()
    ctx.profile_enable(False)
    ctx.set_option(T.OPT_PIPELINE, 1)
    dom_name = max(stats, key=lambda k: stats[k]["bytes"])
    dom = stats[dom_name]
    mean_ms = dom["total_ms"] / max(dom["launches"], 1)
    bytes_per_launch = dom["bytes"] / max(dom["launches"], 1)
    achieved = bytes_per_launch / (mean_ms * 1e-3) / 1e9

    p50, p99 = pct(para_ms, 50), pct(para_ms, 99)
    pp50, pp99 = pct(plain_ms, 50), pct(plain_ms, 99)
    value = cells * a.nt / (p50 * 1e-3)
    plain_value = cells * a.nt / (pp50 * 1e-3)

    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(p, ic, a.nt, a.cpu_seconds)

    line = {
        "metric": "FP64 cell-updates/s batched Parareal vs plain sweep @512²; "
                  "p50/p99 solve ms",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": p50,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (Gaussian hump cell averages)",
        "config": {"workload": f"rotation-{a.n}-parareal", "grid": [a.n, a.n],
                   "T": 1.0, "fine_steps": a.nt, "slices": a.slices,
                   "J": a.nt // a.slices, "kmax": a.kmax, "tol": a.tol,
                   "nu": 1e-3, "bc": "periodic"},
        "p50_ms": p50,
        "p99_ms": p99,
        "plain_value": plain_value,
        "plain_p50_ms": pp50,
        "plain_p99_ms": pp99,
        "plain_graph_p50_ms": (None if a.no_plain_graphs
                               else pct(plain_graph_ms, 50)),
        "speedup_vs_plain": pp50 / p50,
        "parareal_iterations": rep.iterations,
        "parareal_converged": bool(rep.converged),
        "parareal_increments": [float(x) for x in rep.increments],
        "rel_l2_vs_serial_fine": rel,
        "max_abs_vs_serial_fine": maxabs,
        "fine_field_steps": rep.fine_steps_run,
        "coarse_field_steps": rep.coarse_steps_run,
        "kernel_classes": {k: {"launches": v["launches"],
                               "mean_us": 1e3 * v["total_ms"] / max(v["launches"], 1),
                               "GBps": (v["bytes"] / max(v["total_ms"], 1e-12) / 1e6)}
                           for k, v in stats.items() if v["launches"]},
        "roofline": {"bound": "hbm", "kernel": dom_name,
                     "selection": "kernel class moving the most HBM bytes, "
                                  "one-stream instrumented solve",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": (bytes_per_launch * TRAFFIC_RATIO[dom_name]
                                 if dom_name in TRAFFIC_RATIO else None),
                     "mean_launch_us": mean_ms * 1e3,
                     "bytes_per_launch": bytes_per_launch},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
