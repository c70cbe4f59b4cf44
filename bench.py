#!/usr/bin/env python3
"""Benchmark: FP64 cell-updates/s of the 512^2 rotation workload on one
MI355X — batched Parareal against the plain serial-in-time sweep — with p50 /
p99 solve latency, at matched accuracy.

Workload (one "step" = one whole solve): periodic [0,1]^2, 512x512 FP64 cell
averages, solid-body rotation u=-2 pi (y-1/2), v=2 pi (x-1/2), nu=1e-3,
Gaussian hump at (0.5, 0.75) with standard deviation 0.05, T=1 in 1024 fine
steps.  Methods:
  plain        the serial fine loop (1024 Strang steps, two launches each);
  plain-graph  the same launches captured once as a hipGraph and replayed
               (bitwise the same result);
  parareal     64 slices, coarse dt = 16 fine dt, at most 8 iterations, stop
               at a relative L2 increment of 1e-10; all fine propagators of
               an iteration batched; the reference's coarse propagator.
A method is accepted when its result is within 1e-10 relative L2 of the
serial fine solution (the plain methods are that solution).
value = serial-equivalent cell-updates/s = 512^2 * 1024 / p50 of the fastest
accepted method; every method is reported under "methods".  Inputs are
resident in HBM before each timed solve.

Roofline: the kernel class that takes the most time in the headline method's
solve.  Each one-field sweep class is timed in isolation (ccfv_time_sweep:
back-to-back launches between two hipEvents on the context's stream, the
same kernel, grid and arguments the solve uses) and its algorithmic bytes
(8 B in + 8 B out per cell) over that mean launch time give `achieved`.  The
launch-rate floor of the same grid (ccfv_dispatch_floor: an empty kernel, and
a copy of the field with the sweep's memory pattern) is reported beside it.
The Parareal solve's classes are timed per launch in one extra instrumented
solve on one stream.

cpu_baseline: the same scheme (the CPU restatement) on every host core this
process may use, OpenMP over lines, on a bounded number of steps of the plain
sweep, plus a one-thread sample; the CPU model is recorded.

Legs (short runs of the other workloads, as extra keys): the same 512^2
plan on a translating flow where the reference's Parareal converges (its
speed-up over the plain sweep at matched accuracy), the 2048^2 swirl
(plain and Parareal with graph capture and autotuned tiles), the seeded
ragged batch of 256 independent solves, and one Poisson rate of the serving
stream.  --workload ragged|serve|swirl run the full versions alone.
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
# that make the library's copy kernel of a known size read its true bytes)
# (profiles/rotation512_fine_sweep_pmc.md, profiles/rotation512_plain_pmc.md).
# A calibrated constant, not a counter reading of this run.
TRAFFIC_RATIO = {"x_batch": 1.002, "y_batch": 1.002, "y_single": 1.064,
                 "x_single": 1.059, "x_epilog": 1.029}
TRAFFIC_SOURCE = ("calibrated constant: rocprofv3 FETCH_SIZE x2 + WRITE_SIZE x1 per "
                  "dispatch over the algorithmic bytes (profiles/*_pmc.md); not read "
                  "in this run")


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
    ap.add_argument("--no-legs", action="store_true",
                    help="skip the swirl, ragged and serving legs")
    ap.add_argument("--no-graphs", "--no-plain-graphs", dest="no_graphs",
                    action="store_true",
                    help="capture no hipGraph anywhere (the plain sweep replayed "
                         "from a graph is skipped, isolated timings run eagerly, "
                         "the swirl leg issues its phases directly); for "
                         "rocprofv3 --kernel-trace runs: that profiler's process "
                         "has ended in a host segmentation fault in the capture "
                         "of the 2048-launch plain solve")
    ap.add_argument("--ring", type=int, default=8,
                    help="--workload copy: fields of 32 MiB in the ring (8: 256 MiB, "
                         "more than the last-level cache holds)")
    ap.add_argument("--probe", action="store_true",
                    help="short run for counter collection: one 32 MiB "
                         "calibration copy, 64 plain steps and one Parareal "
                         "iteration (--workload swirl: 64 plain steps and one "
                         "Parareal iteration, one stream, no graphs)")
    ap.add_argument("--workload", default="rotation",
                    choices=("rotation", "translation", "ragged", "serve", "swirl", "copy",
                             "accuracy"),
                    help="rotation: the headline (default, with short legs of the "
                         "others); accuracy: the plain method on the periodic sin*sin "
                         "problem, errors and orders on 32^2..256^2; "
                         "translation: the same plan on a translating flow, "
                         "where Parareal converges; ragged: 256 independent solves in one batch; "
                         "serve: Poisson request streams through the serving "
                         "engine; swirl: 2048^2 deformational swirl, 128 slices, "
                         "coarse step 32 fine steps, each Parareal phase one "
                         "hipGraph, autotuned tiles; copy: the library's copy "
                         "kernel (the practical HBM ceiling)")
    ap.add_argument("--count", type=int, default=256,
                    help="ragged: problems; serve: requests per rate")
    ap.add_argument("--rates", default="50,200,500,1000,2000",
                    help="serve: arrival rates (requests/s)")
    ap.add_argument("--leg-rate", type=float, default=1000.0,
                    help="the serving leg's arrival rate (requests/s)")
    ap.add_argument("--adaptive-above", type=int, default=8,
                    help="serve: the adaptive policy's busy threshold")
    return ap.parse_args()


def pct(xs, q):
    return float(np.percentile(np.asarray(xs), q))


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def timed(ctx, reset, fn, reps):
    out = []
    for _ in range(reps):
        reset()
        ctx.sync()
        t = time.perf_counter()
        fn()
        ctx.sync()
        out.append((time.perf_counter() - t) * 1e3)
    return out


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------
# CPU baseline
# ---------------------------------------------------------------------------
def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(p, ic, nt, budget_s):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle  # noqa: E402  (CPU restatement: baseline leg only)
    try:
        threads = len(os.sched_getaffinity(0))
    except AttributeError:
        threads = os.cpu_count() or 1

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
    return {"value": value, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
            "sample": f"median of {len(runs)} runs of the plain sweep at "
                      f"{p.nx}x{p.ny} ("
                      + "; ".join(f"{st} of {nt} fine steps in {el:.1f} s" for st, el in runs)
                      + f"; OpenMP over lines on {threads} threads)",
            "single_thread_value": p.nx * p.ny * steps1 / el1,
            "single_thread_sample": f"{steps1} fine steps on 1 thread ({el1:.1f} s)"}


# ---------------------------------------------------------------------------
# the headline: 512^2 rotation, plain sweep and Parareal
# ---------------------------------------------------------------------------
def hm_p50(methods, name):
    return methods[name]["p50_ms"]


def rotation(a):
    import ccfv
    from ccfv import types as T

    p, ic = T.rotation_problem(a.n), T.rotation_ic()
    ctx = ccfv.Context(p)
    ctx.fill_exact(ic, 0.0)
    u0 = ctx.get_state()
    cells = a.n * a.n
    dt = 1.0 / a.nt

    def reset():
        ctx.set_state(u0)

    plain = lambda: ctx.plain_solve(0.0, 1.0, a.nt)  # noqa: E731
    timed(ctx, reset, plain, a.warmup)
    plain_ms = timed(ctx, reset, plain, a.steps)
    reset()
    plain()
    uref = ctx.get_state()
    methods = {"plain": {"p50_ms": pct(plain_ms, 50), "p99_ms": pct(plain_ms, 99),
                         "rel_l2_vs_serial_fine": 0.0, "accepted": True}}
    if not a.no_graphs:
        ctx.set_option(T.OPT_PLAIN_GRAPHS, 1)
        timed(ctx, reset, plain, a.warmup)
        g_ms = timed(ctx, reset, plain, a.steps)
        ug = ctx.get_state()
        ctx.set_option(T.OPT_PLAIN_GRAPHS, 0)
        methods["plain-graph"] = {"p50_ms": pct(g_ms, 50), "p99_ms": pct(g_ms, 99),
                                  "rel_l2_vs_serial_fine": rel_l2(ug, uref),
                                  "accepted": bool(np.array_equal(ug, uref))}
    log("rotation: plain done")

    box = {}

    def para():
        box["r"] = ctx.parareal_solve(0.0, 1.0, a.nt, a.slices, a.kmax, a.tol)

    timed(ctx, reset, para, a.warmup)
    para_ms = timed(ctx, reset, para, a.steps)
    rep = box["r"]
    upara = ctx.get_state()
    rel = rel_l2(upara, uref)
    accepted = bool(rep.converged) and rel <= a.tol
    methods["parareal"] = {
        "p50_ms": pct(para_ms, 50), "p99_ms": pct(para_ms, 99),
        "rel_l2_vs_serial_fine": rel,
        "max_abs_vs_serial_fine": float(np.abs(upara - uref).max()),
        "accepted": accepted, "iterations": rep.iterations,
        "converged": bool(rep.converged),
        "increments": [float(x) for x in rep.increments],
        "fine_field_steps": rep.fine_steps_run, "coarse_field_steps": rep.coarse_steps_run,
        "speedup_vs_plain": methods["plain"]["p50_ms"] / pct(para_ms, 50)}
    for m in methods.values():
        m["value"] = cells * a.nt / (m["p50_ms"] * 1e-3)
    log("rotation: parareal done")

    # one-field sweep classes of the plain solve, in isolation: 2048
    # back-to-back launches issued one by one (the duration rocprofv3 reports
    # per dispatch) and replayed from a hipGraph (as the plain-graph method
    # issues them), beside the launch-rate floor of the same grid: an empty
    # kernel, and a copy of the field with the sweep's memory pattern and no
    # arithmetic
    g = not a.no_graphs
    n_x, n_y = a.nt + 1, a.nt  # x launches: first x1, nt-1 fused, last x3
    bytes_one = 16.0 * cells
    plain_classes = {}
    for name, d, fused, kname in (("x_single", 0, True,
                                   "fast_x_kernel<8,0,true,0> (one field, two x sub-steps)"),
                                  ("y_single", 1, False, "fast_y_kernel<8,0> (one field)")):
        us = ctx.time_sweep(d, fused, 1, dt, 2048)
        c = {"kernel": kname, "launches_per_solve": n_x if d == 0 else n_y,
             "mean_us": us, "GBps": bytes_one / (us * 1e-6) / 1e9,
             "empty_kernel_floor_us": ctx.dispatch_floor(d, 0, 2048),
             "copy_floor_us": ctx.dispatch_floor(d, 1, 2048)}
        if g:
            c["mean_us_graph"] = ctx.time_sweep(d, fused, 1, dt, 2048, graph=True)
            c["empty_kernel_floor_us_graph"] = ctx.dispatch_floor(d, 0, 2048, graph=True)
            c["copy_floor_us_graph"] = ctx.dispatch_floor(d, 1, 2048, graph=True)
        plain_classes[name] = c
    # the plain solve's pattern without its arithmetic: y and x copies
    # alternating in place, each reading what the previous one wrote
    step_floor = {"launches": 2048, "us_per_step": 2 * ctx.dispatch_floor(0, 2, 2048)}
    if g:
        step_floor["us_per_step_graph"] = 2 * ctx.dispatch_floor(0, 2, 2048, graph=True)
    log("rotation: isolated classes done")

    # Parareal classes, per launch, one instrumented solve on one stream
    reset()
    ctx.set_option(T.OPT_PIPELINE, 0)
    ctx.profile_enable(True)
    para()
    stats = ctx.profile_read()
    ctx.profile_enable(False)
    ctx.set_option(T.OPT_PIPELINE, 1)
    para_classes = {k: {"launches": v["launches"],
                        "mean_us": 1e3 * v["total_ms"] / max(v["launches"], 1),
                        "total_ms": v["total_ms"],
                        "GBps": v["bytes"] / max(v["total_ms"], 1e-12) / 1e6}
                    for k, v in stats.items() if v["launches"]}
    ctx.close()

    # headline: the fastest accepted method
    best = min((m for m in methods if methods[m]["accepted"]),
               key=lambda m: methods[m]["p50_ms"])
    if best == "parareal":
        dom = max(para_classes, key=lambda k: para_classes[k]["total_ms"])
        c = para_classes[dom]
        bpl = stats[dom]["bytes"] / max(stats[dom]["launches"], 1)
        roof = {"kernel": dom, "mean_launch_us": c["mean_us"], "bytes_per_launch": bpl,
                "selection": "kernel class with the most time, one-stream instrumented "
                             "Parareal solve (hipEvents around every launch)"}
    else:
        dom = max(plain_classes, key=lambda k: plain_classes[k]["mean_us"] *
                  plain_classes[k]["launches_per_solve"])
        c = plain_classes[dom]
        nl = n_x + n_y
        roof = {"kernel": dom, "kernel_name": c["kernel"], "mean_launch_us": c["mean_us"],
                "bytes_per_launch": bytes_one,
                "selection": "kernel class with the most time in the plain solve, timed in "
                             "isolation (ccfv_time_sweep: 2048 back-to-back launches "
                             "between two hipEvents on the context's stream)",
                "copy_floor_us": c["copy_floor_us"],
                "empty_kernel_floor_us": c["empty_kernel_floor_us"],
                "frac_of_copy_floor": c["copy_floor_us"] / c["mean_us"],
                "in_solve_mean_launch_us": hm_p50(methods, best) * 1e3 / nl}
    achieved = roof["bytes_per_launch"] / (roof["mean_launch_us"] * 1e-6) / 1e9
    roof.update({"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                 "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                 "traffic": (roof["bytes_per_launch"] * TRAFFIC_RATIO[dom]
                             if dom in TRAFFIC_RATIO else None),
                 "traffic_source": TRAFFIC_SOURCE})
    # the Parareal fine batches' sweep (the bandwidth-bound class) beside it
    fb = None
    if "x_batch" in stats and stats["x_batch"]["launches"]:
        s = stats["x_batch"]
        bpl = s["bytes"] / s["launches"]
        mean_us = 1e3 * s["total_ms"] / s["launches"]
        fb = {"kernel": "x_batch", "mean_launch_us": mean_us, "bytes_per_launch": bpl,
              "achieved": bpl / (mean_us * 1e-6) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
              "frac": bpl / (mean_us * 1e-6) / 1e9 / HBM_PEAK_GBS,
              "traffic": bpl * TRAFFIC_RATIO["x_batch"]}

    hm = methods[best]
    return {
        "metric": "FP64 cell-updates/s batched Parareal vs plain sweep @512²; "
                  "p50/p99 solve ms",
        "value": hm["value"], "unit": "cell-updates/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": hm["p50_ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (Gaussian hump cell averages)",
        "config": {"workload": f"rotation-{a.n}", "grid": [a.n, a.n], "T": 1.0,
                   "fine_steps": a.nt, "slices": a.slices, "J": a.nt // a.slices,
                   "kmax": a.kmax, "tol": a.tol, "nu": 1e-3, "bc": "periodic",
                   "method": best},
        "method": best, "accepted": True,
        "p50_ms": hm["p50_ms"], "p99_ms": hm["p99_ms"],
        "methods": methods,
        "parareal_accepted": accepted,
        "roofline": roof,
        "roofline_parareal_fine_batch": fb,
        "kernel_classes": {"plain": plain_classes, "parareal": para_classes},
        "plain_step_floor": dict(step_floor, **{
            "measured_us_per_step": {m: 1e3 * methods[m]["p50_ms"] / a.nt
                                     for m in methods if m.startswith("plain")},
            "what": "two dependent copy launches per step with the sweeps' memory patterns "
                    "and block mapping, in place, no arithmetic (ccfv_dispatch_floor kind 2)"}),
    }


# ---------------------------------------------------------------------------
# the other workloads
# ---------------------------------------------------------------------------
def ragged(a, reps=None, full=True):
    """256 independent solves of uneven size (seeded), one batch: ragged
    plain sweeps (the serial fine solutions) and ragged Parareal with the
    serial finish (a problem that reaches kmax above tol finishes with its
    serial fine solve), each issued three ways: stage launches with the
    default shapes, with shapes tuned per line class (ccfv_batch_autotune),
    and tuned with every phase replayed as one hipGraph; and (full) the same
    plain solves one context at a time."""
    import ccfv
    from ccfv import types as T
    from ccfv import workloads
    sol = workloads.ragged_batch(a.count)
    n = len(sol)
    t0 = [s.t0 for s in sol]
    t1 = [s.t1 for s in sol]
    nt = [s.nt_fine for s in sol]
    ns = [s.n_slices for s in sol]
    work = float(sum(s.cell_updates for s in sol))
    reps_n = reps or max(1, min(a.steps, 5))

    def fill(b):
        for i, s in enumerate(sol):
            b[i].fill_exact(s.ic, s.t0)

    variants = {}
    with ccfv.Batch([s.problem for s in sol]) as b:
        b.set_option(T.OPT_SERIAL_FINISH, 1)
        fine = None
        for name in ("stage launches", "tuned shapes", "tuned shapes, graphs"):
            if name == "tuned shapes":
                t = time.perf_counter()
                b.autotune(t0, t1, nt, batch_fields=16)
                tune_s = time.perf_counter() - t
            if name.endswith("graphs"):
                if a.no_graphs:
                    continue
                b.set_option(T.OPT_GRAPHS, 1)
            # untimed solves first: tables, arenas, and (graphs) the first
            # sighting and the capture of every phase
            plain_ms, para_ms = [], []
            for r in range(2 + reps_n):
                fill(b)
                t = time.perf_counter()
                b.plain(t0, t1, nt)
                if r >= 2:
                    plain_ms.append((time.perf_counter() - t) * 1e3)
            if fine is None:
                fine = [b[i].get_state() for i in range(n)]
            for r in range(2 + reps_n):
                fill(b)
                t = time.perf_counter()
                reps_ = b.parareal(t0, t1, nt, ns, a.kmax, a.tol)
                if r >= 2:
                    para_ms.append((time.perf_counter() - t) * 1e3)
            errs = [rel_l2(b[i].get_state(), fine[i]) for i in range(n)]
            variants[name] = {
                "plain_batch_p50_ms": pct(plain_ms, 50),
                "parareal_p50_ms": pct(para_ms, 50), "parareal_p99_ms": pct(para_ms, 99),
                "converged": int(sum(r.converged for r in reps_)),
                "finished_serially": int(sum(r.serial_finish for r in reps_)),
                "iterations_mean": float(np.mean([r.iterations for r in reps_])),
                "within_1e-10": int(sum(e <= 1e-10 for e in errs)),
                "rel_l2_vs_serial_fine_max": float(max(errs)),
                "rel_l2_vs_serial_fine_median": float(np.median(errs))}
            log(f"ragged {name}: {json.dumps(variants[name])}")
        variants["tuned shapes"]["autotune_s"] = tune_s
    solo_ms = None
    if full:
        t = time.perf_counter()
        for s in sol:
            with ccfv.Context(s.problem) as c:
                c.fill_exact(s.ic, s.t0)
                c.sync()
                c.plain_solve(s.t0, s.t1, s.nt_fine)
                c.sync()
        solo_ms = (time.perf_counter() - t) * 1e3
    best = min(variants, key=lambda k: min(variants[k]["plain_batch_p50_ms"],
                                          variants[k]["parareal_p50_ms"]))
    v = variants[best]
    pp, pa = v["plain_batch_p50_ms"], v["parareal_p50_ms"]
    return {
        "metric": "FP64 serial-equivalent cell-updates/s, ragged batch of "
                  "independent solves",
        "value": work / (min(pp, pa) * 1e-3), "unit": "cell-updates/s", "n_gpus": 1,
        "steps": reps_n, "higher_is_better": True, "dtype": "f64",
        "data": "synthetic (seeded sums of 3 Gaussians)",
        "config": {"workload": f"ragged-{n}", "seed": workloads.RAGGED_SEED,
                   "grids": list(workloads.GRIDS), "slices": list(workloads.SLICES),
                   "T": workloads.T_END, "fine_cfl": workloads.FINE_CFL,
                   "kmax": a.kmax, "tol": a.tol, "serial_finish": True,
                   "cells_total": int(sum(s.cells for s in sol)),
                   "serial_cell_updates": work},
        "best_variant": best,
        "plain_batch_p50_ms": pp, "plain_batch_value": work / (pp * 1e-3),
        "parareal_p50_ms": pa, "parareal_value": work / (pa * 1e-3),
        "parareal_vs_plain_batch": pp / pa,
        "plain_one_by_one_ms": solo_ms,
        "batching_gain_plain": (solo_ms / pp) if solo_ms else None,
        "within_1e-10": v["within_1e-10"], "finished_serially": v["finished_serially"],
        "variants": variants,
    }


def serve_stream(a, rate, adaptive, serial=False):
    """One Poisson stream of a.count requests at `rate` through a fresh
    engine with the serial finish; every result checked against its serial
    fine solution.  serial: every request is submitted as the plain serial
    solve (n_slices = 0)."""
    import ccfv
    from ccfv import serving, workloads
    stream = workloads.serving_stream(a.count, rate)
    arr = []
    for t, s in stream:
        r = s.request(a.kmax, a.tol)
        if serial:
            r.n_slices = 0
        arr.append((t, r, None))
    with ccfv.Server(max_active=256, max_slices=64, adaptive_above=adaptive,
                     serial_finish=1) as srv:
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
    summ.update({"policy": "serial" if serial else ("adaptive" if adaptive else "as requested"),
                 "serial_finish": True, "rate": rate, "offered_s": stream[-1][0],
                 "ran_serially": int(sum(r.report.n_slices == 0 for r in res.records)),
                 "finished_serially": int(sum(r.report.serial_finish for r in res.records)),
                 "converged": int(sum(r.report.converged for r in res.records)),
                 "status_ok": int(sum(r.status == 0 for r in res.records)),
                 "rel_l2_vs_serial_fine_max": float(max(errs)),
                 "within_1e-10": int(sum(e <= 1e-10 for e in errs)),
                 "cell_updates_per_s": float(sum(s.cell_updates for s in sol)) / res.wall})
    return summ


def serve(a):
    """Poisson request streams (ragged draws plus rotating-flow cases)
    through the serving engine, one stream per arrival rate, with requests
    run as submitted (Parareal, serial finish), with the adaptive policy
    (requests that arrive while the engine is busy run as the plain serial
    solve), and all submitted as plain serial solves."""
    out = {"metric": "serving: solves/s and p50/p99/p999 latency under Poisson "
                     "arrivals", "unit": "solves/s", "n_gpus": 1, "dtype": "f64",
           "data": "synthetic (seeded)", "kmax": a.kmax, "tol": a.tol,
           "adaptive_above": a.adaptive_above, "rates": []}
    for adaptive, serial in ((0, False), (a.adaptive_above, False), (0, True)):
        for rate in [float(x) for x in a.rates.split(",")]:
            summ = serve_stream(a, rate, adaptive, serial)
            out["rates"].append(summ)
            log(json.dumps(summ))
    best = max(out["rates"], key=lambda r: r["solves_per_s"])
    out["value"] = best["solves_per_s"]
    return out


def copy_rate(a):
    """Copies for the practical HBM ceiling: the library's copy kernel
    (ccfv_get_state_device) from one 2048^2 field (32 MiB) to the next, round
    a ring of --ring fields (default 8: 256 MiB), 16 B moved per double.  Each
    call synchronizes, so the rate comes from the kernel durations under
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
    return {"workload": "copy", "ring_fields": nr,
            "copies": copies + nr * max(1, a.warmup),
            "bytes_per_copy": 16 * 2048 * 2048, "kernel": "ccfv::copy_kernel"}


def swirl(a, reps=None):
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
        out = {"probe": "ok", "workload": f"swirl-{n}", "plain_64_ms": plain_ms,
               "parareal_k1_ms": (time.perf_counter() - t) * 1e3,
               "iterations": r.iterations}
        ctx.close()
        return out
    u0 = ctx.get_state()
    t_tune = time.perf_counter()
    tiles = ctx.autotune(ns, 5.0 / nt)
    t_tune = time.perf_counter() - t_tune
    reps = reps or max(1, min(a.steps, 3))

    def reset():
        ctx.set_state(u0)

    plain_ms = timed(ctx, reset, lambda: ctx.plain_solve(0.0, 5.0, nt), reps)
    reset()
    ctx.plain_solve(0.0, 5.0, nt)
    uref = ctx.get_state()
    ctx.set_option(T.OPT_GRAPHS, 0 if a.no_graphs else 1)
    box = {}

    def para():
        box["r"] = ctx.parareal_solve(0.0, 5.0, nt, ns, a.kmax, a.tol)

    timed(ctx, reset, para, 1)  # captures the phases
    para_ms = timed(ctx, reset, para, reps)
    rep = box["r"]
    rel = rel_l2(ctx.get_state(), uref)
    ctx.close()
    cells = n * n
    pa, pp = pct(para_ms, 50), pct(plain_ms, 50)
    accepted = bool(rep.converged) and rel <= a.tol
    return {
        "metric": "FP64 cell-updates/s, large Parareal with graph capture",
        "value": cells * nt / (pp * 1e-3) if not accepted else cells * nt / (min(pa, pp) * 1e-3),
        "unit": "cell-updates/s", "n_gpus": 1,
        "steps": reps, "higher_is_better": True, "dtype": "f64",
        "data": "synthetic (cosine bell)",
        "config": {"workload": f"swirl-{n}", "grid": [n, n], "T": 5.0,
                   "fine_steps": nt, "slices": ns, "J": nt // ns, "kmax": a.kmax,
                   "tol": a.tol, "nu": 1e-4, "graphs": not a.no_graphs},
        "tiles": tiles, "autotune_s": t_tune,
        "plain_p50_ms": pp, "plain_value": cells * nt / (pp * 1e-3),
        "parareal_p50_ms": pa, "parareal_value": cells * nt / (pa * 1e-3),
        "speedup_vs_plain": pp / pa,
        "parareal_iterations": rep.iterations,
        "parareal_converged": bool(rep.converged),
        "parareal_accepted": accepted,
        "parareal_increments": [float(x) for x in rep.increments],
        "rel_l2_vs_serial_fine": rel,
    }


def accuracy(a, reps=None):
    """The plain method on its accuracy workload: periodic [0,1]^2, velocity
    (1, 0.5), nu = 0.01, u0 = sin(2 pi x) sin(2 pi y), T = 1 at CFL 0.8
    (40 / 80 / 160 / 320 steps on 32^2 .. 256^2), against the exact
    decaying translated solution's cell averages (the library's own closed
    form, ccfv_fill_exact): relative L2 and RMS errors and the observed
    orders; the same at a fixed small step (2560 steps), where the spatial
    order shows; and the 128^2 solve's rate (p50 of `reps` solves)."""
    import ccfv
    from ccfv import types as T
    reps = reps or max(1, min(a.steps, 10))
    rows, fixed = [], []
    for n in (32, 64, 128, 256):
        p, ic = T.advection_sinsin_problem(n), T.advection_sinsin_ic()
        nt = T.advection_sinsin_steps(n)
        with ccfv.Context(p) as c:
            ex = c.exact(ic, 1.0)
            ms = []
            for r in range(reps + 1 if n == 128 else 1):
                c.fill_exact(ic, 0.0)
                c.sync()
                t = time.perf_counter()
                c.plain_solve(0.0, 1.0, nt)
                c.sync()
                if r >= 1 or n != 128:
                    ms.append((time.perf_counter() - t) * 1e3)
            u = c.get_state()
            row = {"n": n, "steps": nt, "rel_l2": rel_l2(u, ex),
                   "rms": float(np.sqrt(np.mean((u - ex) ** 2)))}
            if n == 128:
                row["p50_ms"] = pct(ms, 50)
                row["cell_updates_per_s"] = n * n * nt / (pct(ms, 50) * 1e-3)
            rows.append(row)
            if n <= 128:
                c.fill_exact(ic, 0.0)
                c.plain_solve(0.0, 1.0, 2560)
                fixed.append({"n": n, "steps": 2560,
                              "rms": float(np.sqrt(np.mean((c.get_state() - ex) ** 2)))})
    order = lambda e: [float(np.log2(e[i]["rms"] / e[i + 1]["rms"])) for i in range(len(e) - 1)]
    r128 = rows[2]
    return {
        "metric": "FP64 cell-updates/s and observed order, plain method at CFL 0.8",
        "value": r128["cell_updates_per_s"], "unit": "cell-updates/s", "n_gpus": 1,
        "steps": reps, "higher_is_better": True, "dtype": "f64",
        "data": "synthetic (sin(2 pi x) sin(2 pi y) cell averages)",
        "config": {"workload": "sinsin-128-plain", "velocity": [1.0, 0.5], "nu": 0.01,
                   "cfl": 0.8, "T": 1.0, "bc": "periodic", "grids": [32, 64, 128, 256]},
        "errors_vs_exact": rows, "orders_cfl_0.8": order(rows),
        "errors_fixed_step": fixed, "orders_fixed_step": order(fixed),
    }


def smooth(a, reps=None):
    """Where the reference's Parareal converges: the headline's grid, time
    interval, fine steps, slices and tolerance (512^2, T=1 in 1024 fine
    steps, 64 slices of 16, kmax 8, tol 1e-10), but a translating flow
    (u, v) = (0.93, 0.47), nu = 1e-3, and a Gaussian of standard deviation
    0.1 at (0.5, 0.5).  Plain sweep (replayed from a hipGraph) against
    batched Parareal, each against the serial fine solution."""
    import ccfv
    from ccfv import types as T
    p, ic = T.constant_gaussians(a.n, a.n, 0.93, 0.47, 1e-3, [(0.5, 0.5, 0.1, 1.0)])
    ctx = ccfv.Context(p)
    ctx.fill_exact(ic, 0.0)
    u0 = ctx.get_state()
    reps = reps or a.steps

    def reset():
        ctx.set_state(u0)

    plain = lambda: ctx.plain_solve(0.0, 1.0, a.nt)  # noqa: E731
    if not a.no_graphs:
        ctx.set_option(T.OPT_PLAIN_GRAPHS, 1)
    timed(ctx, reset, plain, 3)
    plain_ms = timed(ctx, reset, plain, reps)
    ctx.set_option(T.OPT_PLAIN_GRAPHS, 0)
    reset()
    plain()
    uref = ctx.get_state()
    box = {}

    def para():
        box["r"] = ctx.parareal_solve(0.0, 1.0, a.nt, a.slices, a.kmax, a.tol)

    timed(ctx, reset, para, 3)
    para_ms = timed(ctx, reset, para, reps)
    rep = box["r"]
    rel = rel_l2(ctx.get_state(), uref)
    # the same solve with fewer, longer slices
    by_slices = {}
    for ns in (32, 16):
        if a.nt % ns:
            continue
        r = {}
        f = lambda: r.update(rep=ctx.parareal_solve(0.0, 1.0, a.nt, ns, a.kmax, a.tol))  # noqa: E731
        timed(ctx, reset, f, 2)
        ms = timed(ctx, reset, f, reps)
        e = rel_l2(ctx.get_state(), uref)
        by_slices[str(ns)] = {"p50_ms": pct(ms, 50), "speedup_vs_plain": pct(plain_ms, 50) / pct(ms, 50),
                              "iterations": r["rep"].iterations,
                              "accepted": bool(r["rep"].converged) and e <= a.tol,
                              "rel_l2_vs_serial_fine": e}
    ctx.close()
    cells = a.n * a.n
    pa, pp = pct(para_ms, 50), pct(plain_ms, 50)
    accepted = bool(rep.converged) and rel <= a.tol
    return {
        "metric": "FP64 cell-updates/s, batched Parareal vs plain sweep where Parareal "
                  "converges",
        "value": cells * a.nt / (min(pa, pp) * 1e-3) if accepted else cells * a.nt / (pp * 1e-3),
        "unit": "cell-updates/s", "n_gpus": 1, "steps": reps, "higher_is_better": True,
        "dtype": "f64", "data": "synthetic (Gaussian cell averages)",
        "config": {"workload": f"translation-{a.n}-parareal", "grid": [a.n, a.n], "T": 1.0,
                   "fine_steps": a.nt, "slices": a.slices, "J": a.nt // a.slices,
                   "kmax": a.kmax, "tol": a.tol, "velocity": [0.93, 0.47], "nu": 1e-3,
                   "bc": "periodic"},
        "plain_graph_p50_ms": pp, "plain_graph_p99_ms": pct(plain_ms, 99),
        "parareal_p50_ms": pa, "parareal_p99_ms": pct(para_ms, 99),
        "speedup_vs_plain": pp / pa,
        "parareal_iterations": rep.iterations, "parareal_converged": bool(rep.converged),
        "parareal_accepted": accepted,
        "parareal_increments": [float(x) for x in rep.increments],
        "rel_l2_vs_serial_fine": rel,
        "fine_field_steps": rep.fine_steps_run, "coarse_field_steps": rep.coarse_steps_run,
        "parareal_by_slices": by_slices,
    }


def probe(a):
    """Short run for counter collection (rocprofv3 --pmc): one calibration
    copy between two 2048^2 fields that are not cache resident, 64 plain
    steps of the headline workload and one Parareal iteration."""
    import ccfv
    from ccfv import types as T
    p, ic = T.rotation_problem(a.n), T.rotation_ic()
    big = T.rotation_problem(2048)
    c1, c2 = ccfv.Context(big), ccfv.Context(big)
    c1.fill_exact(ic, 0.0)
    junk = ccfv.Context(big)
    junk.fill_exact(ic, 0.0)  # evict c1's field from the caches
    c1.get_state_device(c2.state_ptr())
    c1.sync()
    ctx = ccfv.Context(p)
    ctx.fill_exact(ic, 0.0)
    ctx.plain_solve(0.0, 64.0 / a.nt, 64)
    ctx.fill_exact(ic, 0.0)
    r = ctx.parareal_solve(0.0, 1.0, a.nt, a.slices, 1, a.tol)
    return {"probe": "ok", "iterations": r.iterations}


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU "
                          "sharding of the slice chain is designed but not "
                          "built", "n_gpus": a.gpus}))
        sys.exit(2)
    if a.workload == "ragged":
        print(json.dumps(ragged(a)))
        return
    if a.workload == "serve":
        print(json.dumps(serve(a)))
        return
    if a.workload == "swirl":
        print(json.dumps(swirl(a)))
        return
    if a.workload == "translation":
        print(json.dumps(smooth(a)))
        return
    if a.workload == "copy":
        print(json.dumps(copy_rate(a)))
        return
    if a.workload == "accuracy":
        print(json.dumps(accuracy(a)))
        return
    if a.probe:
        print(json.dumps(probe(a)))
        return
    from ccfv import types as T
    line = rotation(a)
    if not a.no_legs:
        legs = {}
        t = time.perf_counter()
        legs["sinsin-128-accuracy"] = accuracy(a, reps=min(a.steps, 10))
        log(f"leg accuracy done ({time.perf_counter() - t:.1f} s)")
        t = time.perf_counter()
        legs["translation-512"] = smooth(a, reps=min(a.steps, 10))
        log(f"leg translation done ({time.perf_counter() - t:.1f} s)")
        t = time.perf_counter()
        legs["swirl-2048"] = swirl(a, reps=1)
        log(f"leg swirl done ({time.perf_counter() - t:.1f} s)")
        t = time.perf_counter()
        legs["ragged-256"] = ragged(a, reps=1, full=False)
        log(f"leg ragged done ({time.perf_counter() - t:.1f} s)")
        t = time.perf_counter()
        legs["serve"] = {"as requested": serve_stream(a, a.leg_rate, 0),
                         "adaptive": serve_stream(a, a.leg_rate, a.adaptive_above),
                         "serial": serve_stream(a, a.leg_rate, 0, serial=True)}
        log(f"leg serve done ({time.perf_counter() - t:.1f} s)")
        line["legs"] = legs
    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(T.rotation_problem(a.n), T.rotation_ic(), a.nt, a.cpu_seconds)
    line["cpu_baseline"] = cpu
    print(json.dumps(line))


if __name__ == "__main__":
    main()
