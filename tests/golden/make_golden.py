"""Writes the golden fixtures of tests/golden/ from the CPU oracle.

The reference ships no test vectors, so these are the oracle's own outputs on
small inputs, committed so that (1) any change to the oracle's results shows
up as a CPU test failure (test_golden.py, bitwise), and (2) the HIP path is
checked against fixed data as well as against the live oracle.

Each case stores its input state `u0` (the oracle's cell averages of the
initial data at t0) and its output `u`, both float64, cell (i, j) at index
i*ny + j with j contiguous (the layout of T2S4_Parareal.cu:176-180), in
`<name>.npz` (no pickles).  `cases.json`
holds the problem, the initial data and the operation for every case.

Regenerate (only when the oracle's arithmetic is meant to change):
    make -C oracle && python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (os.path.join(ROOT, "parareal-cc-cfv-split-gpu_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

from ccfv import types as T  # noqa: E402


def problem_dict(p: T.Problem) -> dict:
    return {"nx": p.nx, "ny": p.ny, "bc": p.bc, "vel_kind": p.vel_kind,
            "box": [p.x_lo, p.x_hi, p.y_lo, p.y_hi], "nu": p.nu,
            "vel": list(p.vel), "flags": p.flags, "src_kind": p.src_kind,
            "src": list(p.src)}


def problem_from(d: dict) -> T.Problem:
    p = T.make_problem(d["nx"], d["ny"], bc=d["bc"], vel_kind=d["vel_kind"],
                       vel=d["vel"], nu=d["nu"], box=tuple(d["box"]), flags=d["flags"])
    p.src_kind = d["src_kind"]
    for i, v in enumerate(d["src"]):
        p.src[i] = v
    return p


def ic_dict(ic: T.IC) -> dict:
    return {"kind": ic.kind, "n": ic.n, "p": list(ic.p)}


def ic_from(d: dict) -> T.IC:
    return T.make_ic(d["kind"], d["p"], n=d["n"])


def cases():
    """(name, problem, ic, op) with op a dict naming the solve."""
    out = []
    out.append(("reference32_plain", T.reference_problem(32, 0.01), T.reference_ic(),
                {"op": "plain", "t0": 0.0, "t1": 0.5, "nt": 128}))
    out.append(("reference32_step", T.reference_problem(32, 0.001), T.reference_ic(),
                {"op": "step", "t0": 0.25, "h": 1.0 / 16}))
    n = 32
    out.append(("sinsin32_plain", T.advection_sinsin_problem(n), T.advection_sinsin_ic(),
                {"op": "plain", "t0": 0.0, "t1": 1.0, "nt": T.advection_sinsin_steps(n)}))
    out.append(("rotation64_plain", T.rotation_problem(64), T.rotation_ic(),
                {"op": "plain", "t0": 0.0, "t1": 0.5, "nt": 64}))
    out.append(("rotation64_parareal", T.rotation_problem(64), T.rotation_ic(),
                {"op": "parareal", "t0": 0.0, "t1": 0.5, "nt": 64, "slices": 8,
                 "kmax": 3, "tol": 0.0}))
    out.append(("reference32_parareal", T.reference_problem(32, 0.01), T.reference_ic(),
                {"op": "parareal", "t0": 0.0, "t1": 0.5, "nt": 64, "slices": 8,
                 "kmax": 3, "tol": 0.0}))
    out.append(("swirl32_plain", T.swirl_problem(32, T=1.0), T.cosine_bell_ic(),
                {"op": "plain", "t0": 0.0, "t1": 0.5, "nt": 64}))
    p, ic = T.constant_gaussians(32, 64, -0.7, 0.4, 3e-3,
                                 [(0.3, 0.6, 0.05, 1.0), (0.7, 0.2, 0.08, 0.6)])
    out.append(("gauss32x64_plain", p, ic,
                {"op": "plain", "t0": 0.0, "t1": 0.5, "nt": 64}))
    p, ic, _ = T.manufactured_sinsin(32)
    out.append(("source32_plain", p, ic,
                {"op": "plain", "t0": 0.0, "t1": 0.5, "nt": 40}))
    return out


def run_oracle(oracle, p, u0, op):
    """Returns (u, extra) with extra the Parareal report, if any."""
    if op["op"] == "plain":
        return oracle.plain(p, u0, op["t0"], op["t1"], op["nt"]), {}
    if op["op"] == "step":
        return oracle.step(p, u0, op["t0"], op["h"]), {}
    if op["op"] == "parareal":
        r = oracle.parareal(p, u0, op["t0"], op["t1"], op["nt"], op["slices"],
                            op["kmax"], op["tol"])
        return r["u"], {"iterations": r["iterations"], "increments": r["increments"]}
    raise ValueError(op["op"])


def main():
    import pyoracle as oracle
    oracle.set_threads(min(8, os.cpu_count() or 1))
    manifest = {"layout": "index i*ny + j (j contiguous), float64",
                "generator": "tests/golden/make_golden.py (CPU oracle)",
                "cases": []}
    for name, p, ic, op in cases():
        u0 = oracle.exact(p, ic, op["t0"])
        u, extra = run_oracle(oracle, p, u0, op)
        assert np.all(np.isfinite(u)), name
        np.savez_compressed(os.path.join(HERE, name + ".npz"), u0=u0, u=u)
        manifest["cases"].append({"name": name, "problem": problem_dict(p),
                                  "ic": ic_dict(ic), "op": op, **extra})
        print(f"{name}: {p.nx}x{p.ny} {op['op']} |u|={float(np.linalg.norm(u)):.6g}")
    with open(os.path.join(HERE, "cases.json"), "w") as f:
        json.dump(manifest, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
