"""Writes the full-size Parareal record of the 512^2 rotation workload from the
CPU oracle: tests/golden/rotation512_parareal.json and
tests/golden/rotation512_parareal.npz.

Workload: periodic [0,1]^2, 512x512 cells, solid-body rotation
u = -2 pi (y - 1/2), v = 2 pi (x - 1/2), nu = 1e-3, Gaussian of standard
deviation 0.05 at (0.5, 0.75), T = 1 in 1024 fine steps; Parareal over 64
slices (coarse step = 16 fine steps), 8 iterations, tol = 0 (every iteration
runs, as T2S4_Parareal.cu:1296 does).  The oracle runs the literal
restatement of the reference's loop (T2S4_Parareal.cu:1250-1355): every
slice's fine and coarse propagators in every iteration, no caching.

Stored: the eight relative L2 increments over all slice end-points, the
relative L2 and max-norm distance of U^8(T) to the serial fine solution, the
L2 norms of both, and the sums of both over 16x16 blocks of cells (32x32
values each; float64, block (bi, bj) at index bi*32 + bj).  The full fields
(2 MiB each) are not committed; the block sums pin the end state's shape.

Takes about two minutes on 8 host cores:
    make -C oracle && python tests/golden/make_rotation512.py
"""
from __future__ import annotations

import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (os.path.join(ROOT, "parareal-cc-cfv-split-gpu_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

from ccfv import types as T  # noqa: E402

N, NT, SLICES, KMAX = 512, 1024, 64, 8
BLOCK = 16


def block_sums(u: np.ndarray, n: int = N, b: int = BLOCK) -> np.ndarray:
    return u.reshape(n // b, b, n // b, b).sum(axis=(1, 3)).ravel()


def main():
    import pyoracle as oracle
    oracle.set_threads(os.cpu_count() or 1)
    p, ic = T.rotation_problem(N), T.rotation_ic()
    u0 = oracle.exact(p, ic, 0.0)
    t = time.time()
    fine = oracle.plain(p, u0, 0.0, 1.0, NT)
    t_fine = time.time() - t
    t = time.time()
    r = oracle.parareal(p, u0, 0.0, 1.0, NT, SLICES, KMAX, 0.0)
    t_para = time.time() - t
    u = r["u"]
    rec = {
        "workload": "rotation-512-parareal",
        "generator": "tests/golden/make_rotation512.py (CPU oracle, literal loop)",
        "n": N, "nt_fine": NT, "slices": SLICES, "kmax": KMAX, "tol": 0.0,
        "iterations": r["iterations"],
        "increments": [float(x) for x in r["increments"]],
        "rel_l2_vs_serial_fine": float(np.linalg.norm(u - fine) / np.linalg.norm(fine)),
        "max_abs_vs_serial_fine": float(np.abs(u - fine).max()),
        "norm_parareal": float(np.linalg.norm(u)),
        "norm_serial_fine": float(np.linalg.norm(fine)),
        "block": BLOCK,
        "oracle_seconds": {"serial_fine": round(t_fine, 1), "parareal": round(t_para, 1),
                           "threads": os.cpu_count() or 1},
    }
    np.savez_compressed(os.path.join(HERE, "rotation512_parareal.npz"),
                        parareal_blocks=block_sums(u), serial_fine_blocks=block_sums(fine))
    with open(os.path.join(HERE, "rotation512_parareal.json"), "w") as f:
        json.dump(rec, f, indent=1)
        f.write("\n")
    print(json.dumps(rec))


if __name__ == "__main__":
    main()
