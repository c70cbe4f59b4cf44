"""GPU: the reference-compatible command-line driver end to end: output
files (T2S4_Parareal.cu:93-114), serial-fine snapshots every T_span steps
(:1234-1243), Parareal lines (:1346) and closing lines (:1359-1366), with
the values checked against the library."""
import os
import re
import subprocess

import numpy as np
import pytest

from ccfv import types as T

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "parareal-cc-cfv-split-gpu_amd", "lib", "ccfv_t2s4")
LINE = re.compile(r"^-?\d+\.\d{6} -?\d+\.\d{6} -?\d\.\d{4}e[+-]\d{2,3}$")


def read_field(path, n):
    rows = [l for l in open(path).read().split("\n")]
    # n blocks of n lines, each followed by a blank line
    assert len(rows) == n * (n + 1) + 1 and rows[-1] == ""
    vals = []
    for i in range(n):
        block = rows[i * (n + 1): (i + 1) * (n + 1)]
        assert block[-1] == ""
        for j, l in enumerate(block[:-1]):
            assert LINE.match(l), l
            x, y, c = l.split()
            assert abs(float(x) - (-1 + 2 / n * (i + 0.5))) < 1e-6
            assert abs(float(y) - (-1 + 2 / n * (j + 0.5))) < 1e-6
            vals.append(float(c))
    return np.array(vals)


@pytest.mark.parametrize("t0", [0.0, 0.1])
def test_cli_run_matches_the_library(gpu, tmp_path, t0):
    """t0 = T_min: the initial data are the exact solution at t = 0 either
    way (T2S4_Parareal.cu:1147); the steps start at T_min."""
    n, nt, span, K, t1 = 32, 80, 20, 0.01, 0.5
    (tmp_path / "config.txt").write_text(
        f"T_min:{t0}\nT_max:{t1}\nN:{n}\nNt:{nt}\nT_span:{span}\nK:{K}\n")
    r = subprocess.run([CLI, "config.txt"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    out = r.stdout.splitlines()
    assert out[0] == "open config file ok"
    iters = [l for l in out if l.startswith("Parareal iter")]
    assert len(iters) == 8
    assert out[-8:-1] == ["K:0.01", "dx:0.0625000000000000", "dy:0.0625000000000000",
                          f"dt:{(t1 - t0) / nt:.16f}", "Nx:32", "Ny:32", "Nt:80"]
    assert out[-1].startswith("Total time (including Parareal): ")
    files = sorted(p.name for p in tmp_path.glob("data_T2S4_*.txt"))
    assert files == [f"data_T2S4_{s:06d}.txt" for s in range(0, nt + 1, span)]
    p, ic = T.reference_problem(n, K), T.reference_ic()
    with gpu.Context(p) as c:
        c.fill_exact(ic, 0.0)
        u0 = c.get_state()
        np.testing.assert_allclose(read_field(tmp_path / "data_T2S4_000000.txt", n), u0,
                                   rtol=6e-5, atol=1e-300)
        snaps = {}
        for s in range(span, nt + 1, span):
            c.advance(t0, (t1 - t0) / nt, s - span, s)
            snaps[s] = c.get_state()
        for s, u in snaps.items():
            np.testing.assert_allclose(read_field(tmp_path / f"data_T2S4_{s:06d}.txt", n),
                                       u, rtol=6e-5, atol=1e-300)
        uref = snaps[nt]
        # the k-th line is the k-th iterate of one solve, which is also the
        # end state of a solve stopped after k iterations
        c.set_option(T.OPT_KEEP_ITERATES, 1)
        c.set_state(u0)
        c.parareal_solve(t0, t1, nt, 40, 8, 0.0)
        kept = {}
        for k, l in enumerate(iters, start=1):
            m = re.match(r"^Parareal iter +(\d+): err\(T\) = (\S+)$", l)
            assert m and int(m.group(1)) == k
            kept[k] = c.parareal_iterate(k)
            err = float(np.abs(kept[k] - uref).max())
            assert m.group(2) == f"{err:.3e}", (k, l, err)
        assert np.array_equal(c.get_state(), kept[8])
        c.set_option(T.OPT_KEEP_ITERATES, 0)
        for k in (1, 3):
            c.set_state(u0)
            c.parareal_solve(t0, t1, nt, 40, k, 0.0)
            assert np.array_equal(c.get_state(), kept[k]), k


def test_advance_is_the_plain_loop(gpu):
    """ccfv_advance in pieces equals ccfv_plain_solve (same step times)."""
    p, ic = T.rotation_problem(64), T.rotation_ic()
    with gpu.Context(p) as c:
        c.fill_exact(ic, 0.0)
        c.plain_solve(0.0, 0.3, 30)
        a = c.get_state()
        c.fill_exact(ic, 0.0)
        for n0 in range(0, 30, 7):
            c.advance(0.0, 0.3 / 30, n0, min(30, n0 + 7))
        assert np.array_equal(c.get_state(), a)
