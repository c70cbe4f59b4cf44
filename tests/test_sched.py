"""The serving engine's admission bookkeeping (csrc/ccfv_sched.h: first-fit
arena allocator, arrival-order admission with a bounded bypass, request
footprint), unit-tested in C++ on the CPU: fixed cases, then seeded random
operation sequences against a unit-by-unit model of the arena and against
the admission contract (each check was seen to fail on a planted mistake
in a copy of the header: an off-by-one bypass bound, no merge with the
left neighbour, first fit skipping exact fits)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "parareal-cc-cfv-split-gpu_amd", "csrc")


def test_sched_unit(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path / "sched_test")
    subprocess.run([gxx, "-std=c++17", "-O1", "-Wall", "-Werror", "-I", CSRC,
                    os.path.join(ROOT, "tests", "sched_test.cpp"), "-o", exe],
                   check=True, capture_output=True, text=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip() == "ok"
