"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer: the CPU
oracle over every problem kind (tests/oracle_asan.c) and the serving engine's
admission bookkeeping (tests/sched_test.cpp).  Any report aborts the run."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "parareal-cc-cfv-split-gpu_amd", "csrc")
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all",
       "-fno-omit-frame-pointer", "-g", "-O1"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")


def _run(cmd, exe):
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=ENV)
    assert out.returncode == 0, (out.stdout + out.stderr)[-4000:]
    return out.stdout


def test_oracle_sanitized(tmp_path):
    cc = shutil.which("gcc")
    if cc is None:
        pytest.skip("gcc not available")
    exe = str(tmp_path / "oracle_asan")
    out = _run([cc, "-std=c11", "-ffp-contract=off", "-Wall", "-Wno-unknown-pragmas", *SAN,
                os.path.join(ROOT, "tests", "oracle_asan.c"),
                os.path.join(ROOT, "oracle", "ccfv_oracle.c"), "-lm", "-o", exe], exe)
    assert out.strip().endswith("ok"), out


def test_sched_sanitized(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path / "sched_asan")
    out = _run([gxx, "-std=c++17", "-Wall", *SAN, "-I", CSRC,
                os.path.join(ROOT, "tests", "sched_test.cpp"), "-o", exe], exe)
    assert out.strip() == "ok", out
