"""The swirl feet from line constants (ccfv_feet.h: face_foot_swirl) are
bitwise the general feet (face_foot), and the table form the general sweeps
use (face_foot_swirl_tab, with sin_pi_shift) agrees with them to a few ulp;
checked on the host."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "parareal-cc-cfv-split-gpu_amd", "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def test_swirl_feet_bitwise(tmp_path):
    if not os.path.exists(HIPCC) and shutil.which("hipcc") is None:
        pytest.skip("hipcc not available")
    exe = str(tmp_path / "feet_test")
    # host code only: the device pass is not needed for this check
    r = subprocess.run([HIPCC, "-O3", "-std=c++17", "-x", "hip", "--offload-arch=gfx950",
                        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                        "-I", CSRC, os.path.join(ROOT, "tests", "feet_test.cpp"), "-o", exe],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    out = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout
    assert out.stdout.startswith("ok"), out.stdout
    print(out.stdout)
