"""The reference-compatible command-line driver (cli/ccfv_t2s4.cpp): argument
and config handling on the CPU (T2S4_Parareal.cu:1009-1055)."""
import os
import subprocess

import pytest

import ccfv

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "parareal-cc-cfv-split-gpu_amd", "lib", "ccfv_t2s4")

pytestmark = pytest.mark.skipif(not os.path.exists(CLI), reason="ccfv_t2s4 not built")


def run(args, cwd):
    return subprocess.run([CLI] + args, cwd=cwd, capture_output=True, text=True,
                          timeout=120)


def test_no_config_argument(tmp_path):
    r = run([], tmp_path)
    assert r.returncode == 1 and r.stdout == "please input config file name\n"


def test_config_cannot_be_opened(tmp_path):
    r = run([str(tmp_path / "missing.txt")], tmp_path)
    assert r.returncode == 1 and r.stdout == "Cannot open config file!\n"


def test_missing_key_is_an_error(tmp_path):
    cfg = tmp_path / "c.txt"
    cfg.write_text("T_min:0\nT_max:1\nN:32\nNt:40\nK:0.01\n")  # no T_span
    r = run([str(cfg)], tmp_path)
    assert r.returncode == 1
    assert r.stdout == "open config file ok\n"
    assert "T_span" in r.stderr


def test_unknown_argument(tmp_path):
    cfg = tmp_path / "c.txt"
    cfg.write_text("T_min:0\nT_max:1\nN:32\nNt:40\nT_span:20\nK:0.01\n")
    r = run([str(cfg), "--bogus"], tmp_path)
    assert r.returncode == 1 and "--bogus" in r.stderr


def test_without_device_fails_loudly(tmp_path):
    if ccfv.device_count() > 0:
        pytest.skip("a device is present")
    cfg = tmp_path / "c.txt"
    cfg.write_text("T_min:0\nT_max:1\nN:32\nNt:40\nT_span:20\nK:0.01\n")
    r = run([str(cfg)], tmp_path)
    assert r.returncode == 2 and "no HIP device" in r.stderr
    assert not list(tmp_path.glob("data_T2S4_*.txt"))
