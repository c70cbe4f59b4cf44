"""CPU: the benchmark's command-line contract that needs no GPU — more than
one GPU is refused with a JSON line (this build measures one GPU), and the
help lists every workload."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                          capture_output=True, text=True, timeout=120)


def test_more_than_one_gpu_is_refused():
    r = run("--gpus", "2")
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]


def test_help_lists_every_workload():
    r = run("--help")
    assert r.returncode == 0
    for w in ("rotation", "accuracy", "translation", "ragged", "serve", "swirl", "copy"):
        assert w in r.stdout, w
