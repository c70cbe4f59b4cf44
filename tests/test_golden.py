"""Golden fixtures (tests/golden/): small inputs and the oracle's outputs,
written by tests/golden/make_golden.py and committed.

CPU: the oracle reproduces every fixture bitwise (the per-cell arithmetic is
independent of the thread count; Parareal increments are norms summed in
thread order, so they are compared to 1e-12 relative).  GPU: the HIP path,
through the C ABI, reproduces the fixtures' outputs from their stored inputs
to relative L2 <= 1e-13 for one step and <= 1e-12 for whole solves (the
bound stated for the serial fine solution), and the Parareal increments to
1e-9 relative."""
import json
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
sys.path.insert(0, GOLDEN)
import make_golden  # noqa: E402

with open(os.path.join(GOLDEN, "cases.json")) as _f:
    MANIFEST = json.load(_f)
CASES = MANIFEST["cases"]


def load_case(c):
    with np.load(os.path.join(GOLDEN, c["name"] + ".npz")) as z:
        u0, u = z["u0"].copy(), z["u"].copy()
    return (make_golden.problem_from(c["problem"]), make_golden.ic_from(c["ic"]), u0, u)


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_manifest_matches_generator():
    """The committed manifest lists exactly the generator's cases, with the
    same problems (a regenerated fixture set would differ otherwise)."""
    gen = make_golden.cases()
    assert [c["name"] for c in CASES] == [g[0] for g in gen]
    for c, (name, p, ic, op) in zip(CASES, gen):
        assert c["problem"] == make_golden.problem_dict(p), name
        assert c["ic"] == make_golden.ic_dict(ic), name
        assert c["op"] == op, name


@pytest.mark.parametrize("case", CASES, ids=lambda c: c["name"])
def test_oracle_reproduces_fixture(oracle, case):
    p, ic, u0, u = load_case(case)
    assert u0.shape == u.shape == (p.nx * p.ny,)
    assert np.array_equal(oracle.exact(p, ic, case["op"]["t0"]), u0)
    got, extra = make_golden.run_oracle(oracle, p, u0, case["op"])
    assert np.array_equal(got, u), case["name"]
    if case["op"]["op"] == "parareal":
        assert extra["iterations"] == case["iterations"]
        np.testing.assert_allclose(extra["increments"], case["increments"], rtol=1e-12)


def test_oracle_fixture_independent_of_threads(oracle):
    """One fixture again on 1 thread: the lines are swept independently."""
    case = next(c for c in CASES if c["name"] == "gauss32x64_plain")
    p, ic, u0, u = load_case(case)
    oracle.set_threads(1)
    try:
        got, _ = make_golden.run_oracle(oracle, p, u0, case["op"])
    finally:
        oracle.set_threads(min(16, os.cpu_count() or 1))
    assert np.array_equal(got, u)


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=lambda c: c["name"])
def test_gpu_matches_fixture(gpu, case):
    p, ic, u0, u = load_case(case)
    op = case["op"]
    print(f"golden case {case['name']} {p.nx}x{p.ny} {op['op']}", flush=True)
    with gpu.Context(p) as c:
        c.set_state(u0)
        if op["op"] == "plain":
            c.plain_solve(op["t0"], op["t1"], op["nt"])
            tol = 1e-12
        elif op["op"] == "step":
            c.step(op["t0"], op["h"])
            tol = 1e-13
        else:
            rep = c.parareal_solve(op["t0"], op["t1"], op["nt"], op["slices"], op["kmax"],
                                   op["tol"])
            tol = 1e-12
            assert rep.iterations == case["iterations"]
            for a, b in zip(rep.increments, case["increments"]):
                assert abs(a - b) <= 1e-9 * max(abs(b), 1e-30) + 1e-15
        g = c.get_state()
    assert rel_l2(g, u) <= tol, case["name"]
