"""GPU: the pipelined Parareal schedule with a tolerance defers the last
slice group of iteration k+1's fine propagators until the host has read
iteration k's increment (a converged solve never runs it).  Every result
must be bitwise the single-stream schedule's: a solve that converges
mid-run (with kept iterates), one that converges exactly at kmax, one that
never converges, one that ends with the serial finish, and repeated solves
on one context."""
import numpy as np
import pytest

from ccfv import types as T

pytestmark = pytest.mark.gpu


def translation(n=128):
    return T.constant_gaussians(n, n, 0.93, 0.47, 1e-3, [(0.5, 0.5, 0.1, 1.0)])


def solve(gpu, p, ic, pipeline, ns, nt, kmax, tol, keep=False, finish=False, group=32):
    with gpu.Context(p) as c:
        c.fill_exact(ic, 0.0)
        c.set_option(T.OPT_PIPELINE, pipeline)
        c.set_option(T.OPT_PIPE_GROUP, group)
        c.set_option(T.OPT_KEEP_ITERATES, 1 if keep else 0)
        c.set_option(T.OPT_SERIAL_FINISH, 1 if finish else 0)
        rep = c.parareal_solve(0.0, 1.0, nt, ns, kmax, tol)
        its = [c.parareal_iterate(k + 1) for k in range(rep.iterations)] if keep else []
        return c.get_state(), rep, its


@pytest.mark.parametrize("ns,group", [(16, 32), (16, 5), (32, 8)])
def test_converged_solve_matches_single_stream(gpu, ns, group):
    p, ic = translation()
    us, rs, its_s = solve(gpu, p, ic, 0, ns, 256, 8, 1e-10, keep=True, group=group)
    up, rp, its_p = solve(gpu, p, ic, 1, ns, 256, 8, 1e-10, keep=True, group=group)
    assert rs.converged == 1 and 1 < rs.iterations < 8
    assert rp.iterations == rs.iterations and rp.converged == rs.converged
    assert list(rp.increments) == list(rs.increments)
    assert np.array_equal(up, us)
    for a, b in zip(its_p, its_s):
        assert np.array_equal(a, b)


def test_converges_exactly_at_kmax(gpu):
    """kmax equal to the iterations needed: the decision on the last
    iteration is read after the loop."""
    p, ic = translation()
    _, rs, _ = solve(gpu, p, ic, 0, 16, 256, 8, 1e-10)
    need = rs.iterations
    us, rs2, _ = solve(gpu, p, ic, 0, 16, 256, need, 1e-10)
    up, rp, _ = solve(gpu, p, ic, 1, 16, 256, need, 1e-10)
    assert rs2.converged == 1 and rp.converged == 1
    assert rp.iterations == need
    assert np.array_equal(up, us)


def test_unconverged_and_serial_finish(gpu):
    """Rotation (no convergence within kmax): every speculative iteration is
    a real one; with the serial finish the result is the plain solve."""
    p, ic = T.rotation_problem(64), T.rotation_ic()
    us, rs, _ = solve(gpu, p, ic, 0, 8, 64, 4, 1e-14)
    up, rp, _ = solve(gpu, p, ic, 1, 8, 64, 4, 1e-14)
    assert rs.converged == 0 and rp.converged == 0 and rp.iterations == 4
    assert np.array_equal(up, us)
    assert list(rp.increments) == list(rs.increments)
    uf, rf, _ = solve(gpu, p, ic, 1, 8, 64, 4, 1e-14, finish=True)
    with gpu.Context(p) as c:
        c.fill_exact(ic, 0.0)
        c.plain_solve(0.0, 1.0, 64)
        plain = c.get_state()
    assert rf.serial_finish == 1
    assert np.array_equal(uf, plain)


def test_repeated_solves_on_one_context(gpu):
    """A converged solve, then the same solve again on the same context:
    the deferred group's events of the first solve must not order the
    second."""
    p, ic = translation()
    with gpu.Context(p) as c:
        c.fill_exact(ic, 0.0)
        r1 = c.parareal_solve(0.0, 1.0, 256, 16, 8, 1e-10)
        assert r1.converged == 1
        c.fill_exact(ic, 0.0)
        r2 = c.parareal_solve(0.0, 1.0, 256, 16, 8, 1e-10)
        u2 = c.get_state()
    us, rs, _ = solve(gpu, p, ic, 0, 16, 256, 8, 1e-10)
    assert r2.iterations == rs.iterations
    assert np.array_equal(u2, us)
