"""GPU: the pipelined Parareal schedule (fine propagators of iteration k+1
overlapping the coarse chain of iteration k on a second stream) against the
single-stream schedule.  Same kernels on the same inputs: bitwise equal."""
import numpy as np
import pytest

from ccfv import types as T

pytestmark = pytest.mark.gpu


def solve(gpu, p, ic, pipeline, group, kmax, tol, ns=8, nt=64, t1=0.5, graphs=1):
    with gpu.Context(p) as c:
        c.fill_exact(ic, 0.0)
        c.set_option(T.OPT_PIPELINE, pipeline)
        c.set_option(T.OPT_PIPE_GROUP, group)
        c.set_option(T.OPT_GRAPHS, graphs)
        rep = c.parareal_solve(0.0, t1, nt, ns, kmax, tol)
        return c.get_state(), rep


@pytest.mark.parametrize("make", [
    lambda: (T.rotation_problem(64), T.rotation_ic()),
    lambda: (T.reference_problem(32, 0.01), T.reference_ic()),
    lambda: (T.swirl_problem(32, T=1.0), T.cosine_bell_ic()),
])
@pytest.mark.parametrize("group", [1, 3, 8])
@pytest.mark.parametrize("mode,graphs", [(1, 0), (1, 1), (2, 0)])
def test_pipelined_equals_sequential(gpu, make, group, mode, graphs):
    """mode 1: fine propagators of k+1 beside the chain of k; mode 2: also
    the chains of consecutive iterations beside each other (one speculative
    chain past convergence)."""
    p, ic = make()
    for kmax, tol in ((4, 0.0), (8, 1e-9), (8, 1e-3)):
        us, rs = solve(gpu, p, ic, 0, group, kmax, tol)
        up, rp = solve(gpu, p, ic, mode, group, kmax, tol, graphs=graphs)
        assert np.array_equal(us, up)
        assert rs.iterations == rp.iterations
        assert rs.converged == rp.converged
        np.testing.assert_allclose(rs.increments, rp.increments, rtol=1e-12)


def test_graph_replays_match_eager(gpu, oracle):
    """Captured phases replayed several times, with eager work and host
    synchronisation in between, against the eager (uncaptured) solve."""
    p, ic = T.rotation_problem(64), T.rotation_ic()
    ref, rref = solve(gpu, p, ic, 1, 3, 4, 0.0, graphs=0)
    with gpu.Context(p) as c:
        c.set_option(T.OPT_PIPE_GROUP, 3)
        c.set_option(T.OPT_GRAPHS, 1)
        for rep_i in range(4):
            c.fill_exact(ic, 0.0)
            c.step(0.0, 1e-3)          # eager work between replays
            c.fill_exact(ic, 0.0)
            rep = c.parareal_solve(0.0, 0.5, 64, 8, 4, 0.0)
            u = c.get_state()
            assert np.array_equal(u, ref), rep_i
            np.testing.assert_allclose(rep.increments, rref.increments, rtol=1e-12)


def test_early_stop_reports_convergence(gpu):
    # G == F (one fine step per slice): the first correction changes nothing
    p, ic = T.rotation_problem(64), T.rotation_ic()
    u, rep = solve(gpu, p, ic, 1, 2, 8, 1e-12, ns=8, nt=8)
    assert rep.converged == 1
    assert rep.iterations == 1
    assert rep.increments[0] == 0.0


def test_not_converged_status(gpu):
    p, ic = T.rotation_problem(64), T.rotation_ic()
    with gpu.Context(p) as c:
        c.fill_exact(ic, 0.0)
        with pytest.raises(gpu.CcfvError) as ei:
            c.parareal_solve(0.0, 0.5, 64, 8, 2, 1e-15, raise_on_notconv=True)
        assert ei.value.code == T.CCFV_ENOTCONV


@pytest.mark.parametrize("make", [
    lambda: (T.rotation_problem(64), T.rotation_ic()),
    lambda: (T.reference_problem(32, 0.01), T.reference_ic()),
    lambda: (T.swirl_problem(32, T=1.0), T.cosine_bell_ic()),
])
def test_fused_x_launches_change_nothing(gpu, make):
    """The last x sweep of a step and the first of the next as one launch
    (default) against one launch per sweep: bitwise equal, for the plain
    solve, the fine propagator and both Parareal schedules."""
    p, ic = make()
    out = {}
    for fuse in (0, 1):
        with gpu.Context(p) as c:
            c.set_option(T.OPT_FUSE, fuse)
            assert c.get_option(T.OPT_FUSE) == fuse
            c.fill_exact(ic, 0.0)
            c.plain_solve(0.0, 0.25, 20)
            plain = c.get_state()
            c.fine(0.25, 0.125, 5)
            fine = c.get_state()
            res = [plain, fine]
            for pipe in (0, 1):
                c.set_option(T.OPT_PIPELINE, pipe)
                c.fill_exact(ic, 0.0)
                rep = c.parareal_solve(0.0, 0.5, 64, 8, 4, 0.0)
                res.append(c.get_state())
                res.append(np.array(rep.increments))
            out[fuse] = res
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("make", [
    lambda: (T.rotation_problem(64), T.rotation_ic()),
    lambda: T.constant_gaussians(64, 128, -0.7, 0.4, 3e-3,
                                 [(0.3, 0.6, 0.05, 1.0), (0.7, 0.2, 0.08, 0.6)]),
    lambda: (T.advection_sinsin_problem(128), T.advection_sinsin_ic()),
])
@pytest.mark.parametrize("pipe,fuse,graphs", [(0, 1, 0), (1, 1, 0), (1, 0, 0), (1, 1, 1),
                                              (2, 1, 0)])
def test_panel_layout_changes_nothing(gpu, make, pipe, fuse, graphs):
    """Fine-propagator batches with their states in 8-column panels (default)
    against row-major: bitwise equal final state and increments, for every
    schedule, with and without fused x launches and captured phases."""
    p, ic = make()
    out = []
    for panel in (0, 1):
        with gpu.Context(p) as c:
            c.set_option(T.OPT_PANEL, panel)
            assert c.get_option(T.OPT_PANEL) == panel
            c.set_option(T.OPT_PIPELINE, pipe)
            c.set_option(T.OPT_FUSE, fuse)
            c.set_option(T.OPT_GRAPHS, graphs)
            c.set_option(T.OPT_PIPE_GROUP, 3)
            c.fill_exact(ic, 0.0)
            rep = c.parareal_solve(0.0, 0.5, 64, 8, 4, 0.0)
            out.append((c.get_state(), list(rep.increments)))
    assert np.array_equal(out[0][0], out[1][0])
    assert out[0][1] == out[1][1]


def test_panel_layout_full_size(gpu):
    """The 512^2 rotation with 64 slices of 16 fine steps (groups of 32
    slices), two iterations: panel layout against row-major, bitwise."""
    p, ic = T.rotation_problem(512), T.rotation_ic()
    out = []
    for panel in (0, 1):
        with gpu.Context(p) as c:
            c.set_option(T.OPT_PANEL, panel)
            c.fill_exact(ic, 0.0)
            rep = c.parareal_solve(0.0, 1.0, 1024, 64, 2, 0.0)
            out.append((c.get_state(), list(rep.increments)))
    assert np.array_equal(out[0][0], out[1][0])
    assert out[0][1] == out[1][1]


@pytest.mark.parametrize("mode,graphs", [(0, 0), (1, 0), (1, 1), (2, 0)])
def test_kept_iterates_are_the_stopped_solves(gpu, mode, graphs):
    """OPT_KEEP_ITERATES: iterate k of one solve equals, bitwise, the end
    state of a solve stopped after k iterations, in every schedule."""
    p, ic = T.rotation_problem(64), T.rotation_ic()
    kmax = 5
    with gpu.Context(p) as c:
        c.set_option(T.OPT_PIPELINE, mode)
        c.set_option(T.OPT_GRAPHS, graphs)
        c.set_option(T.OPT_PIPE_GROUP, 3)
        c.set_option(T.OPT_KEEP_ITERATES, 1)
        c.fill_exact(ic, 0.0)
        rep = c.parareal_solve(0.0, 0.5, 64, 8, kmax, 0.0)
        assert rep.iterations == kmax
        kept = [c.parareal_iterate(k) for k in range(1, kmax + 1)]
        assert np.array_equal(kept[-1], c.get_state())
        with pytest.raises(gpu.CcfvError):
            c.parareal_iterate(kmax + 1)
    for k in range(1, kmax + 1):
        u, _ = solve(gpu, p, ic, 0, 3, k, 0.0)
        assert np.array_equal(kept[k - 1], u), k
