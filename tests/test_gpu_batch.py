"""GPU: batches of independent problems of mixed sizes, boundary modes and
velocity fields, issued through shared ragged launches, against the same
problems solved one at a time.  Same per-line arithmetic on the same inputs:
every result must be bitwise equal."""
import numpy as np
import pytest

from ccfv import types as T

pytestmark = pytest.mark.gpu


def mix():
    """(problem, initial condition) pairs covering both sweep paths, both
    boundary modes, every line chunk from 1 to 16 and non-square grids."""
    g, gic = T.constant_gaussians(96, 40, 1.0, 0.5, 1e-3,
                                  [(0.3, 0.4, 0.01, 1.0), (0.7, 0.6, 0.02, -0.5)])
    return [
        (T.rotation_problem(64), T.rotation_ic()),
        (T.reference_problem(32, 0.01), T.reference_ic()),
        (T.swirl_problem(48, T=1.0), T.cosine_bell_ic()),
        (T.rotation_problem(128).copy(ny=64), T.rotation_ic()),
        (g, gic),
        (T.rotation_problem(512), T.rotation_ic()),
        (T.rotation_problem(1024), T.rotation_ic()),
        (T.advection_sinsin_problem(16), T.advection_sinsin_ic()),
    ]


def solo_plain(gpu, p, ic, t0, t1, nt):
    with gpu.Context(p) as c:
        c.fill_exact(ic, t0)
        c.plain_solve(t0, t1, nt)
        return c.get_state()


def solo_parareal(gpu, p, ic, t0, t1, nt, ns, kmax, tol):
    with gpu.Context(p) as c:
        c.fill_exact(ic, t0)
        rep = c.parareal_solve(t0, t1, nt, ns, kmax, tol)
        return c.get_state(), rep


def test_batch_plain_matches_solo(gpu):
    m = mix()
    t0 = [0.0, 0.0, 0.1, 0.0, 0.2, 0.0, 0.0, 0.0]
    t1 = [0.25, 0.3, 0.4, 0.25, 0.5, 0.0625, 0.03125, 0.5]
    nt = [16, 24, 9, 16, 12, 8, 4, 20]
    with gpu.Batch([p for p, _ in m]) as b:
        for i, (_, ic) in enumerate(m):
            b[i].fill_exact(ic, t0[i])
        b.plain(t0, t1, nt)
        got = [b[i].get_state() for i in range(len(m))]
    for i, (p, ic) in enumerate(m):
        ref = solo_plain(gpu, p, ic, t0[i], t1[i], nt[i])
        assert np.array_equal(got[i], ref), i


@pytest.mark.parametrize("kmax,tol", [(3, 0.0), (8, 1e-9)])
def test_batch_parareal_matches_solo(gpu, kmax, tol):
    m = mix()[:5] + [mix()[7]]
    nt = [64, 48, 36, 32, 40, 32]
    ns = [8, 6, 4, 8, 5, 8]
    t1 = [0.5, 0.25, 0.6, 0.25, 0.5, 0.5]
    with gpu.Batch([p for p, _ in m]) as b:
        for i, (_, ic) in enumerate(m):
            b[i].fill_exact(ic, 0.0)
        reps = b.parareal(0.0, t1, nt, ns, kmax, tol)
        got = [b[i].get_state() for i in range(len(m))]
    for i, (p, ic) in enumerate(m):
        ref, rr = solo_parareal(gpu, p, ic, 0.0, t1[i], nt[i], ns[i], kmax, tol)
        assert np.array_equal(got[i], ref), i
        assert reps[i].iterations == rr.iterations, i
        assert reps[i].converged == rr.converged, i
        # the batch runs no speculative fine work: exactly the slices that
        # can still change, (ns - k) * J field-steps in iteration k
        J = nt[i] // ns[i]
        its = reps[i].iterations
        assert reps[i].fine_steps_run == sum((ns[i] - k) * J for k in range(its)), i
        assert reps[i].coarse_steps_run == ns[i] + sum(ns[i] - k - 1 for k in range(its)), i
        # the single-problem schedule may have run one more speculative chain
        assert rr.coarse_steps_run >= reps[i].coarse_steps_run, i
        np.testing.assert_allclose(reps[i].increments, rr.increments, rtol=1e-12)


def test_batch_parareal_large_mix(gpu):
    """512^2 and 1024^2 rotation beside small grids: the chunk-8/16 fast
    path and the chunk-1 generic path in the same rounds."""
    m = [mix()[i] for i in (5, 6, 1, 7)]
    nt, ns = [32, 16, 16, 16], [8, 4, 4, 8]
    t1 = [0.125, 0.0625, 0.1, 0.25]
    with gpu.Batch([p for p, _ in m]) as b:
        for i, (_, ic) in enumerate(m):
            b[i].fill_exact(ic, 0.0)
        reps = b.parareal(0.0, t1, nt, ns, 2, 0.0)
        got = [b[i].get_state() for i in range(len(m))]
    for i, (p, ic) in enumerate(m):
        ref, rr = solo_parareal(gpu, p, ic, 0.0, t1[i], nt[i], ns[i], 2, 0.0)
        assert np.array_equal(got[i], ref), i
        np.testing.assert_allclose(reps[i].increments, rr.increments, rtol=1e-12)


def test_batch_repeated_and_kmax_beyond_slices(gpu):
    """Two solves on one batch (descriptor arenas reused) and kmax above the
    slice count: min(kmax, n_slices) iterations, converged, the serial fine
    solution."""
    m = [mix()[0], mix()[1]]
    with gpu.Batch([p for p, _ in m]) as b:
        for rep_i in range(2):
            for i, (_, ic) in enumerate(m):
                b[i].fill_exact(ic, 0.0)
            reps = b.parareal(0.0, 0.25, 32, 4, 6, 0.0)
            for i, (p, ic) in enumerate(m):
                assert reps[i].iterations == 4 and reps[i].converged == 1
                ref = solo_plain(gpu, p, ic, 0.0, 0.25, 32)
                rel = np.linalg.norm(b[i].get_state() - ref) / np.linalg.norm(ref)
                assert rel <= 1e-12, (rep_i, i)


def test_solo_kmax_beyond_slices(gpu):
    p, ic = mix()[0]
    u, rep = solo_parareal(gpu, p, ic, 0.0, 0.25, 32, 4, 9, 1e-30)
    assert rep.iterations == 4 and rep.converged == 1
    ref = solo_plain(gpu, p, ic, 0.0, 0.25, 32)
    assert np.linalg.norm(u - ref) / np.linalg.norm(ref) <= 1e-12


def test_batch_not_converged_status(gpu):
    m = [mix()[0], mix()[2]]
    with gpu.Batch([p for p, _ in m]) as b:
        for i, (_, ic) in enumerate(m):
            b[i].fill_exact(ic, 0.0)
        with pytest.raises(gpu.CcfvError) as ei:
            b.parareal(0.0, 0.5, 64, 8, 2, 1e-15, raise_on_notconv=True)
        assert ei.value.code == T.CCFV_ENOTCONV


def test_one_failing_problem_leaves_the_others(gpu):
    """A problem whose feet cross (swirl, one step of 1.0 at 64^2) makes the
    batch report CCFV_ECFL; the other problems' results are those they get
    alone."""
    good = mix()[0]
    bad = (T.swirl_problem(64, T=1.0), T.cosine_bell_ic())
    with gpu.Batch([good[0], bad[0]]) as b:
        b[0].fill_exact(good[1], 0.0)
        b[1].fill_exact(bad[1], 0.0)
        with pytest.raises(gpu.CcfvError) as ei:
            b.plain([0.0, 0.0], [0.25, 1.0], [16, 1])
        assert ei.value.code == T.CCFV_ECFL
        got = b[0].get_state()
    assert np.array_equal(got, solo_plain(gpu, good[0], good[1], 0.0, 0.25, 16))


@pytest.mark.parametrize("kmax,ns", [(0, 4), (1, 1), (2, 2)])
def test_batch_small_plans(gpu, kmax, ns):
    """kmax = 0 (the k=0 coarse sweep alone), a single slice, two slices:
    every problem as alone, reports included."""
    m = [mix()[0], mix()[1], (T.rotation_problem(8), T.rotation_ic())]
    nt = [8 * ns, 4 * ns, 2 * ns]
    with gpu.Batch([p for p, _ in m]) as b:
        for i, (_, ic) in enumerate(m):
            b[i].fill_exact(ic, 0.0)
        reps = b.parareal(0.0, 0.25, nt, ns, kmax, 0.0)
        got = [b[i].get_state() for i in range(len(m))]
    for i, (p, ic) in enumerate(m):
        ref, rr = solo_parareal(gpu, p, ic, 0.0, 0.25, nt[i], ns, kmax, 0.0)
        assert np.array_equal(got[i], ref), i
        assert reps[i].iterations == rr.iterations == min(kmax, ns)
        assert reps[i].converged == rr.converged
        assert reps[i].coarse_steps_run == ns + sum(ns - k - 1 for k in range(min(kmax, ns)))


@pytest.mark.parametrize("tune", [False, True])
def test_batch_graphs_and_tuned_shapes_match_solo(gpu, tune):
    """Phases replayed as hipGraphs (the same launch lists recur across
    solves and across the plain rounds), with and without shapes tuned per
    line class: every result bitwise as solved alone, every replay as the
    eager run."""
    m = mix()[:5] + [mix()[7]]
    nt = [64, 48, 36, 32, 40, 32]
    ns = [8, 6, 4, 8, 5, 8]
    t1 = [0.5, 0.25, 0.6, 0.25, 0.5, 0.5]
    with gpu.Batch([p for p, _ in m]) as b:
        if tune:
            b.autotune(0.0, t1, nt, batch_fields=8)
        b.set_option(T.OPT_GRAPHS, 1)
        for rep_i in range(3):
            for i, (_, ic) in enumerate(m):
                b[i].fill_exact(ic, 0.0)
            b.plain(0.0, t1, nt)
            plain = [b[i].get_state() for i in range(len(m))]
            for i, (_, ic) in enumerate(m):
                b[i].fill_exact(ic, 0.0)
            reps = b.parareal(0.0, t1, nt, ns, 3, 0.0)
            para = [b[i].get_state() for i in range(len(m))]
            for i, (p, ic) in enumerate(m):
                assert np.array_equal(plain[i], solo_plain(gpu, p, ic, 0.0, t1[i], nt[i])), \
                    (rep_i, i)
                ref, rr = solo_parareal(gpu, p, ic, 0.0, t1[i], nt[i], ns[i], 3, 0.0)
                assert np.array_equal(para[i], ref), (rep_i, i)
                np.testing.assert_allclose(reps[i].increments, rr.increments, rtol=1e-12)


def many_small(count):
    """`count` problems of one grid class (so one lane and one ragged launch
    per stage holds all of them) with their own velocity, diffusion and
    initial data."""
    rng = np.random.default_rng(7)
    out = []
    for _ in range(count):
        u, v = rng.uniform(-1.0, 1.0, 2)
        nu = 10.0 ** rng.uniform(-4.0, -2.0)
        bumps = [(float(rng.uniform()), float(rng.uniform()), float(rng.uniform(0.05, 0.1)),
                  float(rng.uniform(0.5, 1.0)))]
        out.append(T.constant_gaussians(32, 32, float(u), float(v), float(nu), bumps))
    return out


@pytest.mark.parametrize("count", [65, 150])
def test_batch_many_items_per_launch(gpu, count):
    """More items in one ragged launch than one round of the wave-parallel
    item search covers (64): every block must still find its own problem.
    Step counts differ, so the item lists shrink round by round."""
    m = many_small(count)
    t0 = [0.0] * count
    t1 = [0.125] * count
    nt = [8 + (i % 9) for i in range(count)]
    with gpu.Batch([p for p, _ in m]) as b:
        for i, (_, ic) in enumerate(m):
            b[i].fill_exact(ic, t0[i])
        b.plain(t0, t1, nt)
        got = [b[i].get_state() for i in range(count)]
        for i, (_, ic) in enumerate(m):
            b[i].fill_exact(ic, t0[i])
        reps = b.parareal(t0, t1, [16] * count, [4] * count, 2, 0.0)
        got_p = [b[i].get_state() for i in range(count)]
    for i, (p, ic) in enumerate(m):
        assert np.array_equal(got[i], solo_plain(gpu, p, ic, t0[i], t1[i], nt[i])), i
        ref, rep = solo_parareal(gpu, p, ic, 0.0, 0.125, 16, 4, 2, 0.0)
        assert np.array_equal(got_p[i], ref), i
        assert reps[i].iterations == rep.iterations
