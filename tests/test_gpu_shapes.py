"""GPU: launch shapes and batch composition do not change the arithmetic.

Every line is swept by one wavefront with the same per-line code whatever
the tile, the batch size or the position of a field in a batch, so results
must be bitwise identical across all of them."""
import numpy as np
import pytest

from ccfv import types as T

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("make", [
    lambda: (T.rotation_problem(128), T.rotation_ic()),
    lambda: (T.reference_problem(64, 0.01), T.reference_ic()),
    lambda: (T.swirl_problem(64, T=1.0), T.cosine_bell_ic()),
])
def test_results_independent_of_tiles(gpu, make):
    p, ic = make()
    outs = []
    for tiles in ([4, 4, 8, 4, 1, 1, 2, 1], [2, 2, 16, 4, 2, 2, 8, 4],
                  [8, 4, 4, 2, 4, 4, 4, 2]):
        with gpu.Context(p) as c:
            c.fill_exact(ic, 0.0)
            c.set_tiles(tiles)
            assert c.get_tiles() == tiles
            c.plain_solve(0.0, 0.25, 16)
            outs.append(c.get_state())
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])


def test_parareal_independent_of_tiles(gpu):
    p, ic = T.rotation_problem(64), T.rotation_ic()
    outs = []
    for tiles in ([4, 4, 8, 4, 1, 1, 2, 1], [2, 1, 2, 2, 4, 4, 8, 4]):
        with gpu.Context(p) as c:
            c.fill_exact(ic, 0.0)
            c.set_tiles(tiles)
            c.parareal_solve(0.0, 0.5, 64, 8, 3)
            outs.append(c.get_state())
    assert np.array_equal(outs[0], outs[1])


def test_autotune_picks_valid_shapes_and_keeps_results(gpu):
    p, ic = T.rotation_problem(128), T.rotation_ic()
    with gpu.Context(p) as c:
        c.fill_exact(ic, 0.0)
        u0 = c.get_state()
        c.plain_solve(0.0, 0.25, 16)
        ref = c.get_state()
        tiles = c.autotune(16, 1.0 / 64)
        assert c.get_tiles() == tiles
        np.testing.assert_array_equal(c.get_state(), ref)  # state untouched
        c.set_state(u0)
        c.plain_solve(0.0, 0.25, 16)
        assert np.array_equal(c.get_state(), ref)


def test_invalid_tiles_rejected(gpu):
    with gpu.Context(T.rotation_problem(64)) as c:
        for bad in ([3, 1, 8, 4, 1, 1, 2, 1], [8, 8, 8, 4, 1, 1, 2, 1],
                    [128, 4, 8, 4, 1, 1, 2, 1]):
            with pytest.raises(gpu.CcfvError):
                c.set_tiles(bad)
