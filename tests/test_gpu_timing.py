"""GPU checks of the timing entry points, through the C ABI: ccfv_time_sweep
(one launch class back to back) and ccfv_dispatch_floor (empty kernel, copy
with a sweep's memory pattern, a plain solve's alternating pattern) return a
positive time per launch, eager and replayed from a graph, and leave the
state as it was; invalid arguments are errors."""
import math

import numpy as np
import pytest

import ccfv
from ccfv import types as T

pytestmark = pytest.mark.gpu


def test_timers_leave_the_state_alone(gpu):
    p, ic = T.rotation_problem(128), T.rotation_ic()
    print("timers at 128^2", flush=True)
    with gpu.Context(p) as c:
        c.fill_exact(ic, 0.0)
        u0 = c.get_state()
        got = []
        for graph in (False, True):
            for d in (0, 1):
                got.append(c.time_sweep(d, d == 0, 1, 1.0 / 256, 64, graph=graph))
                for kind in (0, 1):
                    got.append(c.dispatch_floor(d, kind, 64, graph=graph))
            got.append(c.dispatch_floor(0, 2, 64, graph=graph))
        assert all(math.isfinite(v) and v > 0 for v in got), got
        assert np.array_equal(c.get_state(), u0)
        for bad in ((0, 3, 64), (2, 0, 64), (0, 1, 0)):
            with pytest.raises(ccfv.CcfvError):
                c.dispatch_floor(*bad)
