"""GPU check of the context's bounded table cache, through the C ABI.

A context keeps the solve tables of at most 32 (direction, step size) pairs
and drops them all, with the captured graphs that point at them, at the start
of the next one-context call once it holds more.  A caller that changes the
step size every step (40 distinct sizes: 80 tables) must see the results a
fresh context gives for every step, and a plain solve replayed from a graph
captured before the drop must be captured again, not replayed with freed
tables.
"""
import numpy as np
import pytest

from ccfv import types as T

pytestmark = pytest.mark.gpu


def test_many_step_sizes_match_fresh_contexts(gpu):
    p, ic = T.rotation_problem(64), T.rotation_ic()
    hs = [2e-3 * (1.0 + 0.01 * k) for k in range(40)]
    print("table cache: 40 step sizes at 64^2", flush=True)
    with gpu.Context(p) as c:
        c.set_option(T.OPT_PLAIN_GRAPHS, 1)
        c.fill_exact(ic, 0.0)
        u0 = c.get_state()
        # a captured plain solve whose tables the drop below frees
        c.plain_solve(0.0, 0.05, 8)
        plain_first = c.get_state()
        c.set_state(u0)
        t, steps = 0.0, []
        for h in hs:
            c.step(t, h)
            t += h
            steps.append(c.get_state())
        # the same plain solve again, from the same state
        c.set_state(u0)
        c.plain_solve(0.0, 0.05, 8)
        plain_again = c.get_state()
    assert np.array_equal(plain_again, plain_first)
    u, t = u0, 0.0
    for k, h in enumerate(hs):
        with gpu.Context(p) as f:
            f.set_state(u)
            f.step(t, h)
            u = f.get_state()
        assert np.array_equal(u, steps[k]), k
        t += h
    with gpu.Context(p) as f:
        f.set_state(u0)
        f.plain_solve(0.0, 0.05, 8)
        assert np.array_equal(f.get_state(), plain_first)
