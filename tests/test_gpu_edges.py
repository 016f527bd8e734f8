"""Edge shapes of the sync path against the oracle, bit for bit: swarms of one
and two particles, sizes on either side of the 64-particle row padding, the
256/512/1024-particle tiles and the 1024-particle single-workgroup limit, the
smallest and largest dimensions, zero-iteration runs, and every launch mode
(one resident workgroup, a captured graph, one launch per iteration) on the
same swarm."""
import numpy as np
import pytest

import pps_oracle as O
from parity_util import assert_same_gbest, assert_same_state, make_pair, note
from ppswarm import _lib as L

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 255, 257, 1023, 1024, 1025, 2049])
@pytest.mark.parametrize("launch", [L.LAUNCH_AUTO, L.LAUNCH_EAGER, L.LAUNCH_GRAPH])
def test_sizes_and_launch_modes(n, launch):
    note(f"edge n={n} launch={launch}")
    g, o = make_pair(O.ROSENBROCK, 3, n, seed=50 + n, lb=-1.0, ub=10.0, launch=launch)
    assert_same_state(g, o, "init")
    assert_same_gbest(g, o, "init")
    for chunk, w, wdamp in ((1, 0.7298, 1.0), (7, 0.9, 0.95), (0, 0.5, 0.5), (12, 0.7298, 1.0)):
        g.run(chunk, w, wdamp)
        o.run_sync(chunk, w, wdamp)
        assert_same_state(g, o, f"after {chunk}")
        assert_same_gbest(g, o, f"after {chunk}")


@pytest.mark.parametrize("obj,d,n", [(O.SPHERE, 1, 5000), (O.RASTRIGIN, 1, 700),
                                     (O.RASTRIGIN, 64, 3000), (O.ROSENBROCK, 64, 900),
                                     (O.SPHERE, 33, 2500), (O.ROSENBROCK, 2, 1)])
def test_extreme_dimensions(obj, d, n):
    note(f"edge dims obj={obj} d={d} n={n}")
    g, o = make_pair(obj, d, n, seed=d * 1000 + n)
    for chunk in (1, 9):
        g.run(chunk)
        o.run_sync(chunk)
        assert_same_state(g, o, f"after {chunk}")
        assert_same_gbest(g, o, f"after {chunk}")


@pytest.mark.parametrize("mode", [L.MODE_ASYNC, L.MODE_QLOCK])
@pytest.mark.parametrize("n", [1, 65, 1023])
def test_async_and_queue_lock_small_swarms(mode, n):
    # one block: both are deterministic and equal the oracle
    note(f"edge mode={mode} n={n}")
    g, o = make_pair(O.RASTRIGIN, 4, n, seed=7 + n, mode=mode)
    for chunk in (3, 0, 5):
        g.run(chunk, 0.7298, 0.98)
        if mode == L.MODE_ASYNC:
            o.run_async(chunk, 0.7298, 0.98)
        else:
            o.run_qlock(chunk, 0.7298, 0.98)
        assert_same_state(g, o, f"after {chunk}")
    assert_same_gbest(g, o)


def test_one_swarm_through_every_launch_mode():
    # the hand-over state (block keys, candidates, parity) is the same whatever
    # launched the previous iterations: alternate launch modes on one swarm by
    # exporting and re-importing its state
    note("edge launch-mode hand-over")
    import ppswarm as P

    d, n = 5, 1000
    lb, ub = np.full(d, -5.12, np.float32), np.full(d, 5.12, np.float32)
    prob = P.OptimizationProblem(P.Rastrigin(), np.zeros(d, np.float32), lb=lb, ub=ub)
    o = O.Swarm(O.Problem(O.RASTRIGIN, d, lb, ub), n, 17)
    swarms = {m: P.Swarm(P.context(), prob, P.swarm_params(n, L.MODE_SYNC, seed=17, launch=m))
              for m in (L.LAUNCH_WORKGROUP, L.LAUNCH_GRAPH, L.LAUNCH_EAGER)}
    state = None
    for k, m in enumerate([L.LAUNCH_WORKGROUP, L.LAUNCH_GRAPH, L.LAUNCH_EAGER] * 2):
        sw = swarms[m]
        if state is not None:
            sw.set_state(*state)
            pos, cost, _ = o.gbest()
            sw.set_gbest(pos, cost)
        sw.run(4)
        o.run_sync(4)
        assert_same_state(sw, o, f"segment {k} launch {m}")
        x, v, b, bc, it = sw.state()
        state = (x, v, b, bc, it)
