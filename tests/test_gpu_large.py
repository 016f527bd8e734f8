"""Parity at the full single-GPU benchmark shape: 2^20 particles, 10-D
Rastrigin and 10-D Rosenbrock, bounds [-5.12, 5.12] — the whole swarm compared
bit for bit with the oracle for a few iterations, plus size-independent
properties over a longer run; and two states past the Infinity Cache, where
the sync step switches to nontemporal streams."""
import numpy as np
import pytest

import pps_oracle as O
from parity_util import assert_same_gbest, assert_same_state, make_pair, note
from ppswarm import _lib as L

pytestmark = pytest.mark.gpu

N = 1 << 20


@pytest.mark.parametrize("obj", [O.RASTRIGIN, O.ROSENBROCK])
def test_sync_full_shape_bitwise(obj):
    note(f"sync 2^20 obj={obj}")
    g, o = make_pair(obj, 10, N, seed=2 ** 40 + obj, launch=L.LAUNCH_GRAPH)
    assert_same_state(g, o, "after init")
    g.run(3)
    o.run_sync(3)
    assert_same_state(g, o, "after 3 iterations")
    assert_same_gbest(g, o)


@pytest.mark.parametrize("obj,d,n", [(O.RASTRIGIN, 32, 1 << 20), (O.ROSENBROCK, 10, 1 << 22)])
def test_sync_beyond_the_cache_bitwise(obj, d, n):
    # states of 407 MB and 520 MB: the sync step streams them with
    # nontemporal hints (chosen above 288 MiB); values must not change
    note(f"sync beyond cache obj={obj} d={d} n={n}")
    g, o = make_pair(obj, d, n, seed=99 + d, launch=L.LAUNCH_GRAPH)
    g.run(2)
    o.run_sync(2)
    assert_same_state(g, o, "after 2 iterations")
    assert_same_gbest(g, o)


def test_async_full_shape_bitwise():
    note("async 2^20 rastrigin")
    g, o = make_pair(O.RASTRIGIN, 10, N, seed=31337, mode=L.MODE_ASYNC)
    g.run(4)
    o.run_async(4)
    assert_same_state(g, o, "after 4 iterations")
    assert_same_gbest(g, o)


def test_sync_full_shape_properties():
    note("sync 2^20 properties over 100 iterations")
    g, _ = make_pair(O.RASTRIGIN, 10, N, seed=4, launch=L.LAUNCH_GRAPH)
    _, _, _, bc0, _ = g.state()
    prev = g.gbest().cost
    for _ in range(4):
        g.run(25)
        x, v, b, bc, it = g.state()
        gb = g.gbest()
        assert (bc <= bc0).all()                      # best costs never rise
        assert gb.cost <= prev                        # global best never rises
        assert gb.index == int(np.argmin(bc))         # lowest-index argmin
        assert np.float32(gb.cost) == bc.min()
        np.testing.assert_array_equal(gb.position, b[:, gb.index])
        assert (x >= np.float32(-5.12)).all() and (x <= np.float32(5.12)).all()
        bc0, prev = bc, gb.cost
