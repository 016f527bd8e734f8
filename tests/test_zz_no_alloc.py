"""The allocation-free local search, pps_swarm_lbfgs_noalloc, through the C ABI.

Its device scratch is one file-scope array per GPU that every call reuses, so
the test makes two calls in a row on different swarms, the second smaller,
and checks each against pps_swarm_lbfgs (library-allocated buffers) bit for
bit: every particle's local result (AoS), its objective value, and the argmin
(position, value, index).  A swarm whose scratch need exceeds the array is
refused with PPS_E_INVALID before any launch.
"""
import ctypes as C

import numpy as np
import pytest

import ppswarm as P
from parity_util import note
from ppswarm import _lib as L

pytestmark = pytest.mark.gpu


def _swarm(obj, d, n, seed):
    f = P.Rastrigin() if obj == "rastrigin" else P.Rosenbrock()
    p = None if obj == "rastrigin" else [1.0, 100.0]
    prob = P.OptimizationProblem(f, np.zeros(d, np.float32), p, lb=np.full(d, -5.12),
                                 ub=np.full(d, 5.12))
    sw = P.Swarm(P.context(), prob, P.swarm_params(n, L.MODE_SYNC, seed=seed))
    sw.run(20)
    return sw


def _call(fn, sw, lp):
    n, d = sw.n, sw.d
    ox = np.full((n, d), np.nan, np.float32)
    of = np.full(n, np.nan, np.float32)
    bx = np.zeros(d, np.float32)
    bf = C.c_float()
    bi = C.c_int64()
    L.check(fn(sw.h, C.byref(lp), ox.ctypes.data_as(L.fp), of.ctypes.data_as(L.fp),
               bx.ctypes.data_as(L.fp), C.byref(bf), C.byref(bi)))
    return ox, of, bx, np.float32(bf.value), int(bi.value)


def _lp(maxiters=60, history=10):
    lp = L.pps_lbfgs_params()
    L.lib().pps_lbfgs_params_default(C.byref(lp))
    lp.maxiters, lp.history = maxiters, history
    return lp


def test_noalloc_local_search_equals_the_allocating_one():
    lib = L.lib()
    lp = _lp()
    for obj, d, n, seed in (("rastrigin", 16, 6000, 5), ("rosenbrock", 4, 700, 6)):
        note(f"noalloc lbfgs obj={obj} d={d} n={n}")
        sw = _swarm(obj, d, n, seed)
        want = _call(lib.pps_swarm_lbfgs, sw, lp)
        got = _call(lib.pps_swarm_lbfgs_noalloc, sw, lp)
        for name, a, b in zip(("x", "f", "best_x", "best_f", "best_idx"), got, want):
            np.testing.assert_array_equal(a, b, err_msg=f"{name} ({obj}, d={d}, n={n})")
        assert np.isfinite(got[1]).all()
        sw.free()


def test_noalloc_local_search_refuses_more_than_its_scratch():
    lib = L.lib()
    # 2^20 particles x 16-D need ~4 * 2^20 * 33 bytes > 64 MiB of scratch
    note("noalloc lbfgs refusal")
    sw = _swarm("rastrigin", 16, 1 << 20, 7)
    lp = _lp(maxiters=1)
    bf = C.c_float()
    rc = lib.pps_swarm_lbfgs_noalloc(sw.h, C.byref(lp), None, None, None, C.byref(bf), None)
    assert rc == L.E_INVALID
    assert b"scratch" in lib.pps_last_error()
    sw.free()
