"""C-ABI surface and host logic (no GPU calls).

* libpps.so loads and exports every function include/pps.h declares;
* the pure-host entry points (defaults, ABI version, FP64 evaluator) behave;
* the Python mirror of the reference interface validates its inputs the way
  the reference does (bounds checked, unknown objectives rejected).
"""
import ctypes as C
import math

import numpy as np
import pytest


def test_library_exports_every_declared_symbol(pps):
    from ppswarm import _lib

    names = _lib.declared_functions()
    assert "pps_swarm_init" in names and "pps_eval_f64" in names
    L = _lib.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_abi_version_and_defaults(pps):
    from ppswarm import _lib

    L = _lib.lib()
    assert L.pps_abi_version() == _lib.ABI_VERSION == 3
    p = _lib.pps_swarm_params()
    L.pps_swarm_params_default(C.byref(p))
    # reference defaults: c1 = c2 = 1.4962f0 (src/kernels.jl:41-42), queue
    # publication at >= 2 improvers (src/kernels.jl:97)
    assert abs(p.c1 - 1.4962) < 1e-6 and abs(p.c2 - 1.4962) < 1e-6
    assert p.mode == _lib.MODE_SYNC and p.gbest_from == _lib.GBEST_PBEST
    assert p.qlock_min_improvers == 2 and p.launch == _lib.LAUNCH_AUTO


def _prob(pps, f, d, p=None, lb=-5.0, ub=5.0, shift=None):
    return pps.OptimizationProblem(f, np.zeros(d, np.float32), p, lb=np.full(d, lb),
                                   ub=np.full(d, ub), shift=shift)


def test_eval_f64_closed_forms(pps):
    R = _prob(pps, pps.Rosenbrock(), 3, [1.0, 100.0])
    np.testing.assert_allclose(pps.eval_f64(R, [[1, 1, 1], [0, 0, 0]]), [0.0, 2.0])
    A = _prob(pps, pps.Rastrigin(), 10)
    assert pps.eval_f64(A, np.zeros((1, 10)))[0] == 0.0
    x = np.full((1, 10), 0.5)
    assert abs(pps.eval_f64(A, x)[0] - (100 + 10 * (0.25 + 10))) < 1e-9
    sh = np.linspace(-1, 1, 4)
    S = _prob(pps, pps.Rastrigin(), 4, shift=sh)
    assert abs(pps.eval_f64(S, sh[None, :])[0]) < 1e-6
    assert pps.eval_f64(_prob(pps, pps.Beale(), 2), [[3, 0.5]])[0] == 0.0
    assert pps.eval_f64(_prob(pps, pps.Booth(), 2), [[1, 3]])[0] == 0.0
    assert pps.eval_f64(_prob(pps, pps.Himmelblau(), 2), [[3, 2]])[0] == 0.0
    assert pps.eval_f64(_prob(pps, pps.Quadratic(), 2, [1000.0]), [[1, -2]])[0] == 0.0
    assert pps.eval_f64(_prob(pps, pps.BoundaryQuadratic(), 1, lb=0, ub=2), [[0.0]])[0] == 1.0
    assert pps.eval_f64(_prob(pps, pps.Rosen2(), 2, [1.0, 100.0]), [[1, 1]])[0] == 0.0


def test_eval_f64_lotka_volterra_reduce_losses(pps):
    # loss of a zero trajectory against data [1, 2] at 5 samples = 25
    # (test/regression.jl:345-357)
    f = pps.lotka_volterra_sse(np.tile([1.0, 2.0], (5, 1)), u0=(0.0, 0.0), dt=0.01,
                               steps_per_obs=10)
    prob = pps.OptimizationProblem(f, np.zeros(4), lb=np.zeros(4), ub=np.full(4, 5.0))
    assert pps.eval_f64(prob, np.zeros((1, 4)))[0] == 25.0


def test_problem_validation(pps):
    with pytest.raises(ValueError):
        pps.OptimizationProblem(lambda x, p: sum(x), np.zeros(2), lb=[0, 0], ub=[1, 1])
    with pytest.raises(ValueError):
        pps.OptimizationProblem(pps.Rosenbrock(), np.zeros(2), [1.0, 100.0], lb=[0, 0])
    # no bounds at all: the unbounded sampler's problem (src/init.jl:70-74)
    assert pps.OptimizationProblem(pps.Rosenbrock(), np.zeros(2), [1.0, 100.0]).lb is None
    with pytest.raises(ValueError):
        pps.OptimizationProblem(pps.Beale(), np.zeros(3), lb=np.zeros(3), ub=np.ones(3))
    with pytest.raises(ValueError):
        pps.OptimizationProblem(pps.Rosenbrock(), np.zeros(2), None, lb=[0, 0], ub=[1, 1])


def test_algorithm_descriptors_mirror_reference(pps):
    # defaults of src/algorithms.jl:192-249
    a = pps.ParallelPSOKernel(100)
    assert a.global_update is True
    h = pps.HybridPSO()
    assert isinstance(h.pso, pps.ParallelPSOKernel) and h.pso.num_particles == 100
    assert h.pso.global_update is False
    assert isinstance(h.local_opt, pps.SimpleLBFGS)


def test_seed_sequence_is_reproducible(pps):
    from ppswarm import api

    pps.seed_(123)
    a = [api._next_seed() for _ in range(3)]
    pps.seed_(123)
    b = [api._next_seed() for _ in range(3)]
    assert a == b and len(set(a)) == 3


def test_served_mix_generator(pps):
    reqs, arr = pps.served_mix(2000, seed=11, qps=1000.0)
    again, arr2 = pps.served_mix(2000, seed=11, qps=1000.0)
    assert [q.seed for q in reqs] == [q.seed for q in again]
    np.testing.assert_array_equal(arr, arr2)
    ds = np.array([q.prob.d for q in reqs])
    ns = np.array([q.n for q in reqs])
    ts = np.array([q.maxiters for q in reqs])
    assert set(ds) == {2, 4, 8, 16, 32}
    assert ns.min() >= 256 and ns.max() <= 16384
    # log-uniform: the median is near sqrt(256 * 16384) = 2048
    assert 1500 < np.median(ns) < 2800
    assert ts.min() >= 100 and ts.max() <= 1000
    fams = {q.prob.f.name for q in reqs}
    assert fams == {"rosenbrock", "rastrigin"}
    assert all((np.abs(q.prob.shift) <= 1).all() for q in reqs)
    # Poisson arrivals: mean gap 1/qps
    assert abs(np.diff(arr).mean() - 1e-3) < 1e-4


def test_hybrid_refuses_what_the_local_search_does_not_implement(pps):
    from ppswarm import api, hybrid

    # line searches other than Strong-Wolfe / Armijo (the reference throws
    # ArgumentError for anything but StrongWolfeLineSearch, src/hybrid.jl:129-130)
    with pytest.raises(ValueError):
        api.lbfgs_params(10, 10, 0.0, object())
    lp = api.lbfgs_params(10, 10, 0.0, pps.StrongWolfeLineSearch(c2=0.5), reltol=1e-6)
    assert lp.linesearch == 1 and abs(lp.wolfe_c2 - 0.5) < 1e-7 and lp.reltol > 0
    lp = api.lbfgs_params(10, 10, 0.0, pps.ArmijoBacktracking(rho=0.25))
    assert lp.linesearch == 0 and lp.armijo_rho == 0.25
    # the reference's defaults: SimpleLBFGS() searches with Strong-Wolfe
    assert isinstance(pps.SimpleLBFGS().linesearch, pps.StrongWolfeLineSearch)
    prob = _prob(pps, pps.Sphere(), 3)
    with pytest.raises(ValueError):   # local optimisers other than SimpleLBFGS
        pps.init(prob, pps.HybridPSO(local_opt=object()))
    cache = api.HybridPSOCache(None, pps.HybridPSO())
    with pytest.raises(ValueError):
        hybrid.solve_hybrid_cache(cache, 10, reltol=-1.0)
    with pytest.raises(ValueError):
        hybrid.solve_hybrid_cache(cache, 10, linesearch="hager-zhang")
