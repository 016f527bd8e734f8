"""The reference's interface on the MI355X backend: solve / init / solve! /
reinit! semantics and the reference test suite's convergence thresholds
(test/gpu.jl:6-58, README.md:14-31), run through ppswarm -> libpps."""
import numpy as np
import pytest

import pps_oracle as O
import ppswarm as P
from ppswarm import _lib as L
from parity_util import assert_same_state, note

pytestmark = pytest.mark.gpu


def rosen_prob(d):
    lb = np.full(d, -1.0, np.float32)
    ub = np.full(d, 10.0, np.float32)
    return P.OptimizationProblem(P.Rosenbrock(), np.zeros(d, np.float32), [1.0, 100.0],
                                 lb=lb, ub=ub)


@pytest.mark.parametrize("d", [2, 3, 4])
def test_reference_gpu_suite_thresholds(d):
    # test/gpu.jl:6-58 (5000 particles, bounds [-1, 10])
    note(f"reference gpu suite d={d}")
    P.seed_(1234)
    prob = rosen_prob(d)
    f_x0 = prob.f(prob.u0, prob.p)
    sol = P.solve(prob, P.ParallelPSOKernel(5000), maxiters=500)
    assert f_x0 > sol.objective and sol.objective < 2e-3 and sol.retcode == "Default"
    sol = P.solve(prob, P.ParallelPSOKernel(5000, global_update=False), maxiters=1000)
    assert f_x0 > sol.objective
    sol = P.solve(prob, P.ParallelSyncPSOKernel(5000), maxiters=500)
    assert f_x0 > sol.objective and sol.objective < 6e-4


def test_readme_example():
    # README.md:14-31: 3-D Rosenbrock, ParallelSyncPSOKernel(1000), 500 iters
    note("readme example")
    prob = rosen_prob(3)
    sol = P.solve(prob, P.ParallelSyncPSOKernel(1000), maxiters=500, seed=2024)
    assert sol.objective < 1e-3
    assert sol.original.shape == (1000, 3)
    assert (sol.original >= -1).all() and (sol.original <= 10).all()
    assert abs(sol.objective - float(sol.objective_f32)) < 1e-5
    assert sol.stats.time > 0


def test_solve_bang_restarts_from_the_cached_gbest():
    # ParallelSyncPSOKernel's solve! does not store its result into
    # cache.gbest (src/solve.jl:36-59): each call starts from the init gbest
    note("solve! twice, sync")
    prob = rosen_prob(3)
    cache = P.init(prob, P.ParallelSyncPSOKernel(800), seed=55)
    oprob = O.Problem(O.ROSENBROCK, 3, prob.lb, prob.ub, np.array([1, 100], np.float32))
    o = O.Swarm(oprob, 800, 55)
    g0 = o.gbest()
    for _ in range(2):
        P.solve_(cache, maxiters=30)
        o.set_gbest(g0[0], g0[1])
        o.run_sync(30)
        assert_same_state(cache.swarm, o)
    P.reinit_(cache)
    o.reinit()
    assert_same_state(cache.swarm, o)


def test_async_solve_bang_continues_from_device_gbest():
    note("solve! twice, async")
    prob = rosen_prob(4)
    cache = P.init(prob, P.ParallelPSOKernel(1500, global_update=False), seed=9)
    oprob = O.Problem(O.ROSENBROCK, 4, prob.lb, prob.ub, np.array([1, 100], np.float32))
    o = O.Swarm(oprob, 1500, 9)
    for _ in range(2):
        sol = P.solve_(cache, maxiters=40)
        o.run_async(40)
        assert_same_state(cache.swarm, o)
        assert sol.index == o.gbest()[2]


def test_unsupported_dimension_is_rejected():
    note("unsupported dimension")
    prob = P.OptimizationProblem(P.Sphere(), np.zeros(65), lb=np.zeros(65), ub=np.ones(65))
    with pytest.raises(P.PPSError):
        P.init(prob, P.ParallelSyncPSOKernel(10))


def test_lexicographic_clamp_is_the_ode_paths_only():
    # whole-vector clamping belongs to parameter_estim_ode! (src/ode_pso.jl:18-19);
    # the other objectives are compiled with component clamping only
    note("lexicographic clamp refused outside the ODE objective")
    d = 3
    prob = P.OptimizationProblem(P.Rosenbrock(), np.zeros(d, np.float32), [1.0, 100.0],
                                 lb=np.full(d, -1.0), ub=np.full(d, 10.0))
    prm = P.swarm_params(64, L.MODE_SYNC, seed=1, clamp=L.CLAMP_LEXICOGRAPHIC)
    with pytest.raises(P.PPSError, match="lexicographic"):
        P.Swarm(P.context(), prob, prm)
