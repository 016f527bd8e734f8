"""Constrained problems on the GPU: the penalty of src/kernels.jl:28-36 and
src/utils.jl:320-372 in every kernel, bit-exact against the oracle, plus the
reference's own constrained test (test/constraints.jl)."""
import numpy as np
import pytest

import pps_oracle as O
import ppswarm as P
from parity_util import assert_same_gbest, assert_same_state, make_pair, note
from ppswarm import _lib as L

pytestmark = pytest.mark.gpu

# test/constraints.jl:11-14: -x1 + 2 x2 - 1 <= 0, x1 - 2 x2 + 1 <= 0, x1^2/4 + x2^2 - 1 <= 0
REF_CONS = P.QuadraticConstraints(linear=[[-1, 2], [1, -2], [0, 0]],
                                  quadratic=[[0, 0], [0, 0], [0.25, 1]], const=[-1, 1, -1])
P_REF = np.array([2.0, 1.0], np.float32)


def ref_pair(n, seed, mode=L.MODE_SYNC, launch=L.LAUNCH_AUTO):
    # objective (x1 - p1)^2 + (x2 - p2)^2 = Sphere shifted by p; bounds [0, 2]^2
    return make_pair(O.SPHERE, 2, n, seed, mode, lb=0.0, ub=2.0, shift=P_REF, cons=REF_CONS,
                     launch=launch)


@pytest.mark.parametrize("n,launch", [(1000, L.LAUNCH_AUTO), (1000, L.LAUNCH_EAGER),
                                      (5000, L.LAUNCH_GRAPH)])
def test_reference_problem_sync_matches_oracle(n, launch):
    note(f"constrained sync n={n} launch={launch}")
    g, o = ref_pair(n, 17, launch=launch)
    assert_same_state(g, o, "after init")
    assert_same_gbest(g, o, "after init")
    for chunk in (1, 9, 40):
        g.run(chunk)
        o.run_sync(chunk)
        assert_same_state(g, o)
        assert_same_gbest(g, o)


def test_reference_problem_async_matches_oracle():
    note("constrained async")
    g, o = ref_pair(3000, 23, L.MODE_ASYNC)
    g.run(30)
    o.run_async(30)
    assert_same_state(g, o)
    assert_same_gbest(g, o)
    g.run(5)            # a second launch restarts the iteration count of h
    o.run_async(5)
    assert_same_state(g, o)


def test_reference_problem_qlock_single_block_matches_oracle():
    note("constrained queue-lock")
    g, o = ref_pair(1000, 29, L.MODE_QLOCK)
    g.run(25)
    o.run_qlock(25)
    assert_same_state(g, o)
    assert_same_gbest(g, o)


@pytest.mark.parametrize("obj,d", [(O.RASTRIGIN, 6), (O.ROSENBROCK, 10), (O.SPHERE, 33)])
def test_random_constraints_multiblock(obj, d):
    note(f"constrained obj={obj} d={d}")
    rng = np.random.default_rng(d)
    cons = P.QuadraticConstraints(linear=rng.normal(0, 1, (3, d)),
                                  quadratic=np.abs(rng.normal(0, 0.2, (3, d))) * (rng.random((3, d)) < 0.5),
                                  const=rng.normal(-1, 1, 3))
    g, o = make_pair(obj, d, 4099, 31 + d, cons=cons, launch=L.LAUNCH_GRAPH)
    assert_same_state(g, o, "after init")
    g.run(12)
    o.run_sync(12)
    assert_same_state(g, o)
    assert_same_gbest(g, o)


@pytest.mark.parametrize("alg", ["sync", "async"])
def test_reference_constraints_suite(alg):
    # test/constraints.jl:24-33: 1000 particles, 500 iterations,
    # |1 - 2 u2 + u1| < 0.1 (the equality pair's residual)
    note(f"reference constraints suite {alg}")
    prob = P.OptimizationProblem(P.Sphere(), np.ones(2, np.float32), lb=np.zeros(2),
                                 ub=np.full(2, 2.0), shift=P_REF, cons=REF_CONS,
                                 lcons=[-np.inf, -np.inf], ucons=[0.0, 0.0])
    opt = (P.ParallelSyncPSOKernel(1000) if alg == "sync"
           else P.ParallelPSOKernel(1000, global_update=False))
    sol = P.solve(prob, opt, maxiters=500)
    assert abs(1 - 2 * sol.u[1] + sol.u[0]) < 0.1
    # the solution's objective is f alone (src/solve.jl:29-31)
    assert sol.objective == pytest.approx(float((sol.u[0] - 2) ** 2 + (sol.u[1] - 1) ** 2), abs=1e-6)


def test_constraint_refusals():
    note("constraint refusals")
    cons = P.QuadraticConstraints(linear=[[1.0, 1.0]], const=[-1.0])
    beale = P.OptimizationProblem(P.Beale(), np.zeros(2), lb=np.full(2, -4.5), ub=np.full(2, 4.5),
                                  cons=cons)
    with pytest.raises(P.PPSError, match="constraints"):
        P.Swarm(P.context(), beale, P.swarm_params(64, L.MODE_SYNC))
    sph = P.OptimizationProblem(P.Sphere(), np.zeros(2), lb=np.full(2, -1.0), ub=np.full(2, 1.0),
                                cons=cons)
    sw = P.Swarm(P.context(), sph, P.swarm_params(256, L.MODE_SYNC, seed=2))
    with pytest.raises(P.PPSError, match="constraint"):
        sw.lbfgs()
    with pytest.raises(ValueError):
        P.QuadraticConstraints(const=np.zeros(9))
