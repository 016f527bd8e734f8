"""Initial-position samplers on the GPU, bit-exact against the oracle:
uniform (default), unbounded (src/utils.jl:23-55; chosen by init when lb/ub are
absent, src/init.jl:70-74), the device Latin hypercube, and host-given
positions (the QuasiMonteCarlo path, src/init.jl:15-36, src/utils.jl:86-112),
through init, iterations and reinit; plus the reference's regression checks for
these samplers (test/regression.jl:53-199)."""
import numpy as np
import pytest

import pps_oracle as O
import ppswarm as P
from parity_util import assert_same_gbest, assert_same_state, note
from ppswarm import _lib as L

pytestmark = pytest.mark.gpu
F = np.float32


def pair(obj, d, n, seed, *, lb, ub, sampler, mode=L.MODE_SYNC, launch=L.LAUNCH_AUTO, u0=None,
         params=(), positions=None):
    f = {O.ROSENBROCK: P.Rosenbrock, O.RASTRIGIN: P.Rastrigin, O.SPHERE: P.Sphere}[obj]()
    u0 = np.zeros(d, F) if u0 is None else np.asarray(u0, F)
    prob = P.OptimizationProblem(f, u0, list(params) or None, lb=lb, ub=ub)
    sp = P.swarm_params(n, mode, seed=seed, launch=launch, sampler=sampler)
    g = P.Swarm(P.context(), prob, sp, positions=positions)
    inf = np.full(d, np.inf, F)
    olb = -inf if lb is None else np.asarray(lb, F)
    oub = inf if ub is None else np.asarray(ub, F)
    osampler = L.SAMPLER_UNBOUNDED if lb is None else sampler
    o = O.Swarm(O.Problem(obj, d, olb, oub, np.asarray(params, F), u0=u0), n, seed,
                sampler=osampler, positions=positions)
    return g, o


@pytest.mark.parametrize("d,n,launch", [(3, 1000, L.LAUNCH_AUTO), (10, 5000, L.LAUNCH_GRAPH),
                                        (17, 3001, L.LAUNCH_EAGER)])
def test_latin_hypercube_matches_oracle(d, n, launch):
    note(f"LHS d={d} n={n}")
    lb, ub = np.full(d, -5.12, F), np.full(d, 5.12, F)
    g, o = pair(O.RASTRIGIN, d, n, 40 + d, lb=lb, ub=ub, sampler=L.SAMPLER_LATIN_HYPERCUBE,
                launch=launch)
    assert_same_state(g, o, "after init")
    assert_same_gbest(g, o, "after init")
    g.run(7)
    o.run_sync(7)
    assert_same_state(g, o, "after 7 iterations")
    g.reinit()
    o.reinit()
    assert_same_state(g, o, "after reinit")
    assert_same_gbest(g, o, "after reinit")


@pytest.mark.parametrize("mode,n", [(L.MODE_SYNC, 1000), (L.MODE_SYNC, 3000), (L.MODE_ASYNC, 2000)])
def test_unbounded_problem_matches_oracle(mode, n):
    note(f"unbounded mode={mode} n={n}")
    u0 = np.array([0.0, 1.5, -2.0], F)
    g, o = pair(O.ROSENBROCK, 3, n, 11, lb=None, ub=None, sampler=L.SAMPLER_UNIFORM, mode=mode,
                u0=u0, params=(1.0, 100.0), launch=L.LAUNCH_GRAPH if n > 1024 else L.LAUNCH_AUTO)
    assert_same_state(g, o, "after init")
    if mode == L.MODE_SYNC:
        g.run(20)
        o.run_sync(20)
    else:
        g.run(20)
        o.run_async(20)
    assert_same_state(g, o, "after 20 iterations")
    assert_same_gbest(g, o)


def test_given_positions_match_oracle():
    note("given positions")
    d, n = 4, 4099
    rng = np.random.default_rng(8)
    lb, ub = np.full(d, -1, F), np.full(d, 10, F)
    pos = rng.uniform(-1, 10, (n, d)).astype(F)
    g, o = pair(O.ROSENBROCK, d, n, 13, lb=lb, ub=ub, sampler=L.SAMPLER_UNIFORM,
                params=(1.0, 100.0), positions=pos, launch=L.LAUNCH_GRAPH)
    assert np.array_equal(g.positions(), pos)
    assert_same_state(g, o, "after init")
    assert_same_gbest(g, o, "after init")
    g.run(5)
    o.run_sync(5)
    assert_same_state(g, o)
    pos2 = rng.uniform(-1, 10, (n, d)).astype(F)
    g.reinit(pos2)
    o.reinit(pos2)
    assert np.array_equal(g.positions(), pos2)
    assert_same_state(g, o, "after reinit with positions")
    assert_same_gbest(g, o, "after reinit with positions")


class _Grid:
    """A host sampler with QuasiMonteCarlo.sample's signature: a shifted
    lattice, different on every call."""

    def __init__(self):
        self.calls = 0

    def sample(self, n, lb, ub):
        self.calls += 1
        d = len(lb)
        k = (np.arange(n)[None, :] * (2 * np.arange(d)[:, None] + 1) + self.calls) % n
        return (np.asarray(lb)[:, None] + (k + 0.5) / n * (np.asarray(ub) - np.asarray(lb))[:, None])


def test_host_sampler_protocol_init_and_reinit():
    note("host sampler")
    prob = P.OptimizationProblem(P.Sphere(), np.zeros(3), lb=np.full(3, -2.0), ub=np.full(3, 2.0))
    smp = _Grid()
    cache = P.init(prob, P.ParallelSyncPSOKernel(512), seed=3, sampler=smp)
    first = cache.swarm.positions()
    np.testing.assert_array_equal(first, np.asarray(_Grid().sample(512, [-2] * 3, [2] * 3), F).T)
    P.reinit_(cache)
    assert smp.calls == 2 and not np.array_equal(cache.swarm.positions(), first)
    sol = P.solve_(cache, 200)
    assert sol.objective < 1e-4


@pytest.mark.parametrize("N", [2, 3])
def test_reference_sampler_regressions(N):
    # test/regression.jl:5-199, Rosenbrock p = [1, 100], 2000 particles, 500 iterations
    note(f"reference sampler regressions N={N}")
    lb, ub = np.full(N, -1.0, F), np.full(N, 10.0, F)
    prob = P.OptimizationProblem(P.Rosenbrock(), np.zeros(N, F), [1.0, 100.0], lb=lb, ub=ub)
    for alg in (P.ParallelPSOKernel(2000), P.ParallelSyncPSOKernel(2000)):
        sol = P.solve_(P.init(prob, alg, sampler=P.LatinHypercubeSample()), 500)
        assert sol.retcode == "Default" and np.isfinite(sol.objective)
    unb = P.OptimizationProblem(P.Rosenbrock(), np.zeros(N, F), [1.0, 100.0])
    assert P.solve(unb, P.ParallelPSOKernel(2000), maxiters=500).objective < 1e-4
    assert P.solve(unb, P.ParallelSyncPSOKernel(2000), maxiters=500).objective < 2e-2
    # a sampler asked for on an unbounded problem gives way to the unbounded one
    sol = P.solve_(P.init(unb, P.ParallelSyncPSOKernel(2000), sampler=P.LatinHypercubeSample()), 500)
    assert sol.retcode == "Default" and np.isfinite(sol.objective)


def test_sampler_refusals():
    note("sampler refusals")
    mixed = P.OptimizationProblem(P.Sphere(), np.zeros(2), lb=[-np.inf, 0.0], ub=[np.inf, 1.0])
    with pytest.raises(P.PPSError, match="finite"):
        P.Swarm(P.context(), mixed, P.swarm_params(64, L.MODE_SYNC))
    ok = P.OptimizationProblem(P.Sphere(), np.zeros(2), lb=[0.0, 0.0], ub=[1.0, 1.0])
    with pytest.raises(P.PPSError, match="sampler"):
        P.Swarm(P.context(), ok, P.swarm_params(64, L.MODE_SYNC, sampler=7))
    with pytest.raises(P.PPSError, match="not finite"):
        P.Swarm(P.context(), ok, P.swarm_params(4, L.MODE_SYNC),
                positions=np.array([[0, 0], [1, np.nan], [0, 1], [1, 1]], F))
