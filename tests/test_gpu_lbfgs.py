"""HybridPSO local phase on the GPU: projected L-BFGS kernel vs the oracle
(bit for bit: positions and values of every particle, argmin index) with the
Armijo search the benchmark names and the Strong-Wolfe search the reference
uses, the reference's known-answer hybrid tests (test/lbfgs.jl:38-72) through
solve(prob, HybridPSO(...)) (Strong-Wolfe, the default), and the
init / solve! / reinit! flow of a HybridPSO cache."""
import numpy as np
import pytest

import pps_oracle as O
import ppswarm as P
from parity_util import make_pair, note
from ppswarm import _lib as L

pytestmark = pytest.mark.gpu


def _run_both(g, o, maxiters, history=10, abstol=0.0):
    gx, gf, bx, bf, bi = g.lbfgs(maxiters=maxiters, history=history, abstol=abstol)
    ox, of = O.lbfgs_swarm(o, O.lbfgs_params(maxiters=maxiters, history=history, abstol=abstol))
    return gx, gf, bx, bf, bi, ox, of


@pytest.mark.parametrize("obj,d,n,lb,ub", [
    (O.RASTRIGIN, 16, 4096, -5.12, 5.12),   # G = 16, the benchmark layout
    (O.ROSENBROCK, 10, 2000, -5.0, 5.0),    # G = 16 with 6 padding lanes
    (O.SPHERE, 5, 1000, -1.0, 2.0),         # G = 8
    (O.RASTRIGIN, 64, 256, -5.12, 5.12),    # G = 64, one particle per wave
    (O.ROSENBROCK, 20, 512, -5.0, 5.0),     # G = 32 with 12 padding lanes
    (O.ROSENBROCK, 3, 999, -1.0, 10.0),     # G = 4
    (O.BEALE, 2, 500, -4.5, 4.5),
    (O.HIMMELBLAU, 2, 500, -5.0, 5.0),
    (O.BOUNDARY_QUADRATIC, 1, 300, 0.0, 2.0),  # G = 1
])
def test_lbfgs_matches_oracle(obj, d, n, lb, ub):
    note(f"lbfgs parity obj={obj} d={d} n={n}")
    g, o = make_pair(obj, d, n, seed=600 + d, lb=lb, ub=ub, launch=L.LAUNCH_EAGER)
    g.run(5)
    o.run_sync(5)
    gx, gf, bx, bf, bi, ox, of = _run_both(g, o, maxiters=40)
    np.testing.assert_array_equal(gx, ox)
    np.testing.assert_array_equal(gf, of)
    i = int(np.argmin(of))
    assert bi == i and bf == of[i]
    np.testing.assert_array_equal(bx, ox[i])
    # the swarm itself is untouched
    x, v, b, bc, it = g.state()
    np.testing.assert_array_equal(b, o.best)


def test_lbfgs_c2_shape_matches_oracle():
    # 16-D Rastrigin, 2^16 particles, 200 sync iterations, then L-BFGS m=10 x 100
    note("lbfgs C2 shape")
    g, o = make_pair(O.RASTRIGIN, 16, 1 << 16, seed=2, launch=L.LAUNCH_GRAPH)
    g.run(200)
    o.run_sync(200)
    gx, gf, bx, bf, bi, ox, of = _run_both(g, o, maxiters=100, history=10)
    np.testing.assert_array_equal(gf, of)
    np.testing.assert_array_equal(gx, ox)
    assert bi == int(np.argmin(of))
    # the local phase never makes a particle worse than its PSO best
    assert (of <= o.best_cost + 1e-4).all()


SHAPES = [
    (O.RASTRIGIN, 16, 4096, -5.12, 5.12),
    (O.ROSENBROCK, 10, 2000, -5.0, 5.0),
    (O.SPHERE, 5, 1000, -1.0, 2.0),
    (O.RASTRIGIN, 64, 256, -5.12, 5.12),
    (O.ROSENBROCK, 20, 512, -5.0, 5.0),
    (O.ROSENBROCK, 3, 999, -1.0, 10.0),
    (O.BEALE, 2, 500, -4.5, 4.5),
    (O.HIMMELBLAU, 2, 500, -5.0, 5.0),
    (O.BOUNDARY_QUADRATIC, 1, 300, 0.0, 2.0),
]


def _run_both_wolfe(g, o, maxiters, history=10, abstol=0.0, reltol=0.0, **ls):
    sw = P.StrongWolfeLineSearch(**ls)
    gx, gf, bx, bf, bi = g.lbfgs(maxiters=maxiters, history=history, abstol=abstol,
                                 linesearch=sw, reltol=reltol)
    ox, of = O.lbfgs_swarm(o, O.lbfgs_params(
        maxiters=maxiters, history=history, abstol=abstol, linesearch=O.LS_STRONG_WOLFE,
        armijo_c=sw.c1, wolfe_c2=sw.c2, alpha_init=sw.alpha_init, alpha_max=sw.alpha_max,
        ls_maxiters=sw.maxiters, zoom_maxiters=sw.zoom_maxiters, reltol=reltol))
    return gx, gf, bx, bf, bi, ox, of


@pytest.mark.parametrize("obj,d,n,lb,ub", SHAPES)
def test_lbfgs_strong_wolfe_matches_oracle(obj, d, n, lb, ub):
    # the reference's line search (src/hybrid.jl:129-136), kernel vs oracle
    note(f"lbfgs strong-wolfe parity obj={obj} d={d} n={n}")
    g, o = make_pair(obj, d, n, seed=700 + d, lb=lb, ub=ub, launch=L.LAUNCH_EAGER)
    g.run(5)
    o.run_sync(5)
    gx, gf, bx, bf, bi, ox, of = _run_both_wolfe(g, o, maxiters=40)
    np.testing.assert_array_equal(gx, ox)
    np.testing.assert_array_equal(gf, of)
    i = int(np.argmin(of))
    assert bi == i and bf == of[i]
    np.testing.assert_array_equal(bx, ox[i])


@pytest.mark.parametrize("ls", [dict(), dict(c2=0.1, alpha_init=0.25, zoom_maxiters=5)])
def test_lbfgs_strong_wolfe_settings_and_reltol_match_oracle(ls):
    note(f"lbfgs strong-wolfe settings {ls}")
    g, o = make_pair(O.ROSENBROCK, 8, 3000, seed=77, lb=-5.0, ub=5.0, launch=L.LAUNCH_EAGER)
    g.run(3)
    o.run_sync(3)
    for reltol in (0.0, 1e-3):
        gx, gf, bx, bf, bi, ox, of = _run_both_wolfe(g, o, maxiters=60, reltol=reltol, **ls)
        np.testing.assert_array_equal(gx, ox)
        np.testing.assert_array_equal(gf, of)


def test_lbfgs_strong_wolfe_c2_shape_matches_oracle():
    note("lbfgs strong-wolfe C2 shape")
    g, o = make_pair(O.RASTRIGIN, 16, 1 << 16, seed=3, launch=L.LAUNCH_GRAPH)
    g.run(200)
    o.run_sync(200)
    gx, gf, bx, bf, bi, ox, of = _run_both_wolfe(g, o, maxiters=100)
    np.testing.assert_array_equal(gf, of)
    np.testing.assert_array_equal(gx, ox)
    assert bi == int(np.argmin(of))
    assert (of <= o.best_cost + 1e-4).all()


def test_hybrid_init_solve_reinit():
    # init(prob, HybridPSO) -> solve! -> reinit! -> solve! (src/init.jl:143-168,
    # src/hybrid.jl:100-158): solve!(init(...)) equals solve(...) bit for bit,
    # and after reinit! the PSO stage and the local phase equal the oracle's
    # next-epoch swarm and its L-BFGS
    note("hybrid init / solve! / reinit!")
    d, n, T, LT = 8, 3000, 30, 40
    lb, ub = np.full(d, -5.12, np.float32), np.full(d, 5.12, np.float32)
    prob = P.OptimizationProblem(P.Rastrigin(), np.zeros(d, np.float32), lb=lb, ub=ub)
    alg = P.HybridPSO(pso=P.ParallelSyncPSOKernel(n))
    a = P.solve(prob, alg, maxiters=T, local_maxiters=LT, seed=11)
    cache = P.init(prob, alg, seed=11)
    assert isinstance(cache, P.HybridPSOCache) and cache.swarm is cache.pso_cache.swarm
    b = P.solve_(cache, maxiters=T, local_maxiters=LT)
    np.testing.assert_array_equal(a.u, b.u)
    assert a.objective == b.objective and a.objective_f32 == b.objective_f32
    assert a.index == b.index
    P.reinit_(cache)
    c = P.solve_(cache, maxiters=T, local_maxiters=LT)
    oprob = O.Problem(O.RASTRIGIN, d, lb, ub, np.zeros(0, np.float32))
    o = O.Swarm(oprob, n, 11)
    o.reinit()
    o.run_sync(T)
    pos, cost, idx = o.gbest()
    ox, of = O.lbfgs_swarm(o, O.lbfgs_params(maxiters=LT, history=10,
                                            linesearch=O.LS_STRONG_WOLFE))
    i = int(np.argmin(of))
    f_pso = P.eval_f64(prob, pos.astype(np.float64)[None, :])[0]
    f_loc = P.eval_f64(prob, ox[i].astype(np.float64)[None, :])[0]
    want = ox[i] if f_loc < f_pso else pos
    np.testing.assert_array_equal(c.u, want)
    assert c.objective == min(f_pso, f_loc)


KNOWN = [
    (P.Rosen2, [1.0, 100.0], [0.0, 0.0], None, [1, 1], 0.0),
    (P.Rosen2, [1.0, 100.0], [-1.2, 1.0], None, [1, 1], 0.0),
    (P.Rosen2, [1.0, 100.0], [0.0, 0.0], (-2.0, 2.0), [1, 1], 0.0),
    (P.Beale, None, [1.0, 1.0], None, [3, 0.5], 0.0),
    (P.Booth, None, [1.0, 1.0], None, [1, 3], 0.0),
    (P.Himmelblau, None, [1.0, 1.0], None, None, 0.0),
    (P.Quadratic, [1000.0], [4.0, 4.0], (-5.0, 5.0), [1, -2], 0.0),
]


@pytest.mark.parametrize("f,p,u0,bounds,ustar,fstar", KNOWN)
def test_hybrid_known_answers(f, p, u0, bounds, ustar, fstar):
    # test/lbfgs.jl:5-65: HybridPSO(pso=ParallelSyncPSOKernel(64), SimpleLBFGS()),
    # maxiters=50, local_maxiters=200, abstol=1e-10 (unbounded problems get
    # the box [-10, 10]: this backend samples the initial swarm in the bounds)
    note(f"hybrid known answer {f().name} u0={u0}")
    lb, ub = bounds or (-10.0, 10.0)
    prob = P.OptimizationProblem(f(), np.asarray(u0, np.float32), p, lb=np.full(2, lb),
                                 ub=np.full(2, ub))
    sol = P.solve(prob, P.HybridPSO(pso=P.ParallelSyncPSOKernel(64)), maxiters=50,
                  local_maxiters=200, abstol=1e-10, seed=7)
    if ustar is not None:
        np.testing.assert_allclose(sol.u, ustar, atol=1e-3)
    assert sol.objective < 1e-6
    assert (sol.u >= lb).all() and (sol.u <= ub).all()


def test_hybrid_boundary_minimum():
    # test/lbfgs.jl:67-72: (x+1)^2 on [0, 2] -> u = 0, objective = 1
    note("hybrid boundary case")
    prob = P.OptimizationProblem(P.BoundaryQuadratic(), np.array([1.5], np.float32),
                                 lb=[0.0], ub=[2.0])
    sol = P.solve(prob, P.HybridPSO(pso=P.ParallelSyncPSOKernel(64)), maxiters=20,
                  local_maxiters=50, seed=3)
    assert abs(sol.u[0]) < 1e-6 and abs(sol.objective - 1.0) < 1e-6


def test_hybrid_bbob_f8_shape():
    # test/gpu.jl:60-89 (CUDA group): HybridPSO(pso = ParallelSyncPSOKernel(5000))
    # on BBOB F8 in 10-D, bounds [-5, 5], maxiters 150, local_maxiters 50,
    # f - f_opt <= 1e-3.  For D = 10 F8 is Rosenbrock with p = [1, 100] at
    # z = x - x_opt + 1 (its scale max(1, sqrt(D)/8) is 1), i.e. the shifted
    # Rosenbrock with shift = x_opt - 1 and f_opt = 0.  BlackBoxOptimization-
    # Benchmarking's own x_opt (seed 1) is not available here: x_opt is drawn
    # from [-3, 3] (F8's 0.75 * [-4, 4]) by numpy's default_rng(1).
    note("hybrid BBOB F8 shape")
    D = 10
    x_opt = np.random.default_rng(1).uniform(-3, 3, D).astype(np.float32)
    lb, ub = np.full(D, -5.0, np.float32), np.full(D, 5.0, np.float32)
    prob = P.OptimizationProblem(P.Rosenbrock(), np.zeros(D, np.float32), [1.0, 100.0], lb=lb,
                                 ub=ub, shift=x_opt - np.float32(1))
    assert P.eval_f64(prob, x_opt.astype(np.float64)[None, :])[0] < 1e-10   # f(x_opt) = f_opt = 0
    sol = P.solve(prob, P.HybridPSO(pso=P.ParallelSyncPSOKernel(5000)), maxiters=150,
                  local_maxiters=50, abstol=1e-8, reltol=1e-8, seed=42)
    assert np.isfinite(sol.objective) and np.isfinite(sol.u).all()
    assert (sol.u >= lb).all() and (sol.u <= ub).all()
    assert sol.objective <= 1e-3, sol.objective
