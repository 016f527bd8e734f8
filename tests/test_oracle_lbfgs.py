"""Pins of the oracle's projected L-BFGS (HybridPSO local phase).

SimpleOptimization._lbfgs, which the reference calls (src/lbfgs.jl:15-18),
is not vendored, so its iterates are not available: the oracle's restatement
is pinned by
* the known-answer minima of the reference's test/lbfgs.jl:38-72 (Rosenbrock
  from two starts, bounded Rosenbrock, Beale, Booth, Himmelblau, the bounded
  quadratic with p = 1000, and the 1-D case whose minimum sits on the bound),
* first-order optimality measured independently in FP64 (projected gradient of
  the analytic FP64 derivative) at the returned points of 16-D Rastrigin and
  10-D Rosenbrock — a wrong gradient term cannot satisfy it,
* whole-swarm runs equal to single-particle runs, and the degenerate cases
  (zero iterations; the unit-curvature quadratic whose half step is exact).
Tolerances are for FP32 arithmetic (the reference test runs in FP64 with
abstol 1e-10): positions within 1e-3, objectives below 1e-6.
"""
import numpy as np
import pytest


def _P(o, obj, d, par=(), lb=-100.0, ub=100.0):
    return o.Problem(obj, d, np.full(d, lb, np.float32), np.full(d, ub, np.float32),
                     np.asarray(par, np.float32))


KNOWN = [
    # (objective, params, bounds, x0, u*, f*)
    ("ROSEN2", (1, 100), (-100, 100), [0.0, 0.0], [1, 1], 0.0),
    ("ROSEN2", (1, 100), (-100, 100), [-1.2, 1.0], [1, 1], 0.0),
    ("ROSEN2", (1, 100), (-2, 2), [0.0, 0.0], [1, 1], 0.0),
    ("BEALE", (), (-100, 100), [1.0, 1.0], [3, 0.5], 0.0),
    ("BOOTH", (), (-100, 100), [1.0, 1.0], [1, 3], 0.0),
    ("HIMMELBLAU", (), (-100, 100), [1.0, 1.0], None, 0.0),
    ("QUADRATIC", (1000,), (-5, 5), [4.0, 4.0], [1, -2], 0.0),
    ("BOUNDARY_QUADRATIC", (), (0, 2), [1.5], [0.0], 1.0),
]


@pytest.mark.parametrize("name,par,bounds,x0,ustar,fstar", KNOWN)
def test_lbfgs_known_answers(oracle, name, par, bounds, x0, ustar, fstar):
    o = oracle
    d = len(x0)
    P = _P(o, getattr(o, name), d, par, *bounds)
    x, f = o.lbfgs_one(P, x0, o.lbfgs_params(maxiters=200, history=10, abstol=1e-10))
    if ustar is not None:
        np.testing.assert_allclose(x, ustar, atol=1e-3)
    assert abs(float(f) - fstar) < 1e-6
    assert (x >= bounds[0]).all() and (x <= bounds[1]).all()


def _rastrigin_grad64(x):
    x = np.asarray(x, np.float64)
    return 2 * x + 20 * np.pi * np.sin(2 * np.pi * x)


def _rosen_grad64(x, p1=1.0, p2=100.0):
    x = np.asarray(x, np.float64)
    g = np.zeros_like(x)
    g[:-1] += -4 * p2 * x[:-1] * (x[1:] - x[:-1] ** 2) - 2 * (p1 - x[:-1])
    g[1:] += 2 * p2 * (x[1:] - x[:-1] ** 2)
    return g


def _proj_grad_norm(x, g, lb, ub):
    return float(np.max(np.abs(np.clip(x - g, lb, ub) - x)))


def test_lbfgs_first_order_optimality_fp64(oracle):
    o = oracle
    rng = np.random.default_rng(3)
    P = _P(o, o.RASTRIGIN, 16, (), -5.12, 5.12)
    for _ in range(20):
        x0 = rng.uniform(-5.12, 5.12, 16)
        x, f = o.lbfgs_one(P, x0, o.lbfgs_params(maxiters=100, history=10))
        g = _rastrigin_grad64(x)
        # FP32 resolves f ~ 16 only to ~2e-6, i.e. x to ~1e-4 around a local
        # minimum of curvature ~400: |grad| up to ~0.04 is the FP32 floor,
        # while a wrong gradient term leaves O(1) FP64 gradients
        assert _proj_grad_norm(x, g, -5.12, 5.12) < 0.08
        assert f <= o.objective(P, x0.astype(np.float32)) + 1e-4
    P = _P(o, o.ROSENBROCK, 10, (1, 100), -5, 5)
    for _ in range(5):
        x0 = rng.uniform(-2, 2, 10)
        x, f = o.lbfgs_one(P, x0, o.lbfgs_params(maxiters=500, history=10))
        g = _rosen_grad64(x)
        assert _proj_grad_norm(x, g, -5, 5) < 1e-3
        assert f < o.objective(P, x0.astype(np.float32))


def test_lbfgs_swarm_matches_single_runs(oracle):
    o = oracle
    P = _P(o, o.RASTRIGIN, 16, (), -5.12, 5.12)
    s = o.Swarm(P, 64, seed=12)
    s.run_sync(5)
    lp = o.lbfgs_params(maxiters=30, history=10)
    xs, fs = o.lbfgs_swarm(s, lp)
    best = s.best
    for i in (0, 17, 63):
        x, f = o.lbfgs_one(P, best[:, i], lp)
        np.testing.assert_array_equal(x, xs[i])
        assert f == fs[i]


def test_lbfgs_history_one_and_zero_iterations(oracle):
    o = oracle
    P = _P(o, o.SPHERE, 4, (), -1, 1)
    x0 = np.array([0.5, -0.25, 0.75, 0.1], np.float32)
    x, f = o.lbfgs_one(P, x0, o.lbfgs_params(maxiters=0))
    np.testing.assert_array_equal(x, x0)
    # a quadratic with unit curvature: the first steepest-descent step of
    # length 1 overshoots to -x0 (f unchanged, Armijo fails), half a step lands
    # exactly on the minimum
    x, f = o.lbfgs_one(P, x0, o.lbfgs_params(maxiters=1, history=1))
    np.testing.assert_array_equal(x, np.zeros(4, np.float32))
    assert f == 0.0


@pytest.mark.parametrize("ls", [0, 1])
def test_lbfgs_reltol_stops_at_the_first_small_relative_decrease(oracle, ls):
    # the run with reltol r equals the plain run cut at the first iteration k
    # whose step decreased f by at most r |f| (plain runs of increasing
    # maxiters are prefixes of one another)
    o = oracle
    P = _P(o, o.ROSENBROCK, 4, (1, 100), -5, 5)
    x0 = np.array([-1.2, 1.0, -0.5, 0.8], np.float32)
    fk = [o.lbfgs_one(P, x0, o.lbfgs_params(maxiters=k, history=10, linesearch=ls))[1]
          for k in range(0, 80)]
    for r in (1e-2, 1e-4):
        K = next(k for k in range(1, 80) if abs(fk[k - 1] - fk[k]) <= np.float32(r) * abs(fk[k - 1]))
        assert K > 1
        x, f = o.lbfgs_one(P, x0, o.lbfgs_params(maxiters=500, history=10, linesearch=ls,
                                                reltol=r))
        xk, f_k = o.lbfgs_one(P, x0, o.lbfgs_params(maxiters=K, history=10, linesearch=ls))
        np.testing.assert_array_equal(x, xk)
        assert f == f_k
