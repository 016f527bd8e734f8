"""Pins of the oracle's Strong-Wolfe line search and the L-BFGS that uses it.

The reference's HybridPSO accepts only LineSearch.jl's StrongWolfeLineSearch
(src/hybrid.jl:129-136) and passes it to SimpleOptimization._lbfgs; neither
package is vendored, so the oracle restates Nocedal & Wright's Alg. 3.5/3.6
on the projected path phi(a) = f(P(x + a p)).  Pinned by:
* closed forms: on quadratics the zoom's interpolation lands exactly on the
  minimiser along p (1-D sphere: a = 1/2 from a unit step that overshoots);
  a step into a bound ends on the bound where phi' = 0;
* both strong Wolfe conditions re-evaluated in FP64 at every accepted step of
  random searches on Rosenbrock, Rastrigin, Beale and Himmelblau, and the
  accepted point equal to P(x + a p) bit for bit;
* the known-answer minima of the reference's own test/lbfgs.jl:38-72, which
  the reference runs with this line search;
* FP64 first-order optimality of L-BFGS results on 16-D Rastrigin.
"""
import numpy as np
import pytest

C1, C2 = 1e-4, 0.9


def _P(o, obj, d, par=(), lb=-100.0, ub=100.0):
    return o.Problem(obj, d, np.full(d, lb, np.float32), np.full(d, ub, np.float32),
                     np.asarray(par, np.float32))


def _wolfe(o, **kw):
    return o.lbfgs_params(linesearch=o.LS_STRONG_WOLFE, **kw)


def test_wolfe_sphere_half_step_exact(oracle):
    o = oracle
    P = _P(o, o.SPHERE, 1, (), -10, 10)
    ok, a, xn, fn, gn, f, g = o.wolfe_search(P, [1.0], [-2.0])
    # phi(a) = (1 - 2a)^2: phi(1) = phi(0) fails sufficient decrease, the
    # quadratic through phi(0), phi'(0), phi(1) has its minimum at 1/2
    assert ok and a == np.float32(0.5)
    assert xn[0] == 0.0 and fn == 0.0 and gn[0] == 0.0


def test_wolfe_quadratic_minimiser_along_p(oracle):
    o = oracle
    # f = p1 (x1 - 1)^2 + (x2 + 2)^2: along p = -g the minimiser is
    # a* = g'g / p'Hp with H = diag(2 p1, 2).  With c2 = 0.1 a short first
    # step (0.1, 0.2, 0.4 against a* ~ 0.5) leaves phi' too steep and is
    # doubled until 0.8 overshoots; the zoom's quadratic through phi(0.4),
    # phi'(0.4), phi(0.8) is exact on a quadratic up to FP32 rounding.  (With
    # the default c2 = 0.9, a = 0.1 itself already meets the curvature
    # condition: phi'(0.1) = 0.8 phi'(0).)
    P = _P(o, o.QUADRATIC, 2, (0.01,), -100, 100)
    x = np.array([3.0, -1.0], np.float32)
    g = np.array([2 * 0.01 * (x[0] - 1), 2 * (x[1] + 2)], np.float64)
    p = -g
    H = np.diag([2 * 0.01, 2.0])
    astar = float(g @ g / (p @ H @ p))
    ok, a, xn, fn, gn, f, gg = o.wolfe_search(P, x, p.astype(np.float32),
                                              _wolfe(o, alpha_init=0.1, wolfe_c2=0.1))
    assert ok and abs(float(a) - astar) < 1e-5 * astar
    # curvature condition holds far below the c2 bound: phi'(a) ~ 0
    assert abs(float(np.dot(gn.astype(np.float64), p))) < 1e-5 * abs(float(g @ p))


def test_wolfe_first_trial_accepted_when_it_meets_both_conditions(oracle):
    # the same quadratic with the default c2 = 0.9: the first trial, 0.1,
    # already has phi'(0.1) = 0.8 phi'(0) and sufficient decrease, so it is
    # the step (no further doubling, no zoom)
    P = _P(o := oracle, o.QUADRATIC, 2, (0.01,), -100, 100)
    x = np.array([3.0, -1.0], np.float32)
    g = np.array([2 * 0.01 * (x[0] - 1), 2 * (x[1] + 2)], np.float32)
    ok, a, *_ = o.wolfe_search(P, x, -g, _wolfe(o, alpha_init=0.1))
    assert ok and a == np.float32(0.1)
    # a first trial of 0.04 is too short (phi'(0.04) = 0.92 phi'(0)): doubled once
    ok, a, *_ = o.wolfe_search(P, x, -g, _wolfe(o, alpha_init=0.04))
    assert ok and a == np.float32(0.08)


def _box(o, obj, lb, ub, par=()):
    return o.Problem(obj, len(lb), np.asarray(lb, np.float32), np.asarray(ub, np.float32),
                     np.asarray(par, np.float32))


def test_wolfe_derivative_counts_only_moving_lanes(oracle):
    o = oracle
    # sphere on [1, 10] x [-10, 10] from (1.5, 4) along -g = (-3, -8): x1
    # reaches its bound at a = 1/6 and stops; beyond it phi'(a) is x2's term
    # alone, zero at a = 1/2.  With c2 = 0.01 the step must land within
    # |a - 1/2| <= 0.005; counting x1's frozen gradient (2 * 1 * -3) would
    # move the zero of phi' to a ~ 0.547.
    P = _box(o, o.SPHERE, [1.0, -10.0], [10.0, 10.0])
    ok, a, xn, *_ = o.wolfe_search(P, [1.5, 4.0], [-3.0, -8.0], _wolfe(o, wolfe_c2=0.01))
    assert ok and abs(float(a) - 0.5) <= 0.005 and xn[0] == 1.0


def test_wolfe_initial_slope_skips_a_lane_pushing_out_of_the_box(oracle):
    o = oracle
    # from (1, 4) on the bound x1 = 1 along (-2, -8): x1 cannot move, so
    # phi'(0) = 2 * 4 * -8 = -64 (not -68).  With alpha_init = 0.04:
    # |phi'(0.04)| = 58.88 > 0.9 * 64 = 57.6 (rejected; it would pass the
    # bound 0.9 * 68 = 61.2), and the doubled step 0.08 is accepted.
    P = _box(o, o.SPHERE, [1.0, -10.0], [10.0, 10.0])
    ok, a, *_ = o.wolfe_search(P, [1.0, 4.0], [-2.0, -8.0], _wolfe(o, alpha_init=0.04))
    assert ok and a == np.float32(0.08)


def test_wolfe_step_into_a_bound(oracle):
    o = oracle
    # (x + 1)^2 on [0, 2] from 1.5 along -g = -5: the unit step ends on the
    # bound 0 (P clips -3.5), where the path is flat: phi'(1) = 0 is accepted
    P = _P(o, o.BOUNDARY_QUADRATIC, 1, (), 0.0, 2.0)
    ok, a, xn, fn, gn, f, g = o.wolfe_search(P, [1.5], [-5.0])
    assert ok and a == 1.0 and xn[0] == 0.0 and fn == 1.0


def test_wolfe_refuses_an_ascent_direction(oracle):
    o = oracle
    P = _P(o, o.SPHERE, 2, (), -10, 10)
    ok, *_ = o.wolfe_search(P, [1.0, 1.0], [1.0, 0.0])
    assert not ok


def _fd64(o, P, name, x):
    """FP64 value and gradient of the test objectives."""
    x = np.asarray(x, np.float64)
    if name == "ROSENBROCK":
        f = np.sum(100 * (x[1:] - x[:-1] ** 2) ** 2 + (1 - x[:-1]) ** 2)
        g = np.zeros_like(x)
        g[:-1] += -400 * x[:-1] * (x[1:] - x[:-1] ** 2) - 2 * (1 - x[:-1])
        g[1:] += 200 * (x[1:] - x[:-1] ** 2)
    elif name == "RASTRIGIN":
        f = 10 * len(x) + np.sum(x * x - 10 * np.cos(2 * np.pi * x))
        g = 2 * x + 20 * np.pi * np.sin(2 * np.pi * x)
    elif name == "BEALE":
        a, b = x
        t = [(1.5 - a + a * b), (2.25 - a + a * b * b), (2.625 - a + a * b ** 3)]
        f = sum(v * v for v in t)
        g = np.array([2 * t[0] * (b - 1) + 2 * t[1] * (b * b - 1) + 2 * t[2] * (b ** 3 - 1),
                      2 * t[0] * a + 2 * t[1] * 2 * a * b + 2 * t[2] * 3 * a * b * b])
    else:  # HIMMELBLAU
        a, b = x
        u, v = a * a + b - 11, a + b * b - 7
        f = u * u + v * v
        g = np.array([4 * a * u + 2 * v, 2 * u + 4 * b * v])
    return float(f), g


@pytest.mark.parametrize("name,d,par,lb,ub", [("ROSENBROCK", 6, (1, 100), -2, 2),
                                              ("RASTRIGIN", 5, (), -5.12, 5.12),
                                              ("BEALE", 2, (), -4.5, 4.5),
                                              ("HIMMELBLAU", 2, (), -5, 5)])
def test_wolfe_conditions_hold_in_fp64(oracle, name, d, par, lb, ub):
    o = oracle
    P = _P(o, getattr(o, name), d, par, lb, ub)
    rng = np.random.default_rng(len(name) + d)
    accepted = 0
    for trial in range(60):
        x = rng.uniform(lb, ub, d).astype(np.float32)
        f32, g32 = o.value_grad(P, x)
        # steepest descent, or a random descent direction of random length
        p = -g32 if trial % 2 == 0 else rng.normal(size=d).astype(np.float32)
        if float(np.dot(p, g32)) > 0:
            p = -p
        p = (p * np.float32(rng.uniform(0.01, 2.0) / (np.linalg.norm(p) + 1e-30))).astype(
            np.float32)
        ok, a, xn, fn, gn, f, g = o.wolfe_search(P, x, p, _wolfe(o))
        if not ok:
            continue
        # the accepted point is the projected step, bit for bit
        want = np.clip((x + (a * p).astype(np.float32)).astype(np.float32), lb, ub)
        np.testing.assert_array_equal(xn, want.astype(np.float32))
        # phi'(0) and phi'(a) in FP64 with the oracle's free-lane rule
        f0, g0 = _fd64(o, P, name, x)
        fa, ga = _fd64(o, P, name, xn)
        free0 = ~(((x <= lb) & (p < 0)) | ((x >= ub) & (p > 0)))
        u = (x + a * p).astype(np.float32)
        freea = (u > lb) & (u < ub)
        d0 = float(np.sum(np.where(free0, g0 * p, 0.0)))
        da = float(np.sum(np.where(freea, ga * p, 0.0)))
        scale = abs(f0) + 1.0
        assert fa <= f0 + C1 * float(a) * d0 + 2e-6 * scale, (trial, fa, f0, d0, float(a))
        # every accepted step of these smooth problems meets the curvature
        # condition too (up to FP32 gradient rounding): the zoom keeps a
        # bracket that holds a strong-Wolfe step
        assert abs(da) <= -C2 * d0 + 1e-4 * (np.abs(ga).max() + 1) * np.abs(p).sum(), trial
        accepted += 1
    assert accepted >= 50


KNOWN = [
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
def test_wolfe_lbfgs_known_answers(oracle, name, par, bounds, x0, ustar, fstar):
    # test/lbfgs.jl:38-72 (local_maxiters 200, abstol 1e-10), FP32 tolerances
    o = oracle
    d = len(x0)
    P = _P(o, getattr(o, name), d, par, *bounds)
    x, f = o.lbfgs_one(P, x0, _wolfe(o, maxiters=200, history=10, abstol=1e-10))
    if ustar is not None:
        np.testing.assert_allclose(x, ustar, atol=1e-3)
    assert abs(float(f) - fstar) < 1e-6
    assert (x >= bounds[0]).all() and (x <= bounds[1]).all()


def test_wolfe_lbfgs_first_order_optimality_fp64(oracle):
    o = oracle
    rng = np.random.default_rng(5)
    P = _P(o, o.RASTRIGIN, 16, (), -5.12, 5.12)
    for _ in range(20):
        x0 = rng.uniform(-5.12, 5.12, 16)
        x, f = o.lbfgs_one(P, x0, _wolfe(o, maxiters=100, history=10))
        g = 2 * x.astype(np.float64) + 20 * np.pi * np.sin(2 * np.pi * x.astype(np.float64))
        assert float(np.max(np.abs(np.clip(x - g, -5.12, 5.12) - x))) < 0.08
        assert f <= o.objective(P, x0.astype(np.float32)) + 1e-4
