"""Pins of the CPU oracle (oracle/pps_oracle.c) before it is trusted as the
checker of the HIP kernels.

Each check is chosen so that a plausible mistake in the oracle (wrong Philox
round, dropped velocity term, swapped operand, wrong clamp semantics, wrong
argmin tie rule, wrong RK4 weights) fails it:

* Philox4x32-10 against the published Random123 known-answer vectors;
* cos(2πx) against FP64 cos and its exact special values;
* objectives against closed-form minima (incl. the minima of the reference's
  test/lbfgs.jl:38-72) and FP64 evaluations;
* the Lotka-Volterra loss against the exact value of the reference's
  _reduce_losses! unit test (test/regression.jl:345-357) and RK4 against the
  fourth-order amplification factor of u' = λu;
* one velocity/position move against an FP64 recomputation from the same
  random draws, and Julia's NaN-propagating max/min;
* run-level invariants and the reference's own convergence thresholds
  (test/regression.jl:69,87; test/gpu.jl:35,47,57).
"""
import math

import numpy as np
import pytest


# -- Philox ------------------------------------------------------------------
# Known-answer vectors of Philox4x32-10 from the Random123 distribution
# (Salmon et al., SC'11): (counter, key) -> output.
PHILOX_KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,out", PHILOX_KAT)
def test_philox_known_answers(oracle, ctr, key, out):
    assert oracle.philox(ctr, key) == out


# -- cos(2 pi x) ----------------------------------------------------------------
def test_cos2pi_accuracy_and_exact_points(oracle):
    rng = np.random.default_rng(0)
    xs = np.concatenate([np.linspace(-6, 6, 20001), rng.uniform(-100, 100, 5000)]).astype(np.float32)
    err = max(abs(oracle.cos2pi(float(x)) - math.cos(2 * math.pi * float(x))) for x in xs)
    assert err < 3e-7
    for x, v in [(0.0, 1.0), (1.0, 1.0), (-3.0, 1.0), (0.5, -1.0), (2.5, -1.0)]:
        assert oracle.cos2pi(x) == v
    assert abs(oracle.cos2pi(0.25)) < 1e-7 and abs(oracle.cos2pi(-0.75)) < 1e-7
    # even and 1-periodic, bit for bit (the reduction x - rint(x) is exact);
    # dyadic points so that x + 4 is itself exact in FP32
    for k in rng.integers(-4096, 4096, 200):
        x = float(k) / 1024.0
        assert oracle.cos2pi(x) == oracle.cos2pi(-x)
        assert oracle.cos2pi(x) == oracle.cos2pi(x + 4.0)
    s_err = max(abs(oracle.sin2pi(float(x)) - math.sin(2 * math.pi * float(x))) for x in xs[::5])
    assert s_err < 3e-7


# -- objectives --------------------------------------------------------------
def _prob(oracle, obj, d, params=(), lb=-10, ub=10, shift=None):
    return oracle.Problem(obj, d, np.full(d, lb, np.float32), np.full(d, ub, np.float32),
                          np.asarray(params, np.float32), shift)


def test_objective_closed_forms(oracle):
    o = oracle
    R = _prob(o, o.ROSENBROCK, 3, (1, 100))
    assert o.objective(R, [1, 1, 1]) == 0.0
    assert o.objective(R, [0, 0, 0]) == 2.0  # 1 + 1 at the reference's x0 = zeros
    A = _prob(o, o.RASTRIGIN, 10)
    assert o.objective(A, np.zeros(10)) == 0.0
    shift = np.linspace(-1, 1, 10).astype(np.float32)
    As = _prob(o, o.RASTRIGIN, 10, shift=shift)
    assert abs(o.objective(As, shift)) < 1e-5
    # minima of the reference's L-BFGS known-answer tests (test/lbfgs.jl:38-72)
    assert o.objective(_prob(o, o.BEALE, 2), [3, 0.5]) == 0.0
    assert o.objective(_prob(o, o.BOOTH, 2), [1, 3]) == 0.0
    assert o.objective(_prob(o, o.HIMMELBLAU, 2), [3, 2]) == 0.0
    assert o.objective(_prob(o, o.QUADRATIC, 2, (1000,)), [1, -2]) == 0.0
    assert o.objective(_prob(o, o.BOUNDARY_QUADRATIC, 1), [0.0]) == 1.0
    assert o.objective(_prob(o, o.ROSEN2, 2, (1, 100)), [1, 1]) == 0.0
    assert o.objective(_prob(o, o.SPHERE, 5), np.ones(5)) == 5.0


def _f64(obj, x, p):
    x = np.asarray(x, np.float64)
    if obj == 1:
        return sum(p[1] * (x[i + 1] - x[i] ** 2) ** 2 + (p[0] - x[i]) ** 2 for i in range(len(x) - 1))
    if obj == 2:
        return 10 * len(x) + sum(x ** 2 - 10 * np.cos(2 * np.pi * x))
    if obj == 3:
        return float(np.sum(x ** 2))
    raise ValueError(obj)


@pytest.mark.parametrize("obj,d", [(1, 2), (1, 3), (1, 10), (1, 32), (2, 3), (2, 10), (2, 16), (3, 7)])
def test_objectives_against_fp64(oracle, obj, d):
    rng = np.random.default_rng(obj * 100 + d)
    P = _prob(oracle, obj, d, (1, 100))
    xs = rng.uniform(-5, 5, (200, d)).astype(np.float32)
    got = oracle.evaluate(P, xs)
    ref = np.array([_f64(obj, x, (1.0, 100.0)) for x in xs])
    np.testing.assert_allclose(got, ref, rtol=2e-5, atol=2e-4)


# -- Lotka-Volterra loss / RK4 -------------------------------------------------
def _lv(oracle, obs, theta, u0=(0.0, 0.0), dt=0.01, spo=10):
    obs = np.asarray(obs, np.float32).reshape(-1, 2)
    P = oracle.Problem(oracle.LOTKA_VOLTERRA, 4, np.zeros(4, np.float32), np.full(4, 5, np.float32),
                       np.array([u0[0], u0[1], dt, spo], np.float32), obs=obs)
    return oracle.objective(P, np.asarray(theta, np.float32))


def test_lv_loss_matches_reduce_losses_unit_test(oracle):
    # test/regression.jl:345-357: data [1, 2] at m = 5 samples against a zero
    # trajectory gives 5 * (1 + 4) = 25 for every particle, exactly.
    assert _lv(oracle, [[1, 2]] * 5, [0, 0, 0, 0]) == 25.0


def test_rk4_fourth_order_factor(oracle):
    # with β = γ = δ = 0: u1' = α u1, u2' = 0.  One RK4 step multiplies u1 by
    # 1 + z + z²/2 + z³/6 + z⁴/24 (z = α dt); read it back through the loss
    # against zero observations: SSE = u1(0)^2 + u2^2 + u1(1)^2 + u2^2.
    for alpha, dt in [(1.5, 0.01), (-2.0, 0.1), (3.0, 0.05)]:
        z = alpha * dt
        g = 1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24
        sse = _lv(oracle, [[0, 0], [0, 0]], [alpha, 0, 0, 0], u0=(1.0, 0.0), dt=dt, spo=1)
        assert abs(float(sse) - (1 + g * g)) < 5e-6 * (1 + g * g)


def _lv_f64_traj(theta, u0, dt, n_obs, spo):
    a, b, c, d = theta
    f = lambda u: np.array([a * u[0] - b * u[0] * u[1], -c * u[1] + d * u[0] * u[1]])
    u = np.array(u0, np.float64)
    out = [u.copy()]
    for _ in range(n_obs - 1):
        for _ in range(spo):
            k1 = f(u); k2 = f(u + dt / 2 * k1); k3 = f(u + dt / 2 * k2); k4 = f(u + dt * k3)
            u = u + dt / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
        out.append(u.copy())
    return np.array(out)


def test_lv_sse_against_fp64_trajectory(oracle):
    theta = (1.5, 1.0, 3.0, 1.0)
    traj = _lv_f64_traj(theta, (1.0, 1.0), 0.01, 101, 10)
    rng = np.random.default_rng(7)
    obs = traj + rng.normal(0, 0.05, traj.shape)
    sse64 = float(np.sum((obs - traj) ** 2))
    sse32 = float(_lv(oracle, obs, theta, u0=(1.0, 1.0)))
    assert abs(sse32 - sse64) < 1e-3 * sse64
    # first integral of LV: V = δu1 − γ ln u1 + βu2 − α ln u2 is conserved by
    # the flow; FP64 RK4 at dt=0.01 drifts by O(dt^4) only (pins the trajectory
    # the SSE is computed on)
    V = lambda u: theta[3] * u[0] - theta[2] * np.log(u[0]) + theta[1] * u[1] - theta[0] * np.log(u[1])
    assert abs(V(traj[-1]) - V(traj[0])) < 1e-6


# -- one move: FP64 recomputation from the same draws --------------------------
def _u01(u):
    return (u >> 8) * 2.0 ** -24


@pytest.mark.parametrize("d", [1, 2, 3, 5, 10])
def test_move_against_fp64_formula(oracle, d):
    rng = np.random.default_rng(d)
    lb, ub = -2.0, 3.0
    P = _prob(oracle, oracle.ROSENBROCK, d, (1, 100), lb=lb, ub=ub)
    seed, epoch = 0x1234_5678_9ABC_DEF0, 3
    key = [seed & 0xFFFFFFFF, seed >> 32]
    for trial in range(20):
        x = rng.uniform(lb, ub, d).astype(np.float32)
        v = rng.uniform(-1, 1, d).astype(np.float32)
        pb = rng.uniform(lb, ub, d).astype(np.float32)
        gb = rng.uniform(lb, ub, d).astype(np.float32)
        part = np.concatenate([x, v, [0.0], pb, [1.0]]).astype(np.float32)
        i, t = 17 + trial, 5 + trial
        w, c1, c2 = 0.7298, 1.4962, 1.4962
        out = oracle.move(P, part, gb, w, c1, c2, i, t, epoch, seed)
        for j in range(d):
            r = oracle.philox([i, t, j // 2, epoch], key)
            r1, r2 = (_u01(r[0]), _u01(r[1])) if j % 2 == 0 else (_u01(r[2]), _u01(r[3]))
            vn = w * v[j] + c1 * r1 * (pb[j] - x[j]) + c2 * r2 * (gb[j] - x[j])
            xn = min(max(x[j] + vn, lb), ub)
            assert abs(out[d + j] - vn) < 1e-5 * (1 + abs(vn))
            assert abs(out[j] - xn) < 1e-5 * (1 + abs(xn))
        # best position and cost untouched by a move
        np.testing.assert_array_equal(out[2 * d + 1: 3 * d + 1], pb)


def test_move_clamp_is_nan_propagating(oracle):
    # Julia's max/min propagate NaN (a NaN position stays NaN; maxNum would
    # have returned the bound).
    P = _prob(oracle, oracle.ROSENBROCK, 2, (1, 100), lb=-1, ub=1)
    part = np.array([0, 0, np.nan, 0, 0, 0, 0, 1], np.float32)
    out = oracle.move(P, part, [0, 0], 0.5, 1.4962, 1.4962, 0, 1, 0, 1)
    assert np.isnan(out[0]) and not np.isnan(out[1]) and -1 <= out[1] <= 1
    part = np.array([0, 0, np.inf, -np.inf, 0, 0, 0, 1], np.float32)
    out = oracle.move(P, part, [0, 0], 0.5, 1.4962, 1.4962, 0, 1, 0, 1)
    assert out[0] == 1.0 and out[1] == -1.0


def test_lexicographic_clamp(oracle):
    # max(x, lb) / min(x, ub) on whole vectors (src/ode_pso.jl:18-19): only the
    # first differing component decides.
    P = oracle.Problem(oracle.SPHERE, 3, np.zeros(3, np.float32), np.full(3, 5, np.float32),
                       np.zeros(0, np.float32), clamp=oracle.CLAMP_LEXICOGRAPHIC)
    z = np.zeros(3, np.float32)
    for x, want in [([1, -3, 9], [1, -3, 9]),   # x > lb lexicographically, x < ub
                    ([-1, 3, 3], [0, 0, 0]),    # x < lb -> lb
                    ([6, 0, 0], [5, 5, 5]),     # x > ub -> ub
                    ([0, -1, 7], [0, 0, 0]),    # tie on x1, x2 < 0 -> lb
                    ([5, 6, -1], [5, 5, 5])]:   # tie with ub on x1, x2 > 5 -> ub
        part = np.concatenate([x, z, [0], z, [1]]).astype(np.float32)
        out = oracle.move(P, part, z, 0.0, 0.0, 0.0, 0, 1, 0, 1)
        np.testing.assert_array_equal(out[:3], np.asarray(want, np.float32))


# -- run-level invariants ----------------------------------------------------
def test_init_layout_and_bounds(oracle):
    P = _prob(oracle, oracle.RASTRIGIN, 5, lb=-5.12, ub=5.12)
    s = oracle.Swarm(P, 3000, seed=99)
    assert (s.x >= -5.12).all() and (s.x <= 5.12).all()
    assert (s.v == 0).all()
    np.testing.assert_array_equal(s.best, s.x)
    np.testing.assert_array_equal(s.best_cost, oracle.evaluate(P, s.x.T))
    pos, cost, idx = s.gbest()
    assert idx == int(np.argmin(s.best_cost)) and cost == s.best_cost.min()
    np.testing.assert_array_equal(pos, s.best[:, idx])
    # uniform: mean of all coordinates near 0, spread near (b-a)^2/12
    assert abs(s.x.mean()) < 0.05 and abs(s.x.var() - 10.24 ** 2 / 12) < 0.2
    s.reinit()
    assert not np.array_equal(s.x, oracle.Swarm(P, 3000, seed=99).x)


@pytest.mark.parametrize("gf", [0, 1])
def test_sync_invariants(oracle, gf):
    P = _prob(oracle, oracle.ROSENBROCK, 3, (1, 100), lb=-1, ub=10)
    s = oracle.Swarm(P, 500, seed=5)
    prev_best = s.best_cost
    prev_g = s.gbest()[1]
    for _ in range(30):
        s.run_sync(1, gbest_from=gf)
        bc = s.best_cost
        assert (bc <= prev_best).all()
        assert (s.x >= -1).all() and (s.x <= 10).all()
        pos, cost, idx = s.gbest()
        assert idx == int(np.argmin(bc)) and cost == bc.min() and cost <= prev_g
        np.testing.assert_array_equal(pos, (s.x if gf == 1 else s.best)[:, idx])
        prev_best, prev_g = bc, cost


def test_zero_coefficients_freeze_the_swarm(oracle):
    P = _prob(oracle, oracle.SPHERE, 4, lb=-1, ub=1)
    s = oracle.Swarm(P, 100, seed=1)
    x0 = s.x
    s.run_sync(10, w=0.0, c1=0.0, c2=0.0)
    np.testing.assert_array_equal(s.x, x0)
    s.run_async(10, w=0.0, c1=0.0, c2=0.0)
    np.testing.assert_array_equal(s.x, x0)


def test_argmin_ties_lowest_index(oracle):
    # every particle starts at the same corner -> identical costs
    P = _prob(oracle, oracle.SPHERE, 2, lb=1, ub=1)
    s = oracle.Swarm(P, 64, seed=3)
    assert s.gbest()[2] == 0
    s.run_sync(3)
    assert s.gbest()[2] == 0


# -- the reference's own convergence thresholds --------------------------------
ROSEN_LB, ROSEN_UB = -1.0, 10.0


def _rosen(oracle, d):
    return _prob(oracle, oracle.ROSENBROCK, d, (1, 100), lb=ROSEN_LB, ub=ROSEN_UB)


@pytest.mark.parametrize("d", [2, 3])
def test_reference_threshold_sync_cpu(oracle, d):
    # test/regression.jl:80-87: ParallelSyncPSOKernel(2000; CPU), 500 iters < 3e-3
    s = oracle.Swarm(_rosen(oracle, d), 2000, seed=123 + d)
    s.run_sync(500, gbest_from=oracle.GBEST_CURRENT)
    assert s.gbest()[1] < 3e-3


@pytest.mark.parametrize("d", [2, 3])
def test_reference_threshold_qlock_cpu(oracle, d):
    # test/regression.jl:63-69: ParallelPSOKernel(2000; CPU), 500 iters < 1e-4
    s = oracle.Swarm(_rosen(oracle, d), 2000, seed=223 + d)
    s.run_qlock(500)
    assert s.gbest()[1] < 1e-4


@pytest.mark.parametrize("d", [2, 3, 4])
def test_reference_thresholds_gpu_suite(oracle, d):
    # test/gpu.jl:31-57 with 5000 particles: queue-lock 500 iters < 2e-3, async
    # 1000 iters improves on f(x0), sync 500 iters < 6e-4
    P = _rosen(oracle, d)
    f_x0 = oracle.objective(P, np.zeros(d))
    s = oracle.Swarm(P, 5000, seed=1234 + d)
    s.run_qlock(500)
    assert s.gbest()[1] < 2e-3 and s.gbest()[1] < f_x0
    s = oracle.Swarm(P, 5000, seed=2234 + d)
    s.run_async(1000)
    assert s.gbest()[1] < f_x0
    s = oracle.Swarm(P, 5000, seed=3234 + d)
    s.run_sync(500)
    assert s.gbest()[1] < 6e-4


def test_lv_fixture_loss_at_true_parameters(oracle):
    # the committed observations (tests/golden/lv_obs.json): FP32 in-kernel
    # RK4 loss at the true parameters equals the FP64 noise energy
    import json
    import os

    doc = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "lv_obs.json")))
    obs = np.array(doc["obs"])
    traj = np.array(doc["trajectory"])
    sse64 = float(np.sum((obs - traj) ** 2))
    sse32 = float(_lv(oracle, obs, doc["theta"], u0=doc["u0"], dt=doc["dt"], spo=doc["steps_per_obs"]))
    assert abs(sse32 - sse64) < 1e-3 * sse64
    # and the noise has the stated scale
    assert abs(np.std(obs - traj) - doc["sigma"]) < 0.005


def _bits(a):
    a = np.asarray(a, np.float32)
    return np.where(np.isnan(a), np.uint32(0x7FC00000), a.view(np.uint32))


def test_folded_draw_scale_equals_the_scaled_draw():
    # The kernels compute the attraction coefficient c * u01(u) as
    # (c * 2^-24) * m with m = the draw's top 24 bits as a float; the oracle
    # computes c * (m * 2^-24).  Both round the same real number once when
    # c * 2^-24 is exact: c = 0, c not finite, or |c| >= 2^-102.
    rng = np.random.default_rng(5)
    m = np.concatenate([rng.integers(0, 1 << 24, 20000), [0, 1, 2, (1 << 24) - 1]])
    m = m.astype(np.float32)
    s = np.float32(2.0 ** -24)
    allowed = [1.4962, -1.4962, 2.0 ** -102, -(2.0 ** -102), 1.3 * 2.0 ** -100, 1e-20, 7.0,
               1e30, 3.4e38, 0.0, -0.0, np.inf, -np.inf, np.nan]
    for c in allowed:
        c = np.float32(c)
        with np.errstate(over="ignore", invalid="ignore"):  # c = 3.4e38 overflows, inf * 0
            folded = np.float32(c * s) * m
            scaled = c * (m * s)
        np.testing.assert_array_equal(_bits(folded), _bits(scaled), err_msg=f"c = {c}")
    # below the limit c * 2^-24 is subnormal and loses bits: the forms differ,
    # which is why the library refuses such coefficients
    c = np.float32(1.7 * 2.0 ** -110)
    assert np.any(_bits(np.float32(c * s) * m) != _bits(c * (m * s)))
