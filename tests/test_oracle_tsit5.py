"""Pins of the oracle's Tsit5 Lotka-Volterra solver (CPU).

The reference's default ode_alg is DiffEqGPU's GPUTsit5 (src/ode_pso.jl:54-170),
not vendored; parity with it is unpinned.  What pins the restatement:
* the tableau as stored satisfies the Runge-Kutta order conditions: all 17 of
  order <= 5 for the solution weights, all 8 of order <= 4 (and not all of
  order 5) for the embedded weights — a wrong digit breaks them;
* fixed steps converge at 5th order to an FP64 reference trajectory;
* root5 (the controller's 5th root from IEEE basic operations) is within 2 ulp;
* the adaptive controller meets its tolerances: the error falls as they tighten,
  and at tight tolerances it reaches the FP32 floor;
* the FP64 evaluator (pps_eval_f64, product host code) agrees with the FP32
  oracle on noisy data.
Saveat semantics (samples read from Tsit5's continuous extension between steps):
* the extension's weights satisfy the 8 continuous order conditions of order
  <= 4 at every theta, reduce to b at theta = 1 and vanish at theta = 0;
* with samples on step ends it reproduces step-end sampling;
* with steps 2-8 times the sample spacing it still converges at 5th order (FP32
  oracle and FP64 evaluator), and the adaptive form meets its tolerances.
"""
import ctypes as C

import numpy as np
import pytest

import pps_oracle as O

F = np.float32
C_NODES = np.array([0, 0.161, 0.327, 0.9, 0.9800255409045097, 1, 1])


def _conditions(A, b, c, order):
    e = np.ones(7)
    out = [(b @ e, 1.0)]
    if order >= 2:
        out.append((b @ c, 1 / 2))
    if order >= 3:
        out += [(b @ c ** 2, 1 / 3), (b @ A @ c, 1 / 6)]
    if order >= 4:
        out += [(b @ c ** 3, 1 / 4), (b @ (c * (A @ c)), 1 / 8), (b @ A @ c ** 2, 1 / 12),
                (b @ A @ A @ c, 1 / 24)]
    if order >= 5:
        out += [(b @ c ** 4, 1 / 5), (b @ (c ** 2 * (A @ c)), 1 / 10), (b @ ((A @ c) ** 2), 1 / 20),
                (b @ (c * (A @ c ** 2)), 1 / 15), (b @ (c * (A @ A @ c)), 1 / 30),
                (b @ A @ c ** 3, 1 / 20), (b @ A @ (c * (A @ c)), 1 / 40),
                (b @ A @ A @ c ** 2, 1 / 60), (b @ A @ A @ A @ c, 1 / 120)]
    return np.array([abs(x - y) for x, y in out])


def test_tsit5_tableau_order_conditions():
    a, e = O.tsit5_tableau()
    A = np.zeros((7, 7))
    A[:, :6] = a.astype(np.float64)
    c = A.sum(axis=1)
    np.testing.assert_allclose(c, C_NODES, atol=2e-6)
    b = A[6].copy()
    bhat = b - e.astype(np.float64)
    assert _conditions(A, b, c, 5).max() < 2e-6              # 17 conditions, order 5
    assert len(_conditions(A, b, c, 5)) == 17
    assert _conditions(A, bhat, c, 4).max() < 2e-6           # embedded: order 4
    assert _conditions(A, bhat, c, 5).max() > 1e-4           # ... and not order 5


def test_root5():
    assert O.root5(0.0) == 0 and O.root5(-1.0) == 0 and O.root5(np.nan) == 0
    assert O.root5(1.0) == 1 and O.root5(32.0) == 2 and O.root5(np.inf) == np.inf
    rng = np.random.default_rng(0)
    xs = np.exp(rng.uniform(np.log(1e-30), np.log(1e30), 5000)).astype(F)
    for x in xs:
        ref = np.float64(x) ** 0.2
        assert abs(float(O.root5(x)) - ref) <= 2 * np.spacing(np.float32(ref))


THETA = np.array([1.5, 1.0, 3.0, 1.0])


def _reference_trajectory(obs_dt=0.5, n_obs=21, sub=2000):
    def rhs(u):
        return np.array([THETA[0] * u[0] - THETA[1] * u[0] * u[1],
                         -THETA[2] * u[1] + THETA[3] * u[0] * u[1]])
    u = np.array([1.0, 1.0])
    out = [u.copy()]
    h = obs_dt / sub
    for _ in range(1, n_obs):
        for _ in range(sub):
            k1 = rhs(u)
            k2 = rhs(u + h / 2 * k1)
            k3 = rhs(u + h / 2 * k2)
            k4 = rhs(u + h * k3)
            u = u + h / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
        out.append(u.copy())
    return np.array(out)


def _lv(obj, params, obs):
    return O.Problem(obj, 4, np.zeros(4, F), np.full(4, 5, F), np.asarray(params, F),
                     obs=np.asarray(obs, F))


def test_fixed_tsit5_converges_at_fifth_order():
    traj = _reference_trajectory()
    sse = {}
    for spo in (2, 4):
        pr = _lv(O.LOTKA_VOLTERRA_TSIT5, [1, 1, 0.5 / spo, spo], traj)
        sse[spo] = float(O.objective(pr, THETA.astype(F)))
    ratio = np.sqrt(sse[2] / sse[4])          # error ratio for h -> h/2
    assert 2 ** 4.5 < ratio < 2 ** 7, ratio


def test_adaptive_tsit5_meets_its_tolerances():
    traj = _reference_trajectory()
    sse = [float(O.objective(_lv(O.LOTKA_VOLTERRA_ATSIT5, [1, 1, 0.5, 0.01, a, r], traj),
                             THETA.astype(F)))
           for a, r in ((1e-3, 1e-3), (1e-6, 1e-3), (1e-6, 1e-6))]
    assert sse[0] > sse[1] > sse[2]
    assert sse[0] < 0.1 and sse[2] < 1e-8


def test_adaptive_gives_up_with_inf():
    # parameters that make the solution explode end with +Inf
    traj = _reference_trajectory()
    pr = _lv(O.LOTKA_VOLTERRA_ATSIT5, [1, 1, 0.5, 0.01, 1e-9, 1e-9], traj)
    assert O.objective(pr, np.array([5, 0, 0, 5], F)) == np.inf


def test_fp64_evaluator_agrees_with_the_fp32_oracle(pps):
    from ppswarm import _lib as L

    rng = np.random.default_rng(1)
    noisy = (_reference_trajectory() + rng.normal(0, 0.05, (21, 2))).astype(F)
    for obj, params in ((O.LOTKA_VOLTERRA_TSIT5, [1, 1, 0.125, 4]),
                        (O.LOTKA_VOLTERRA_ATSIT5, [1, 1, 0.5, 0.01, 1e-7, 1e-6])):
        f32 = float(O.objective(_lv(obj, params, noisy), THETA.astype(F)))
        pr = L.pps_problem()
        par = np.asarray(params, F)
        lb, ub = np.zeros(4, F), np.full(4, 5, F)
        pr.objective, pr.d = obj, 4
        pr.lb, pr.ub = lb.ctypes.data_as(L.fp), ub.ctypes.data_as(L.fp)
        pr.params, pr.n_params = par.ctypes.data_as(L.fp), len(par)
        pr.obs, pr.n_obs = noisy.ctypes.data_as(L.fp), 21
        x = THETA.astype(np.float64)
        out = C.c_double()
        L.check(L.lib().pps_eval_f64(C.byref(pr), 1, x.ctypes.data_as(C.POINTER(C.c_double)),
                                     C.byref(out)))
        # FP32 rounding over the trajectory: ~1e-5 relative on this loss
        assert abs(out.value - f32) < 5e-5 * f32 + 1e-7, (obj, out.value, f32)


# ---------------------------------------------------------------------------
# saveat semantics: observations read from Tsit5's continuous extension
# (GPUTsit5 filling `saveat` between steps, as the reference's benchmark does
# with dt = 0.1 and samples every 0.01,
# benchmarks/Lotka_Volterra/lotka_volterra.jl:124-127).

def test_interpolant_weights_satisfy_the_continuous_order_conditions():
    # a continuous extension of order 4: sum_i b_i(theta) Phi_i = theta^r / gamma
    # for the 8 trees of order <= 4 at every theta, b_i(1) = b_i, b_i(0) = 0,
    # and (inside the step) not order 5
    a, _ = O.tsit5_tableau()
    A = np.zeros((7, 7))
    A[:, :6] = a.astype(np.float64)
    c = A.sum(axis=1)
    Ac = A @ c
    for th in (0.1, 0.25, 0.37, 0.5, 0.83, 1.0):
        B = O.tsit5_interp_weights(th).astype(np.float64)
        conds = [B.sum() - th, B @ c - th ** 2 / 2, B @ c ** 2 - th ** 3 / 3, B @ Ac - th ** 3 / 6,
                 B @ c ** 3 - th ** 4 / 4, B @ (c * Ac) - th ** 4 / 8,
                 B @ (A @ c ** 2) - th ** 4 / 12, B @ (A @ Ac) - th ** 4 / 24]
        assert max(abs(x) for x in conds) < 2e-6, th
    np.testing.assert_allclose(O.tsit5_interp_weights(1.0), A[6], atol=2e-6)
    assert not O.tsit5_interp_weights(0.0).any()
    B = O.tsit5_interp_weights(0.5).astype(np.float64)
    assert abs(B @ c ** 4 - 0.5 ** 5 / 5) > 1e-4


def _dense_reference():
    return _reference_trajectory(obs_dt=0.05, n_obs=61, sub=200)   # t in [0, 3]


def test_saveat_on_step_ends_matches_step_end_sampling():
    # saveat = spo * dt puts every observation on a step end, where the
    # extension (theta = 1) must give the step's own value: losses against
    # any data agree to rounding
    rng = np.random.default_rng(2)
    for obs in (np.zeros((21, 2)), rng.uniform(0, 5, (21, 2))):
        for spo in (1, 2, 5):
            a = float(O.objective(_lv(O.LOTKA_VOLTERRA_TSIT5, [1, 1, 0.5 / spo, spo], obs),
                                  THETA.astype(F)))
            b = float(O.objective(_lv(O.LOTKA_VOLTERRA_TSIT5_SAVEAT, [1, 1, 0.5, 0.5 / spo], obs),
                                  THETA.astype(F)))
            assert abs(a - b) <= 2e-6 * a, (spo, a, b)


def test_saveat_fixed_steps_converge_at_fifth_order():
    # steps of 0.4 .. 0.1 with samples every 0.05: most samples come from the
    # extension; a wrong weight would drop the order to 1 or 2
    traj = _dense_reference()
    err = [np.sqrt(float(O.objective(_lv(O.LOTKA_VOLTERRA_TSIT5_SAVEAT, [1, 1, 0.05, dt], traj),
                                     THETA.astype(F))))
           for dt in (0.4, 0.2, 0.1)]
    for e0, e1 in zip(err, err[1:]):
        assert 4.5 < np.log2(e0 / e1) < 7, err


def test_saveat_adaptive_meets_its_tolerances():
    traj = _dense_reference()
    err = [np.sqrt(float(O.objective(
        _lv(O.LOTKA_VOLTERRA_ATSIT5_SAVEAT, [1, 1, 0.05, 0.01, tol, tol], traj), THETA.astype(F))))
        for tol in (1e-3, 1e-5, 1e-7)]
    assert err[0] > err[1] > err[2] and err[2] < 2e-5, err


def test_saveat_adaptive_gives_up_with_inf():
    traj = _reference_trajectory()
    pr = _lv(O.LOTKA_VOLTERRA_ATSIT5_SAVEAT, [1, 1, 0.5, 0.01, 1e-9, 1e-9], traj)
    assert O.objective(pr, np.array([5, 0, 0, 5], F)) == np.inf


def _f64(obj, params, obs):
    from ppswarm import _lib as L

    pr = L.pps_problem()
    par = np.asarray(params, F)
    lb, ub = np.zeros(4, F), np.full(4, 5, F)
    obs = np.ascontiguousarray(obs, F)
    pr.objective, pr.d = obj, 4
    pr.lb, pr.ub = lb.ctypes.data_as(L.fp), ub.ctypes.data_as(L.fp)
    pr.params, pr.n_params = par.ctypes.data_as(L.fp), len(par)
    pr.obs, pr.n_obs = obs.ctypes.data_as(L.fp), len(obs)
    x = THETA.astype(np.float64)
    out = C.c_double()
    L.check(L.lib().pps_eval_f64(C.byref(pr), 1, x.ctypes.data_as(C.POINTER(C.c_double)),
                                 C.byref(out)))
    return out.value


def test_saveat_fp64_evaluator_converges_and_agrees_with_the_oracle(pps):
    traj = _dense_reference()
    err = [np.sqrt(_f64(O.LOTKA_VOLTERRA_TSIT5_SAVEAT, [1, 1, 0.05, dt], traj))
           for dt in (0.4, 0.2, 0.1, 0.05)]
    for e0, e1 in zip(err, err[1:]):
        assert 4.5 < np.log2(e0 / e1) < 7, err
    rng = np.random.default_rng(3)
    noisy = (traj + rng.normal(0, 0.05, traj.shape)).astype(F)
    for obj, params in ((O.LOTKA_VOLTERRA_TSIT5_SAVEAT, [1, 1, 0.05, 0.03]),
                        (O.LOTKA_VOLTERRA_ATSIT5_SAVEAT, [1, 1, 0.05, 0.01, 1e-7, 1e-6])):
        f32 = float(O.objective(_lv(obj, params, noisy), THETA.astype(F)))
        f64 = _f64(obj, params, noisy)
        assert abs(f64 - f32) < 5e-5 * f32 + 1e-7, (obj, f64, f32)
