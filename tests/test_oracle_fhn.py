"""Pins of the oracle's FitzHugh-Nagumo model (CPU): the second ODE model of
the estimation path (benchmarks/Fitzhugh_Nagumo/ode_estimation.jl:6-12),
dv = v - v^3/3 - w + l, dw = tau_inv (v + a - b w), rates (a, b, tau_inv, l).

* at an equilibrium (v*, w*) (the real root of the cubic the two nullclines
  give, from numpy.roots) every solver leaves the state where it is: the loss
  against constant data is 0; a wrong sign or coefficient moves it;
* RK4 converges at 4th order and Tsit5 (saveat form, FP64 evaluator) at 5th
  order to an FP64 trajectory;
* the FP64 evaluator agrees with the FP32 oracle on noisy data, for the RK4
  and the adaptive saveat forms.
"""
import numpy as np
import pytest

import pps_oracle as O

F = np.float32
THETA = np.array([0.7, 0.8, 0.08, 0.5])     # the reference benchmark's rates
SOLVERS = [(O.FITZHUGH_NAGUMO, [0.1, 10]), (O.FITZHUGH_NAGUMO_TSIT5, [0.1, 10]),
           (O.FITZHUGH_NAGUMO_ATSIT5, [1.0, 0.01, 1e-6, 1e-6]),
           (O.FITZHUGH_NAGUMO_TSIT5_SAVEAT, [0.1, 0.03]),
           (O.FITZHUGH_NAGUMO_ATSIT5_SAVEAT, [0.1, 0.01, 1e-6, 1e-6])]


def _equilibrium(th):
    a, b, _, l = th
    # w = (v + a) / b on the w-nullcline; v - v^3/3 - w + l = 0 on the v-nullcline
    r = np.roots([-1 / 3, 0.0, 1 - 1 / b, l - a / b])
    v = r[np.abs(r.imag) < 1e-12].real[0]
    return v, (v + a) / b


def _prob(obj, params, obs):
    return O.Problem(obj, 4, np.zeros(4, F), np.full(4, 5, F), np.asarray(params, F),
                     obs=np.asarray(obs, F))


def _rhs(u, th=THETA):
    a, b, tinv, l = th
    return np.array([u[0] - u[0] ** 3 / 3 - u[1] + l, tinv * (u[0] + a - b * u[1])])


def _trajectory(obs_dt, n_obs, sub, u0=(1.0, 1.0)):
    u = np.array(u0, np.float64)
    out = [u.copy()]
    h = obs_dt / sub
    for _ in range(1, n_obs):
        for _ in range(sub):
            k1 = _rhs(u)
            k2 = _rhs(u + h / 2 * k1)
            k3 = _rhs(u + h / 2 * k2)
            k4 = _rhs(u + h * k3)
            u = u + h / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
        out.append(u.copy())
    return np.array(out)


@pytest.mark.parametrize("th", [THETA, np.array([0.3, 1.2, 0.4, 0.1])])
@pytest.mark.parametrize("obj,rest", SOLVERS)
def test_equilibrium_is_fixed_for_every_solver(obj, rest, th):
    v, w = _equilibrium(th)
    assert abs(_rhs((v, w), th)).max() < 1e-12
    data = np.tile([v, w], (31, 1))
    assert float(O.objective(_prob(obj, [v, w] + rest, data), th.astype(F))) < 1e-10
    # the same data from a point off the equilibrium is far from it
    assert float(O.objective(_prob(obj, [v + 0.5, w] + rest, data), th.astype(F))) > 1e-2


def test_rk4_converges_at_fourth_order():
    traj = _trajectory(0.5, 21, 2000)
    err = [np.sqrt(float(O.objective(_prob(O.FITZHUGH_NAGUMO, [1, 1, 0.5 / spo, spo], traj),
                                     THETA.astype(F))))
           for spo in (1, 2, 4)]
    for e0, e1 in zip(err, err[1:]):
        assert 3.5 < np.log2(e0 / e1) < 5.5, err


def _f64(pps, obj, params, obs):
    from ppswarm import _lib as L
    import ctypes as C

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


def test_tsit5_saveat_converges_at_fifth_order_and_agrees(pps):
    traj = _trajectory(0.1, 101, 400)
    e64 = [np.sqrt(_f64(pps, O.FITZHUGH_NAGUMO_TSIT5_SAVEAT, [1, 1, 0.1, dt], traj))
           for dt in (0.4, 0.2)]
    assert 4.5 < np.log2(e64[0] / e64[1]) < 7.5, e64
    e32 = np.sqrt(float(O.objective(_prob(O.FITZHUGH_NAGUMO_TSIT5_SAVEAT, [1, 1, 0.1, 0.4], traj),
                                    THETA.astype(F))))
    assert abs(e32 - e64[0]) < 0.02 * e64[0], (e32, e64[0])


def test_fp64_evaluator_agrees_with_the_fp32_oracle(pps):
    traj = _trajectory(0.1, 101, 400)
    noisy = (traj + np.random.default_rng(3).normal(0, 0.05, traj.shape)).astype(F)
    for obj, params in ((O.FITZHUGH_NAGUMO, [1, 1, 0.01, 10]),
                        (O.FITZHUGH_NAGUMO_ATSIT5_SAVEAT, [1, 1, 0.1, 0.01, 1e-7, 1e-6])):
        f32 = float(O.objective(_prob(obj, params, noisy), THETA.astype(F)))
        f64 = _f64(pps, obj, params, noisy)
        assert abs(f64 - f32) < 5e-5 * f32, (obj, f64, f32)
