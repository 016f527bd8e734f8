"""Lotka-Volterra parameter estimation (fused PSO step + in-kernel FP32 RK4, and
Tsit5 with fixed or adaptive steps).

At the benchmark shape (2^14 particles, 4 parameters in [0, 5], 101
observations, RK4 dt = 0.01): bit-exact parity with the oracle over the first
iterations, and a full 300-iteration run that recovers the true parameters
(1.5, 1, 3, 1) to within the noise of the committed data."""
import json
import os

import numpy as np
import pytest

import pps_oracle as O
import ppswarm as P
from parity_util import assert_same_gbest, assert_same_state, make_pair, note
from ppswarm import _lib as L

pytestmark = pytest.mark.gpu

DOC = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "lv_obs.json")))
OBS = np.array(DOC["obs"], np.float32)


def test_lv_full_shape_parity():
    note("LV 2^14 parity 3 iterations")
    g, o = make_pair(O.LOTKA_VOLTERRA, 4, 1 << 14, seed=77, lb=0.0, ub=5.0, obs=OBS,
                     clamp=L.CLAMP_LEXICOGRAPHIC, launch=L.LAUNCH_GRAPH)
    assert_same_state(g, o, "after init")
    g.run(3)
    o.run_sync(3)
    assert_same_state(g, o, "after 3 iterations")
    assert_same_gbest(g, o)


@pytest.mark.parametrize("gbest_from", [L.GBEST_PBEST, L.GBEST_CURRENT])
def test_lv_pair_lanes_ragged_multi_tile(gbest_from):
    # The two-lanes-per-particle kernel with a partial last tile (20001 is not
    # a multiple of 64) and several tiles per workgroup (one workgroup per
    # CU), under both global-best sources
    note(f"LV RK4 20001 particles, grid_per_cu=1, gbest_from={gbest_from}")
    g, o = make_pair(O.LOTKA_VOLTERRA, 4, 20001, seed=91, lb=0.0, ub=5.0, obs=OBS,
                     clamp=L.CLAMP_LEXICOGRAPHIC, launch=L.LAUNCH_GRAPH, grid_per_cu=1,
                     gbest_from=gbest_from)
    g.run(2)
    o.run_sync(2, gbest_from=gbest_from)
    assert_same_state(g, o, "after 2 iterations")
    assert_same_gbest(g, o)


def test_lv_estimation_recovers_parameters():
    note("LV estimation 300 iterations")
    sol, sw = P.parameter_estim_ode(OBS, np.zeros(4), np.full(4, 5.0), 1 << 14, 300, seed=5)
    theta = np.array(DOC["theta"])
    # FP64 re-evaluated loss at the noise floor of the data (SSE of the noise
    # itself is ~0.54) and parameters within a few percent
    assert sol.objective < 0.6
    np.testing.assert_allclose(sol.u, theta, rtol=0.05)
    assert abs(sol.objective - sol.objective_f32) < 1e-3


def _tsit5_pair(n, seed, adaptive, mode=L.MODE_SYNC, launch=L.LAUNCH_GRAPH):
    f = P.lotka_volterra_sse(OBS, u0=(1.0, 1.0), dt=0.01, steps_per_obs=10, ode_alg="tsit5",
                             adaptive=adaptive, abstol=1e-6, reltol=1e-3)
    lb, ub = np.zeros(4, np.float32), np.full(4, 5.0, np.float32)
    prob = P.OptimizationProblem(f, np.zeros(4), lb=lb, ub=ub)
    g = P.Swarm(P.context(), prob, P.swarm_params(n, mode, seed=seed, clamp=L.CLAMP_LEXICOGRAPHIC,
                                                  launch=launch))
    obj = O.LOTKA_VOLTERRA_ATSIT5 if adaptive else O.LOTKA_VOLTERRA_TSIT5
    o = O.Swarm(O.Problem(obj, 4, lb, ub, f.params(), clamp=O.CLAMP_LEXICOGRAPHIC, obs=OBS), n,
                seed)
    return g, o


@pytest.mark.parametrize("adaptive", [False, True])
@pytest.mark.parametrize("mode,n", [(L.MODE_SYNC, 4096), (L.MODE_SYNC, 1000), (L.MODE_ASYNC, 2048)])
def test_lv_tsit5_parity(adaptive, mode, n):
    # the reference's default GPUTsit5, fixed (Val(false)) and adaptive (Val(true))
    note(f"LV Tsit5 adaptive={adaptive} mode={mode} n={n}")
    g, o = _tsit5_pair(n, 31 + n, adaptive, mode,
                       L.LAUNCH_GRAPH if mode == L.MODE_SYNC and n > 1024 else L.LAUNCH_AUTO)
    assert_same_state(g, o, "after init")
    assert_same_gbest(g, o, "after init")
    g.run(3)
    if mode == L.MODE_SYNC:
        o.run_sync(3)
    else:
        o.run_async(3)
    assert_same_state(g, o, "after 3 iterations")
    assert_same_gbest(g, o)


@pytest.mark.parametrize("adaptive", [False, True])
def test_lv_tsit5_estimation_recovers_parameters(adaptive):
    note(f"LV Tsit5 estimation adaptive={adaptive}")
    sol, sw = P.parameter_estim_ode(OBS, np.zeros(4), np.full(4, 5.0), 1 << 14, 300, seed=6,
                                    ode_alg="tsit5", adaptive=adaptive)
    theta = np.array(DOC["theta"])
    assert sol.objective < 0.6
    np.testing.assert_allclose(sol.u, theta, rtol=0.05)


def _saveat_pair(n, seed, adaptive, mode, launch):
    # steps of 0.03 (fixed) or from 0.01 (adaptive) against samples every 0.1:
    # most samples are read from Tsit5's continuous extension
    f = P.lotka_volterra_sse(OBS, u0=(1.0, 1.0), dt=0.01 if adaptive else 0.03, ode_alg="tsit5",
                             adaptive=adaptive, abstol=1e-6, reltol=1e-3, saveat=0.1)
    lb, ub = np.zeros(4, np.float32), np.full(4, 5.0, np.float32)
    prob = P.OptimizationProblem(f, np.zeros(4), lb=lb, ub=ub)
    g = P.Swarm(P.context(), prob, P.swarm_params(n, mode, seed=seed, clamp=L.CLAMP_LEXICOGRAPHIC,
                                                  launch=launch))
    obj = O.LOTKA_VOLTERRA_ATSIT5_SAVEAT if adaptive else O.LOTKA_VOLTERRA_TSIT5_SAVEAT
    o = O.Swarm(O.Problem(obj, 4, lb, ub, f.params(), clamp=O.CLAMP_LEXICOGRAPHIC, obs=OBS), n,
                seed)
    return g, o


@pytest.mark.parametrize("adaptive", [False, True])
@pytest.mark.parametrize("mode,n", [(L.MODE_SYNC, 4096), (L.MODE_SYNC, 1000), (L.MODE_ASYNC, 2048)])
def test_lv_tsit5_saveat_parity(adaptive, mode, n):
    # GPUTsit5 with saveat between steps (src/ode_pso.jl:54-170 passing
    # saveat/dt through to the solver)
    note(f"LV Tsit5 saveat adaptive={adaptive} mode={mode} n={n}")
    g, o = _saveat_pair(n, 57 + n, adaptive, mode,
                        L.LAUNCH_GRAPH if mode == L.MODE_SYNC and n > 1024 else L.LAUNCH_AUTO)
    assert_same_state(g, o, "after init")
    assert_same_gbest(g, o, "after init")
    g.run(3)
    if mode == L.MODE_SYNC:
        o.run_sync(3)
    else:
        o.run_async(3)
    assert_same_state(g, o, "after 3 iterations")
    assert_same_gbest(g, o)


def test_reference_lv_benchmark_shape():
    # benchmarks/Lotka_Volterra/lotka_volterra.jl:21-37,91-127: 10000
    # particles, 100 iterations, rates in [0, 10], GPUTsit5 with dt = 0.1 and
    # saveat every 0.01 on [0, 3] against noise-free data
    note("LV reference benchmark shape (saveat 0.01, dt 0.1)")
    doc = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                      "lv_obs_benchmark.json")))
    obs = np.array(doc["obs"], np.float32)
    sol, sw = P.parameter_estim_ode(obs, np.zeros(4), np.full(4, 10.0), 10000, 100, seed=1,
                                    ode_alg="tsit5", dt=0.1, saveat=0.01)
    np.testing.assert_allclose(sol.u, doc["theta"], rtol=1e-3)
    assert sol.objective < 1e-5


FHN_OBS = np.array(json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                               "fhn_obs_benchmark.json")))["obs"], np.float32)


@pytest.mark.parametrize("solver,kw", [
    (O.FITZHUGH_NAGUMO, dict(ode_alg="rk4", dt=0.01, steps_per_obs=10)),
    (O.FITZHUGH_NAGUMO_TSIT5, dict(ode_alg="tsit5", dt=0.05, steps_per_obs=2)),
    (O.FITZHUGH_NAGUMO_ATSIT5, dict(ode_alg="tsit5", adaptive=True, dt=0.01, steps_per_obs=10)),
    (O.FITZHUGH_NAGUMO_TSIT5_SAVEAT, dict(ode_alg="tsit5", dt=0.1, saveat=0.1)),
    (O.FITZHUGH_NAGUMO_ATSIT5_SAVEAT, dict(ode_alg="tsit5", adaptive=True, dt=0.01, saveat=0.1)),
])
@pytest.mark.parametrize("mode,n", [(L.MODE_SYNC, 3000), (L.MODE_ASYNC, 1000)])
def test_fhn_parity(solver, kw, mode, n):
    # FitzHugh-Nagumo (benchmarks/Fitzhugh_Nagumo/ode_estimation.jl:6-12) with
    # every solver: samples 0.1 apart on [0, 3]
    note(f"FHN parity obj={solver} mode={mode} n={n}")
    obs = FHN_OBS[::10]
    f = P.fitzhugh_nagumo_sse(obs, u0=(1.0, 1.0), **kw)
    assert f.code == solver
    lb, ub = np.zeros(4, np.float32), np.full(4, 5.0, np.float32)
    prob = P.OptimizationProblem(f, np.zeros(4), lb=lb, ub=ub)
    g = P.Swarm(P.context(), prob, P.swarm_params(n, mode, seed=71 + solver,
                                                  clamp=L.CLAMP_LEXICOGRAPHIC,
                                                  launch=L.LAUNCH_GRAPH if n > 1024 else L.LAUNCH_AUTO))
    o = O.Swarm(O.Problem(solver, 4, lb, ub, f.params(), clamp=O.CLAMP_LEXICOGRAPHIC, obs=obs), n,
                71 + solver)
    assert_same_state(g, o, "after init")
    g.run(3)
    if mode == L.MODE_SYNC:
        o.run_sync(3)
    else:
        o.run_async(3)
    assert_same_state(g, o, "after 3 iterations")
    assert_same_gbest(g, o)


def test_reference_fhn_benchmark_shape():
    # benchmarks/Fitzhugh_Nagumo/ode_estimation.jl:100-140: 10000 particles, 100
    # iterations, rates in [0, 5], GPUTsit5 dt = 0.1 with saveat every 0.01 on
    # [0, 3] against noise-free data.  Three time units determine l sharply;
    # a, b and tau_inv act through the slow variable and are only loosely
    # determined after 100 iterations, so the fit is judged by its loss
    note("FHN reference benchmark shape (saveat 0.01, dt 0.1)")
    doc = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                      "fhn_obs_benchmark.json")))
    sol, sw = P.parameter_estim_ode(FHN_OBS, np.zeros(4), np.full(4, 5.0), 10000, 100, seed=0,
                                    ode_alg="tsit5", dt=0.1, saveat=0.01, model="fitzhugh_nagumo")
    assert sol.objective < 1e-4
    assert abs(sol.u[3] / doc["theta"][3] - 1) < 1e-3


@pytest.mark.parametrize("saveat", [False, True])
@pytest.mark.parametrize("mode", [L.MODE_SYNC, L.MODE_ASYNC, L.MODE_QLOCK])
def test_adaptive_early_stop_keeps_state(monkeypatch, saveat, mode):
    # An adaptive loss stops once it reaches the particle's best cost; the
    # swarm must be exactly the one with every loss summed in full
    # (PPS_ODE_FULL_LOSS=1), and after enough iterations for stiff parameter
    # vectors and non-improving particles to occur, both equal the oracle.
    note(f"adaptive early stop saveat={saveat} mode={mode}")
    # queue-lock: one workgroup (n <= 1024), where it is deterministic
    n, T = (1000 if mode == L.MODE_QLOCK else 3000), 12
    mk = _saveat_pair if saveat else _tsit5_pair
    launch = L.LAUNCH_GRAPH if mode == L.MODE_SYNC else L.LAUNCH_AUTO
    g, o = mk(n, 5 + mode, True, mode, launch)
    monkeypatch.setenv("PPS_ODE_FULL_LOSS", "1")
    gf, _ = mk(n, 5 + mode, True, mode, launch)
    monkeypatch.delenv("PPS_ODE_FULL_LOSS")
    g.run(T)
    gf.run(T)
    for a, b, name in zip(g.state()[:4], gf.state()[:4], ("x", "v", "best", "best_cost")):
        np.testing.assert_array_equal(a, b, err_msg=name)
    ga, gb = g.gbest(), gf.gbest()
    assert ga.index == gb.index and ga.cost == gb.cost
    if mode == L.MODE_SYNC:
        o.run_sync(T)
    elif mode == L.MODE_ASYNC:
        o.run_async(T)
    else:
        o.run_qlock(T)
    assert_same_state(g, o, f"after {T} iterations")
    assert_same_gbest(g, o)
