"""ODE parameter estimation with PSO (reference parameter_estim_ode!,
src/ode_pso.jl:54-170, fixed or adaptive time steps).

The reference steps every particle (ode_update_particle_states!, with its
whole-vector max/min clamp, src/ode_pso.jl:18-19), solves one ODE per particle
with DiffEqGPU, reduces the squared error against the data
(_reduce_losses!, :47-52), updates best costs and takes the argmin of the best
costs as the next global best.  Here the model's right-hand side is
integrated inside the fused PSO kernel (FP32 RK4, or Tsit5 with fixed or
adaptive steps as the reference's default GPUTsit5), so one launch per
iteration does all of it.  Models: Lotka-Volterra and FitzHugh-Nagumo, the
reference's two estimation benchmarks.
"""
from __future__ import annotations

import time

import numpy as np

from . import _lib as L
from . import api


def parameter_estim_ode(obs, lb, ub, num_particles=1 << 14, maxiters=300, *, u0=(1.0, 1.0),
                        dt=0.01, steps_per_obs=10, w=0.7298, wdamp=1.0, seed=None,
                        clamp="lexicographic", launch=L.LAUNCH_AUTO, ode_alg="rk4",
                        adaptive=False, abstol=1e-6, reltol=1e-3, saveat=None,
                        model="lotka_volterra"):
    """Estimate the four rates of `model` ("lotka_volterra": (alpha, beta,
    gamma, delta); "fitzhugh_nagumo": (a, b, tau_inv, l)) from `obs` (n_obs x
    2, sample k at t = k * dt * steps_per_obs, or k * saveat when `saveat` is
    given).  ode_alg "rk4" or "tsit5"; adaptive=True (Tsit5 only) is the
    reference's Val(true) path with `dt` as the initial step; saveat (Tsit5
    only) decouples the steps from the samples, which are read from Tsit5's
    interpolant.  Returns (solution, swarm)."""
    f = api.ode_sse(model, obs, u0=u0, dt=dt, steps_per_obs=steps_per_obs, ode_alg=ode_alg,
                    adaptive=adaptive, abstol=abstol, reltol=reltol, saveat=saveat)
    prob = api.OptimizationProblem(f, np.zeros(4, np.float32), lb=np.asarray(lb, np.float32),
                                   ub=np.asarray(ub, np.float32))
    seed = api._next_seed() if seed is None else int(seed)
    cl = L.CLAMP_LEXICOGRAPHIC if clamp == "lexicographic" else L.CLAMP_COMPONENT
    sw = api.Swarm(api.context(), prob, api.swarm_params(num_particles, L.MODE_SYNC, seed=seed,
                                                         clamp=cl, launch=launch))
    t0 = time.perf_counter()
    sw.run(maxiters, w, wdamp)
    t1 = time.perf_counter()
    g = sw.gbest()
    obj = float(api.eval_f64(prob, g.position[None, :])[0])
    sol = api.OptimizationSolution(g.position, obj, g.cost, None,
                                   api.Stats(t1 - t0, sw.elapsed_ms()), "Default", g.index)
    return sol, sw
