"""HybridPSO: particle-swarm search followed by per-particle local refinement
(reference SciMLBase.solve! for HybridPSO{<:SimpleLBFGS}, src/hybrid.jl:100-158).

Flow, as in the reference: run the configured PSO, then projected L-BFGS from
every particle's best position on the device, take the argmin of the local
results (lowest index on ties), and keep it when it beats the PSO result.
`stats.time` is PSO time + local time.
"""
from __future__ import annotations

import time

import numpy as np

from . import api


def solve_hybrid(prob, alg, maxiters=100, *, seed=None, w=0.7298, wdamp=1.0,
                 local_maxiters=50, abstol=None, reltol=None, **_kw):
    lo = alg.local_opt
    if not isinstance(lo, api.SimpleLBFGS):
        raise ValueError("HybridPSO on the MI355X backend supports local_opt=SimpleLBFGS()")
    cache = api.init(prob, alg.pso, seed=seed)
    sol_pso = api.solve_(cache, maxiters, w=w, wdamp=wdamp)
    best_u, best_obj = sol_pso.u, sol_pso.objective
    t0 = time.perf_counter()
    _, _, bx, bf, bi = cache.swarm.lbfgs(
        maxiters=local_maxiters, history=lo.threshold,
        abstol=0.0 if abstol is None else float(abstol), armijo_c=lo.armijo_c,
        armijo_rho=lo.armijo_rho, max_backtracks=lo.max_backtracks, want_all=False)
    local_ms = cache.swarm.elapsed_ms()
    local_obj = float(api.eval_f64(prob, bx[None, :])[0]) if np.isfinite(bf) else np.inf
    index = sol_pso.index
    f32 = sol_pso.objective_f32
    if local_obj < best_obj:
        best_u, best_obj, index, f32 = bx, local_obj, bi, bf
    t1 = time.perf_counter()
    return api.OptimizationSolution(
        best_u, best_obj, f32, None,
        api.Stats(sol_pso.stats.time + (t1 - t0), sol_pso.stats.device_ms + local_ms),
        "Default", index)
