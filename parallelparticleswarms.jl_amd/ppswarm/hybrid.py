"""HybridPSO: particle-swarm search followed by per-particle local refinement
(reference SciMLBase.solve! for HybridPSOCache with HybridPSO{<:SimpleLBFGS},
src/hybrid.jl:100-158).

Flow, as in the reference: run the PSO stage's solve! on the cache, then
projected L-BFGS from every particle's best position on the device (history
= the local optimiser's threshold, line search from ``linesearch`` or the
local optimiser's), take the argmin of the local results (lowest index on
ties), and keep it when it beats the PSO result.  `stats.time` is PSO time +
local time.  The local search stops on ``abstol`` (max-norm of the projected
gradient step), on ``reltol`` (a step with |f - f_new| <= reltol |f|) and on
``local_maxiters``; SimpleOptimization's own stopping rules are not
available to compare with.  A line search other than StrongWolfeLineSearch
or ArmijoBacktracking is refused, as the reference refuses anything but
StrongWolfeLineSearch (src/hybrid.jl:129-130).
"""
from __future__ import annotations

import time

import numpy as np

from . import api


def solve_hybrid_cache(cache, maxiters=100, *, w=0.7298, wdamp=1.0, local_maxiters=50,
                       abstol=None, reltol=None, linesearch=None, original=False):
    alg = cache.alg
    lo = alg.local_opt
    ls = lo.linesearch if linesearch is None else linesearch
    rt = 0.0 if reltol is None else float(reltol)
    if not (rt >= 0.0 and np.isfinite(rt)):
        raise ValueError(f"reltol must be finite and >= 0, got {reltol}")
    api.lbfgs_params(local_maxiters, lo.threshold, 0.0, ls, rt)  # refuses other line searches
    sol_pso = api.solve_(cache.pso_cache, maxiters, w=w, wdamp=wdamp, original=original)
    best_u, best_obj = sol_pso.u, sol_pso.objective
    t0 = time.perf_counter()
    sw = cache.pso_cache.swarm
    _, _, bx, bf, bi = sw.lbfgs(maxiters=local_maxiters, history=lo.threshold,
                                abstol=0.0 if abstol is None else float(abstol),
                                want_all=False, linesearch=ls, reltol=rt)
    local_ms = sw.elapsed_ms()
    local_obj = float(api.eval_f64(cache.pso_cache.prob, bx[None, :])[0]) if np.isfinite(bf) \
        else np.inf
    index = sol_pso.index
    f32 = sol_pso.objective_f32
    if local_obj < best_obj:
        best_u, best_obj, index, f32 = bx, local_obj, bi, bf
    t1 = time.perf_counter()
    return api.OptimizationSolution(
        best_u, best_obj, f32, sol_pso.original,
        api.Stats(sol_pso.stats.time + (t1 - t0), sol_pso.stats.device_ms + local_ms),
        "Default", index)
