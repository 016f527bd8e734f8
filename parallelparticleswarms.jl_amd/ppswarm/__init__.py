"""ppswarm — Python host interface of the MI355X particle-swarm backend (libpps).

Mirrors ParallelParticleSwarms.jl's ``solve(OptimizationProblem, alg)`` API for
``ParallelSyncPSOKernel``, ``ParallelPSOKernel`` and ``HybridPSO`` with the
swarm resident in HBM and every iteration in the HIP kernels of libpps.
"""
from ._lib import PPSError, declared_functions, lib  # noqa: F401
from .api import (  # noqa: F401
    Beale, Booth, BoundaryQuadratic, Context, GBest, GPUUnboundedSampler, GPUUniformSampler,
    ArmijoBacktracking, Himmelblau, HybridPSO, HybridPSOCache, LatinHypercubeSample, MI355XBackend,
    Objective, OptimizationProblem, OptimizationSolution, ParallelPSOKernel, QuadraticConstraints,
    ParallelSyncPSOKernel, PSOCache, Quadratic, Rastrigin, Rosen2, Rosenbrock, SimpleLBFGS,
    Sphere, StrongWolfeLineSearch, Swarm, autotune_sync, tuned_launch, context, eval_f64, init, lotka_volterra_sse, fitzhugh_nagumo_sse, ode_sse, reinit_, seed_, solve, solve_,
    swarm_params)
from .serve import Request, Result, Server, batch_solve, solve_alone  # noqa: F401
from .workloads import served_mix  # noqa: F401
from .ode import parameter_estim_ode  # noqa: F401
