"""The reference's regression suite for the GPU algorithms
(test/regression.jl:5-342), every ParallelPSOKernel / ParallelSyncPSOKernel
case through the product path: Rosenbrock p = [1, 100], x0 = 0, 2000
particles, bounds [-1, 10], infinite bounds or none, the default or the Latin
hypercube sampler, and the reference's objective thresholds.

The thresholds are quality bars of the reference's own (Julia RNG) runs; here
the random stream is Philox, so they are met as bars, not as values.  The seeds
are the reference's: 123 for N = 2, 3 and 1 for N = 4 (regression.jl:8,208).
ParallelSerialPSO and ParallelPSOArray are CPU-only algorithms of the reference
and are not mirrored."""
import numpy as np
import pytest

import ppswarm as P
from parity_util import note

pytestmark = pytest.mark.gpu
F = np.float32


def _prob(N, bounds):
    kw = {}
    if bounds == "box":
        kw = dict(lb=np.full(N, -1.0, F), ub=np.full(N, 10.0, F))
    elif bounds == "inf":
        kw = dict(lb=np.full(N, -np.inf, F), ub=np.full(N, np.inf, F))
    return P.OptimizationProblem(P.Rosenbrock(), np.zeros(N, F), [1.0, 100.0], **kw)


def _alg(name):
    return P.ParallelPSOKernel(2000) if name == "pso" else P.ParallelSyncPSOKernel(2000)


# (N, bounds, algorithm, maxiters, threshold), regression.jl line of the assertion
CASES = [
    (2, "box", "pso", 500, 1e-4),    # :63-69
    (2, "box", "sync", 500, 3e-3),   # :81-87
    (2, "inf", "pso", 500, 1e-4),    # :121-127
    (2, "inf", "sync", 500, 2e-4),   # :139-145
    (2, "none", "pso", 500, 1e-4),   # :176-182
    (2, "none", "sync", 500, 2e-2),  # :192-198
    (3, "box", "pso", 500, 1e-4),
    (3, "box", "sync", 500, 3e-3),
    (3, "inf", "pso", 500, 1e-4),
    (3, "inf", "sync", 500, 2e-4),
    (3, "none", "pso", 500, 1e-4),
    (3, "none", "sync", 500, 2e-2),
    (4, "box", "pso", 2000, 2e-2),   # :257-263
    (4, "box", "sync", 2000, 3e-2),  # :275-281
    (4, "none", "pso", 1000, 2e-3),  # :318-324
    (4, "none", "sync", 2000, 4e-1),  # :336-342
]


@pytest.mark.parametrize("N,bounds,alg,maxiters,tol", CASES)
def test_reference_regression_threshold(N, bounds, alg, maxiters, tol):
    note(f"regression N={N} bounds={bounds} alg={alg}")
    seed = 1 if N == 4 else 123
    sol = P.solve(_prob(N, bounds), _alg(alg), maxiters, seed=seed)
    assert sol.retcode == "Default"
    assert sol.objective < tol, (sol.objective, tol)
    if bounds == "box":
        assert (sol.u >= -1).all() and (sol.u <= 10).all()


# Latin hypercube runs: the reference checks the return code only
# (regression.jl:53-61,71-79,111-119,129-137,166-174,184-190,247-255,265-273,
# 308-316,326-334)
@pytest.mark.parametrize("N,bounds,maxiters", [(2, "box", 500), (2, "inf", 500), (2, "none", 500),
                                               (3, "box", 500), (3, "inf", 500), (3, "none", 500),
                                               (4, "box", 2000), (4, "none", 1000)])
@pytest.mark.parametrize("alg", ["pso", "sync"])
def test_reference_regression_latin_hypercube(N, bounds, maxiters, alg):
    note(f"regression LHS N={N} bounds={bounds} alg={alg}")
    seed = 1 if N == 4 else 123
    cache = P.init(_prob(N, bounds), _alg(alg), seed=seed, sampler=P.LatinHypercubeSample())
    sol = P.solve_(cache, maxiters)
    assert sol.retcode == "Default" and np.isfinite(sol.objective)
