"""Synthetic workloads of the benchmark configurations.

`served_mix` draws the served/batched mix: independent problems, each from
its own seed, Rosenbrock (p = [1, 100]) or Rastrigin with probability 1/2,
dimension uniform in {2, 4, 8, 16, 32}, swarm size log-uniform in
[256, 16384], iteration budget uniform in [100, 1000], optimum shifted by a
vector uniform in [-1, 1]^d, bounds [-5.12, 5.12]; arrival times form a
Poisson process at the given rate (requests per second).
"""
from __future__ import annotations

import math

import numpy as np

from . import api
from .serve import Request

DIMS = (2, 4, 8, 16, 32)


def served_mix(count: int = 4096, seed: int = 2024, qps: float = float("inf"),
               n_range=(256, 16384), t_range=(100, 1000), bound: float = 5.12):
    rng = np.random.default_rng(seed)
    arr_rng = np.random.default_rng([seed, 1])  # arrivals do not perturb the problems
    reqs, arrivals = [], []
    t = 0.0
    for k in range(count):
        fam = rng.integers(0, 2)
        d = int(DIMS[rng.integers(0, len(DIMS))])
        n = int(round(math.exp(rng.uniform(math.log(n_range[0]), math.log(n_range[1])))))
        T = int(rng.integers(t_range[0], t_range[1] + 1))
        shift = rng.uniform(-1, 1, d).astype(np.float32)
        s = int(rng.integers(0, 2 ** 63))
        f = api.Rosenbrock() if fam == 0 else api.Rastrigin()
        p = [1.0, 100.0] if fam == 0 else None
        prob = api.OptimizationProblem(f, np.zeros(d, np.float32), p, lb=np.full(d, -bound),
                                       ub=np.full(d, bound), shift=shift)
        reqs.append(Request(prob, n, T, s))
        if math.isfinite(qps):
            t += arr_rng.exponential(1.0 / qps)
        arrivals.append(t)
    return reqs, np.array(arrivals)


def particle_iters(reqs) -> int:
    return int(sum(q.n * q.maxiters for q in reqs))
