"""A seeded sweep of swarm configurations against the oracle, bit for bit:
objective (Rosenbrock, Rastrigin, sphere), dimension (the exact-dimension
kernels 1, 2, 3, 4, 8, 10, 16, 32 and runtime-dimension ones between), swarm
size, launch mode (one resident workgroup, captured graph, one launch per
iteration), particles per workgroup and workgroups per CU where compiled,
global-best source, coefficients, inertia and damping, bounds and a shifted
optimum; the async kernel over the same space at any size, and the queue-lock
kernel in one workgroup (n <= 1024, where it is deterministic).  Each case runs
two chunks of iterations on the GPU and in the oracle and compares the whole
state and the global best."""
import numpy as np
import pytest

import pps_oracle as O
from parity_util import PARAMS, assert_same_gbest, assert_same_state, note
from ppswarm import _lib as L

pytestmark = pytest.mark.gpu

EXACT = (1, 2, 3, 4, 8, 10, 16, 32)
RUNTIME = (5, 6, 7, 12, 20, 27)


def _cases(count=160, seed=20261204):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        obj = int(rng.choice([O.ROSENBROCK, O.RASTRIGIN, O.SPHERE]))
        d = int(rng.choice(EXACT + RUNTIME))
        if obj == O.ROSENBROCK and d == 1:
            d = 2
        launch = int(rng.choice([L.LAUNCH_AUTO, L.LAUNCH_GRAPH, L.LAUNCH_EAGER,
                                 L.LAUNCH_WORKGROUP]))
        n = int(rng.integers(2, 1025)) if launch == L.LAUNCH_WORKGROUP else \
            int(rng.choice([int(rng.integers(2, 5000)), int(rng.integers(5000, 40000))]))
        ppw = 0
        if d in EXACT and launch != L.LAUNCH_WORKGROUP:
            ppw = int(rng.choice([0, 256, 512] + ([1024] if d <= 16 else [])))
            if d == 32 and ppw == 1024:
                ppw = 512
        grid = int(rng.choice([0, 1, 2, 3, 5, 8]))
        gbest_from = int(rng.choice([L.GBEST_PBEST, L.GBEST_CURRENT]))
        c1 = float(rng.choice([1.4962, 0.5, 2.0, 0.0]))
        c2 = float(rng.choice([1.4962, 0.75, 2.5]))
        w = float(rng.choice([0.7298, 0.9, 0.4]))
        wdamp = float(rng.choice([1.0, 0.99, 0.9]))
        half = float(rng.choice([1.0, 5.12, 10.0]))
        shifted = bool(rng.integers(0, 2))
        chunks = (int(rng.integers(1, 6)), int(rng.integers(1, 9)))
        out.append((k, obj, d, n, launch, ppw, grid, gbest_from, c1, c2, w, wdamp, half,
                    shifted, chunks, int(rng.integers(1, 1 << 30))))
    return out


@pytest.mark.parametrize("case", _cases(), ids=lambda c: f"case{c[0]}")
def test_random_config(case):
    (k, obj, d, n, launch, ppw, grid, gbest_from, c1, c2, w, wdamp, half, shifted, chunks,
     seed) = case
    note(f"random config {k}: obj={obj} d={d} n={n} launch={launch} ppw={ppw} grid={grid} "
         f"gbest_from={gbest_from} c=({c1},{c2}) w={w} wdamp={wdamp} half={half} "
         f"shift={shifted} chunks={chunks} seed={seed}")
    import ppswarm as P

    lb = np.full(d, -half, np.float32)
    ub = np.full(d, half, np.float32)
    shift = (np.random.default_rng(seed).uniform(-0.5, 0.5, d).astype(np.float32)
             if shifted else None)
    params = PARAMS.get(obj, ())
    marker = {O.ROSENBROCK: P.Rosenbrock, O.RASTRIGIN: P.Rastrigin, O.SPHERE: P.Sphere}[obj]()
    prob = P.OptimizationProblem(marker, np.zeros(d, np.float32), params if params else None,
                                 lb=lb, ub=ub, shift=shift)
    g = P.Swarm(P.context(), prob,
                P.swarm_params(n, L.MODE_SYNC, seed=seed, gbest_from=gbest_from, launch=launch,
                               grid_per_cu=grid, particles_per_workgroup=ppw, c1=c1, c2=c2))
    o = O.Swarm(O.Problem(obj, d, lb, ub, np.asarray(params, np.float32), shift), n, seed)
    assert_same_state(g, o, "init")
    for chunk in chunks:
        g.run(chunk, w, wdamp)
        o.run_sync(chunk, w, wdamp, c1=c1, c2=c2, gbest_from=gbest_from)
        assert_same_state(g, o, f"after {chunk}")
        assert_same_gbest(g, o, f"after {chunk}")
    g.free()


def _other_cases(mode, count, seed):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        obj = int(rng.choice([O.ROSENBROCK, O.RASTRIGIN, O.SPHERE]))
        d = int(rng.choice(EXACT + RUNTIME))
        if obj == O.ROSENBROCK and d == 1:
            d = 2
        n = int(rng.integers(2, 1025)) if mode == L.MODE_QLOCK else int(rng.integers(2, 40000))
        c1 = float(rng.choice([1.4962, 0.5, 0.0]))
        c2 = float(rng.choice([1.4962, 0.75, 2.5]))
        w = float(rng.choice([0.7298, 0.9, 0.4]))
        wdamp = float(rng.choice([1.0, 0.99, 0.9]))
        half = float(rng.choice([1.0, 5.12, 10.0]))
        min_improvers = int(rng.choice([1, 2, 3]))
        chunks = (int(rng.integers(1, 6)), int(rng.integers(1, 9)))
        out.append((k, obj, d, n, c1, c2, w, wdamp, half, min_improvers, chunks,
                    int(rng.integers(1, 1 << 30))))
    return out


@pytest.mark.parametrize("mode,case",
                         [(L.MODE_ASYNC, c) for c in _other_cases(L.MODE_ASYNC, 40, 11)] +
                         [(L.MODE_QLOCK, c) for c in _other_cases(L.MODE_QLOCK, 40, 12)],
                         ids=lambda v: f"case{v[0]}" if isinstance(v, tuple) else f"mode{v}")
def test_random_async_and_queue_lock(mode, case):
    k, obj, d, n, c1, c2, w, wdamp, half, min_improvers, chunks, seed = case
    note(f"random mode {mode} config {k}: obj={obj} d={d} n={n} c=({c1},{c2}) w={w} "
         f"wdamp={wdamp} half={half} min_improvers={min_improvers} chunks={chunks} seed={seed}")
    import ppswarm as P

    lb = np.full(d, -half, np.float32)
    ub = np.full(d, half, np.float32)
    params = PARAMS.get(obj, ())
    marker = {O.ROSENBROCK: P.Rosenbrock, O.RASTRIGIN: P.Rastrigin, O.SPHERE: P.Sphere}[obj]()
    prob = P.OptimizationProblem(marker, np.zeros(d, np.float32), params if params else None,
                                 lb=lb, ub=ub)
    g = P.Swarm(P.context(), prob,
                P.swarm_params(n, mode, seed=seed, c1=c1, c2=c2,
                               qlock_min_improvers=min_improvers))
    o = O.Swarm(O.Problem(obj, d, lb, ub, np.asarray(params, np.float32), None), n, seed)
    for chunk in chunks:
        g.run(chunk, w, wdamp)
        if mode == L.MODE_ASYNC:
            o.run_async(chunk, w, wdamp, c1=c1, c2=c2)
        else:
            o.run_qlock(chunk, w, wdamp, c1=c1, c2=c2, min_improvers=min_improvers)
        assert_same_state(g, o, f"after {chunk}")
    assert_same_gbest(g, o)
    g.free()
