"""Packed solves of seeded random requests against the oracle, bit for bit:
Rosenbrock or Rastrigin at any dimension 1..32 (not only the served mix's
2, 4, 8, 16, 32), swarm sizes either side of the one-workgroup / several-parts
split (n * d = 8192) and of the 256-particle part tiles, 0 to 60 iterations,
damped inertia, bounds of several widths and shifted optima, all in one batch
and again through a server fed with arrivals.  Each result must equal
`Swarm(...).run_sync(T)` of the oracle for the same seed: position, FP32 cost
and particle index."""
import numpy as np
import pytest

import pps_oracle as O
import ppswarm as P
from parity_util import note

pytestmark = pytest.mark.gpu


def _requests(count, seed):
    rng = np.random.default_rng(seed)
    reqs = []
    for k in range(count):
        rosen = bool(rng.integers(0, 2))
        d = int(rng.integers(2 if rosen else 1, 33))
        # around the split at n * d = 8192 and the 256-particle tiles
        n = int(rng.choice([rng.integers(2, 300), rng.integers(300, 3000),
                            max(2, 8192 // d + int(rng.integers(-40, 40))),
                            256 * int(rng.integers(1, 40)) + int(rng.integers(-1, 2)),
                            rng.integers(3000, 20000)]))
        half = float(rng.choice([1.0, 2.0, 5.12]))
        shift = rng.uniform(-0.5, 0.5, d).astype(np.float32) if rng.integers(0, 2) else None
        f = P.Rosenbrock() if rosen else P.Rastrigin()
        prob = P.OptimizationProblem(f, np.zeros(d, np.float32), [1.0, 100.0] if rosen else None,
                                     lb=np.full(d, -half), ub=np.full(d, half), shift=shift)
        T = 0 if rng.random() < 0.1 else int(rng.choice([rng.integers(1, 10),
                                                          rng.integers(10, 61)]))
        w = float(rng.choice([0.7298, 0.9]))
        wdamp = float(rng.choice([1.0, 0.98]))
        reqs.append(P.Request(prob, n, T, int(rng.integers(1, 1 << 30)), w, wdamp))
    return reqs


def _oracle(q):
    pr = q.prob
    s = O.Swarm(O.Problem(pr.f.code, pr.d, pr.lb, pr.ub, pr.p, pr.shift), q.n, q.seed)
    s.run_sync(q.maxiters, w=q.w, wdamp=q.wdamp)
    return s.gbest()


def _check(results, reqs, what):
    for k, (r, q) in enumerate(zip(results, reqs)):
        pos, cost, idx = _oracle(q)
        tag = f"{what} request {k}: d={q.prob.d} n={q.n} T={q.maxiters} seed={q.seed}"
        assert r.index == idx, tag
        assert np.float32(r.cost_f32) == cost, tag
        np.testing.assert_array_equal(r.u, pos, err_msg=tag)


def test_random_batch_equals_the_oracle():
    reqs = _requests(96, seed=77)
    note(f"random packed batch of {len(reqs)}")
    res, _ = P.batch_solve(reqs)
    _check(res, reqs, "batched")


def test_random_arrivals_equal_the_oracle():
    # the same kind of requests submitted one by one while earlier ones run
    reqs = _requests(48, seed=78)
    note(f"random served arrivals of {len(reqs)}")
    with P.Server() as sv:
        tickets = [sv.submit(q) for q in reqs]
        res = [sv.wait(t) for t in tickets]
    _check(res, reqs, "served")
