"""Packed (batched / served) results against the CPU oracle, at the served-mix
shape: 4096 independent problems, Rosenbrock (p = [1, 100]) or Rastrigin,
d in {2, 4, 8, 16, 32}, swarm size log-uniform in [256, 16384], iteration
budget uniform in [100, 1000], optimum shifted uniformly in [-1, 1]^d.

Each packed problem is a ParallelSyncPSOKernel solve (the GPU backend's
semantics, global best = best position of the argmin, src/kernels.jl:127-177,
src/lowerlevel_solve.jl:1-36) of f(x - shift), so its result must equal the
oracle's `Swarm(...).run_sync(T)` for the same seed bit for bit: position,
FP32 cost and particle index.  The objective tolerance of the served mix is
stated as a relative one: |f32(u) - f64(u)| <= 1e-4 * max(1, |f64(u)|) for
the FP64 re-evaluation of the returned position (costs of 32-D Rastrigin
reach the hundreds, where FP32 alone carries ~1e-5 relative error).
"""
import time

import numpy as np
import pytest

import pps_oracle as O
import ppswarm as P
from parity_util import note

pytestmark = pytest.mark.gpu

REL_TOL = 1e-4


def oracle_solve(q):
    pr = q.prob
    op = O.Problem(pr.f.code, pr.d, pr.lb, pr.ub, pr.p, pr.shift)
    s = O.Swarm(op, q.n, q.seed)
    s.run_sync(q.maxiters, w=q.w, wdamp=q.wdamp)
    return s.gbest()


def same_as_oracle(r, q, what=""):
    pos, cost, idx = oracle_solve(q)
    assert r.index == idx, f"index {r.index} != {idx} {what}"
    assert np.float32(r.cost_f32) == cost, f"cost {r.cost_f32} != {cost} {what}"
    np.testing.assert_array_equal(r.u, pos, err_msg=f"position {what}")


def same(a, b, what=""):
    np.testing.assert_array_equal(a.u, b.u, err_msg=what)
    assert a.cost_f32 == b.cost_f32 and a.index == b.index, what


def within_objective_tolerance(r):
    return abs(r.cost_f64 - r.cost_f32) <= REL_TOL * max(1.0, abs(r.cost_f64))


def test_packed_results_equal_the_oracle():
    # 96 requests of the served mix packed together; the oracle checks the
    # largest-work problem of the whole 4096 mix (a PART problem stepped over
    # hundreds of slices), a shifted problem of every (objective, d), and
    # problems with damped inertia
    full, _ = P.served_mix(4096, seed=2024)
    big = max(full, key=lambda q: q.n * q.prob.d * q.maxiters)
    reqs, _ = P.served_mix(95, seed=91)
    reqs = [big] + reqs
    for q in reqs[1:12]:
        q.wdamp = 0.995
    chosen = {0}
    seen = set()
    for k, q in enumerate(reqs):
        key = (q.prob.f.code, q.prob.d)
        if key not in seen:
            seen.add(key)
            chosen.add(k)
    chosen.update(range(1, 6))
    note(f"packed vs oracle: 96 problems, largest n={big.n} d={big.prob.d} T={big.maxiters}")
    res, _ = P.batch_solve(reqs)
    assert big.n * big.prob.d >= 8192 and big.prob.d == 32
    for k in sorted(chosen):
        same_as_oracle(res[k], reqs[k], f"request {k} (n={reqs[k].n} d={reqs[k].prob.d} "
                                        f"T={reqs[k].maxiters})")
    assert all(within_objective_tolerance(r) for r in res)


def test_full_served_mix_against_the_oracle():
    # all 4096 requests in one batch: every result finite and within the
    # objective tolerance, and every 128th equal to the oracle
    reqs, _ = P.served_mix(4096, seed=2024)
    note("full served mix: 4096 problems")
    res, ms = P.batch_solve(reqs)
    assert len(res) == 4096
    assert all(np.isfinite(r.u).all() and np.isfinite(r.cost_f32) for r in res)
    rel = [abs(r.cost_f64 - r.cost_f32) / max(1.0, abs(r.cost_f64)) for r in res]
    assert max(rel) <= REL_TOL, max(rel)
    for k in range(0, 4096, 128):
        same_as_oracle(res[k], reqs[k], f"request {k}")


def test_every_request_is_the_same_alone_and_in_the_mix():
    # 64 requests of the served-mix distribution arriving as a Poisson stream
    # (so they overlap on the device), each also solved alone on the server
    reqs, arr = P.served_mix(64, seed=77, qps=3000.0)
    note("alone vs mixed: 64 problems")
    with P.Server() as sv:
        t0 = time.perf_counter()
        tickets = []
        for q, a in zip(reqs, arr):
            while time.perf_counter() - t0 < a:
                pass
            tickets.append(sv.submit(q))
        mixed = [sv.wait(t, timeout=120) for t in tickets]
        assert sv.stats()["launches"] >= 2
    # the requests did share the device: some were admitted before others
    # finished
    assert any(mixed[j].t_admit < mixed[i].t_done and j > i
               for i in range(8) for j in range(i + 1, 64))
    with P.Server() as sv:
        for k, q in enumerate(reqs):
            alone = sv.wait(sv.submit(q), timeout=120)
            same(alone, mixed[k], f"request {k}")
    for k in (0, 1, 2, 63):
        same_as_oracle(mixed[k], reqs[k], f"request {k}")
