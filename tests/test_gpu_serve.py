"""Served / batched solving on the GPU.

* every problem of a packed launch gets exactly (bit for bit) the result of
  the plain per-problem path (init + one launch per iteration) and of the
  resident-workgroup path;
* a problem's result is the same alone and in a mix (the requirement that
  results do not depend on what they are packed with);
* requests arriving while others run are admitted as HBM frees up (tight
  budget), queued requests can be cancelled, and oversize or unsupported
  requests are refused at submit.
"""
import time

import numpy as np
import pytest

import ppswarm as P
from parity_util import note
from ppswarm import _lib as L

pytestmark = pytest.mark.gpu


def small_mix(count, seed, n_range=(64, 3000), t_range=(20, 60)):
    reqs, _ = P.served_mix(count, seed=seed, n_range=n_range, t_range=t_range)
    return reqs


def same(a, b):
    np.testing.assert_array_equal(a.u, b.u)
    assert a.cost_f32 == b.cost_f32 and a.index == b.index


def test_batch_equals_plain_per_problem_path():
    reqs = small_mix(40, seed=1)
    # odd dimensions too (the packed kernel streams any d <= 32)
    for q, d in zip(reqs[:4], (1, 3, 7, 31)):
        f = q.prob.f
        q.prob = P.OptimizationProblem(f, np.zeros(d, np.float32),
                                       None if f.name == "rastrigin" else [1.0, 100.0],
                                       lb=np.full(d, -5.12), ub=np.full(d, 5.12),
                                       shift=np.linspace(-1, 1, d))
    # and large swarms, which run as several workgroups per iteration
    for q, (n, d) in zip(reqs[4:7], ((16000, 32), (9000, 8), (70000, 2))):
        f = q.prob.f
        q.prob = P.OptimizationProblem(f, np.zeros(d, np.float32),
                                       None if f.name == "rastrigin" else [1.0, 100.0],
                                       lb=np.full(d, -5.12), ub=np.full(d, 5.12))
        q.n = n
    note("batch vs plain, 40 problems")
    res, ms = P.batch_solve(reqs)
    for q, r in zip(reqs, res):
        same(r, P.solve_alone(q, L.LAUNCH_EAGER))
        if q.n <= 1024:
            same(r, P.solve_alone(q, L.LAUNCH_WORKGROUP))
        assert abs(r.cost_f64 - r.cost_f32) <= 1e-4 * max(1.0, abs(r.cost_f64))


def test_alone_and_in_the_mix_agree():
    reqs = small_mix(64, seed=2)
    probe = reqs[17]
    note("alone vs mix")
    with P.Server() as sv:
        alone = sv.wait(sv.submit(probe))
    with P.Server() as sv:
        tickets = [sv.submit(q) for q in reqs]
        mixed = [sv.wait(t) for t in tickets]
    same(alone, mixed[17])


def test_poisson_arrivals_overlap_and_complete():
    reqs, arr = P.served_mix(48, seed=3, qps=2000.0, n_range=(256, 4096), t_range=(50, 200))
    note("poisson arrivals 48 problems")
    with P.Server() as sv:
        t0 = time.perf_counter()
        tickets = []
        for q, a in zip(reqs, arr):
            while time.perf_counter() - t0 < a:
                pass
            tickets.append(sv.submit(q))
        res = [sv.wait(t, timeout=60) for t in tickets]
        st = sv.stats()
    assert st["completed"] == 48 and st["launches"] >= 2
    for q, r in zip(reqs[:8], res[:8]):
        same(r, P.solve_alone(q))
    assert all(r.t_submit <= r.t_admit <= r.t_done for r in res)


def test_tight_hbm_budget_queues_until_memory_frees():
    reqs = small_mix(12, seed=4, n_range=(2000, 3000), t_range=(30, 40))
    biggest = max(q.prob.d for q in reqs)
    # room for about two problems at a time
    budget = 2 * (3 * 4 * biggest * 3072 + 4 * 3072 + 4096)
    note("tight budget")
    with P.Server(hbm_budget=budget) as sv:
        tickets = [sv.submit(q) for q in reqs]
        res = [sv.wait(t, timeout=60) for t in tickets]
        st = sv.stats()
    assert st["peak_bytes"] <= budget
    for q, r in zip(reqs, res):
        same(r, P.solve_alone(q))


def test_submit_refusals_and_cancel():
    note("refusals and cancel")
    prob = P.OptimizationProblem(P.Sphere(), np.zeros(3), lb=np.full(3, -1.0), ub=np.ones(3))
    wide = P.OptimizationProblem(P.Rastrigin(), np.zeros(40), lb=np.full(40, -1.0),
                                 ub=np.ones(40))
    with P.Server(hbm_budget=1 << 20) as sv:
        with pytest.raises(P.PPSError):   # objective outside the packed set
            sv.submit(P.Request(prob, 100, 10, 1))
        with pytest.raises(P.PPSError):   # d > 32
            sv.submit(P.Request(wide, 100, 10, 1))
        big = P.OptimizationProblem(P.Rastrigin(), np.zeros(32), lb=np.full(32, -1.0),
                                    ub=np.ones(32))
        with pytest.raises(P.PPSError):   # larger than the whole budget
            sv.submit(P.Request(big, 100000, 10, 1))
        # a long problem fills the budget; the next one queues and is cancelled
        hog = P.OptimizationProblem(P.Rastrigin(), np.zeros(16), lb=np.full(16, -1.0),
                                    ub=np.ones(16))
        t_hog = sv.submit(P.Request(hog, 3000, 3000, 5))
        t_q = sv.submit(P.Request(hog, 3000, 10, 6))
        sv.cancel(t_q)
        with pytest.raises(P.PPSError):
            sv.wait(t_q)
        assert sv.wait(t_hog, timeout=120).cost_f32 >= 0
        # a final status is returned once; the ticket is then released
        with pytest.raises(P.PPSError):
            sv.poll(t_hog)
        with pytest.raises(P.PPSError):
            sv.wait(t_q)
