"""Pins of the oracle's initial-position samplers (CPU).

* Latin hypercube: the keyed permutation is a bijection of [0, n) for every n,
  key, epoch and coordinate, and the design has exactly one particle per stratum
  of every coordinate (the definition of a Latin hypercube);
* unbounded sampler (src/utils.jl:23-55): x_j = u0_j + r |u0_j|, or r when
  u0_j = 0, which is the uniform sampler's draw on [u0_j, u0_j + |u0_j|] (for
  u0_j < 0: [u0_j, 0]) or on [0, 1], bit for bit;
* given positions (the QuasiMonteCarlo path, src/utils.jl:86-112) are used as
  they are, costs are f at them and the global best is their argmin;
* the reference's unbounded regression thresholds (test/regression.jl:147-199).
"""
import numpy as np
import pytest

import pps_oracle as O

F = np.float32


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 7, 16, 17, 64, 255, 1000, 4099])
def test_lhs_permutation_is_a_bijection(n):
    for j, epoch, seed in [(0, 0, 1), (3, 0, 1), (0, 5, 1), (11, 2, 2 ** 40 + 3)]:
        p = [O.lhs_permute(i, n, j, epoch, seed) for i in range(n)]
        assert sorted(p) == list(range(n)), (n, j, epoch, seed)


def test_lhs_permutation_depends_on_key_epoch_and_coordinate():
    n = 257
    base = [O.lhs_permute(i, n, 0, 0, 9) for i in range(n)]
    assert base != list(range(n))
    for j, epoch, seed in [(1, 0, 9), (0, 1, 9), (0, 0, 10)]:
        assert [O.lhs_permute(i, n, j, epoch, seed) for i in range(n)] != base


@pytest.mark.parametrize("n,lo,hi", [(1, -1.0, 10.0), (7, 0.0, 1.0), (64, -5.12, 5.12),
                                     (1000, -1.0, 10.0), (3001, 2.0, 3.0)])
def test_lhs_design_has_one_particle_per_stratum(n, lo, hi):
    d = 3
    pr = O.Problem(O.SPHERE, d, np.full(d, lo, F), np.full(d, hi, F))
    for epoch in (0, 1):
        s = O.Swarm(pr, n, 77, sampler=O.SAMPLER_LATIN_HYPERCUBE)
        if epoch:
            s.reinit()
        x = s.x.astype(np.float64)
        assert (x >= lo).all() and (x <= hi).all()
        for j in range(d):
            t = np.sort((x[j] - lo) / (hi - lo) * n)
            k = np.arange(n)
            # the k-th smallest point lies in stratum k (FP32 rounding may put
            # a point on the upper edge of its stratum)
            assert (t >= k - 1e-3).all() and (t <= k + 1 + 1e-3).all(), (n, j, epoch)


def test_unbounded_sampler_is_the_uniform_draw_on_the_matching_interval():
    d, n = 4, 500
    u0 = np.array([0.0, 2.0, -3.0, 0.5], F)
    unb = O.Swarm(O.Problem(O.SPHERE, d, np.full(d, -np.inf, F), np.full(d, np.inf, F), u0=u0),
                  n, 31, sampler=O.SAMPLER_UNBOUNDED)
    lo = np.array([0.0, 2.0, -3.0, 0.5], F)
    hi = np.array([1.0, 4.0, 0.0, 1.0], F)
    uni = O.Swarm(O.Problem(O.SPHERE, d, lo, hi), n, 31)
    assert np.array_equal(unb.x, uni.x)
    assert (unb.x[1] >= 2) .all() and (unb.x[1] <= 4).all()
    assert (unb.x[0] >= 0).all() and (unb.x[0] < 1).all()


def test_given_positions_are_used_as_they_are():
    d, n = 3, 200
    rng = np.random.default_rng(4)
    pos = rng.uniform(-1, 10, (n, d)).astype(F)
    pr = O.Problem(O.ROSENBROCK, d, np.full(d, -1, F), np.full(d, 10, F), np.array([1, 100], F))
    s = O.Swarm(pr, n, 5, positions=pos)
    assert np.array_equal(s.x, pos.T)
    assert np.array_equal(s.v, np.zeros((d, n), F))
    f = O.evaluate(pr, pos)
    assert np.array_equal(s.cost, f) and np.array_equal(s.best_cost, f)
    assert s.gbest()[2] == int(np.argmin(f))


@pytest.mark.parametrize("N", [2, 3])
def test_reference_unbounded_thresholds(N):
    # test/regression.jl:147-199: Rosenbrock, x0 = 0, lb = ub = nothing, 2000
    # particles, 500 iterations: ParallelPSOKernel < 1e-4, ParallelSyncPSOKernel < 2e-2
    inf = np.full(N, np.inf, F)
    pr = O.Problem(O.ROSENBROCK, N, -inf, inf, np.array([1, 100], F), u0=np.zeros(N, F))
    q = O.Swarm(pr, 2000, 123 + N, sampler=O.SAMPLER_UNBOUNDED)
    q.run_qlock(500)
    s = O.Swarm(pr, 2000, 321 + N, sampler=O.SAMPLER_UNBOUNDED)
    s.run_sync(500)
    for sw, tol in ((q, 1e-4), (s, 2e-2)):
        pos = sw.gbest()[0][:N]
        assert float(O.evaluate(pr, pos[None, :])[0]) < tol
