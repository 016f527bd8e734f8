"""Pins of the oracle's constraint penalty (CPU).

The reference adds calc_penalty (src/utils.jl:320-372) to the cost through
handle_constraints (src/kernels.jl:28-36).  Each expected value below is derived
by hand from the reference's text, not from the oracle's code:
  theta(q) = 10 (q < 0.001), 20 (q <= 0.1), 100 (q <= 1), 300; the thresholds are
  Float64 literals, so Float32(0.001) > 0.001 and Float32(0.1) > 0.1 after
  promotion; gamma(q) = 1 (q < 1) else 2; penalty = h(iter) * (sum theta + sum
  theta * q^gamma); h = sqrt; `iter` is 1 at init, the 1-based particle index + 1
  in the sync and queue-lock kernels, the 1-based iteration of the launch + 1 in
  the async kernel.
"""
import numpy as np
import pytest

import pps_oracle as O
from ppswarm.api import QuadraticConstraints

F = np.float32

# test/constraints.jl:11-14
REF_CONS = QuadraticConstraints(linear=[[-1, 2], [1, -2], [0, 0]],
                                quadratic=[[0, 0], [0, 0], [0.25, 1]], const=[-1, 1, -1])


def ref_problem():
    # objective (x1 - p1)^2 + (x2 - p2)^2, p = (2, 1): Sphere shifted by p
    return O.Problem(O.SPHERE, 2, np.zeros(2, F), np.full(2, 2, F), shift=np.array([2, 1], F),
                     cons=REF_CONS.rows(2))


def one_dim(c=0.0):
    # g(x) = x + c on [-5, 5], objective x^2
    return O.Problem(O.SPHERE, 1, np.full(1, -5, F), np.full(1, 5, F),
                     cons=np.array([[1, 0, c]], F))


@pytest.mark.parametrize("q,expected", [
    (0.0, 10.0),                                   # theta 10, gamma 1
    (-3.0, 10.0),                                  # q = max(g, 0) = 0
    (0.0005, F(10.0) + F(10.0) * F(0.0005)),
    (F(0.001), F(20.0) + F(20.0) * F(0.001)),      # Float32(0.001) is above 0.001
    (0.05, F(20.0) + F(20.0) * F(0.05)),
    (F(0.1), F(100.0) + F(100.0) * F(0.1)),        # Float32(0.1) is above 0.1
    (0.5, 150.0),
    (1.0, 200.0),                                  # theta 100, gamma 2
    (2.0, 1500.0),                                 # 300 + 300 * 2^2
])
def test_theta_gamma_schedule(q, expected):
    assert O.penalty(one_dim(), np.array([q], F), 1.0) == F(expected)


def test_nan_constraint_gives_nan_penalty():
    assert np.isnan(O.penalty(one_dim(), np.array([np.nan], F), 1.0))


def test_reference_problem_at_x0():
    # x0 = (1, 1): g = (0, 0, 1/4 + 1 - 1) -> q = (0, 0, 0.25): 10 + 10 + 100 + 100 * 0.25
    pr = ref_problem()
    x = np.array([1, 1], F)
    assert [O.constraint(pr, k, x) for k in range(3)] == [F(0), F(0), F(0.25)]
    assert O.penalty(pr, x, 1.0) == F(145)
    assert O.objective(pr, x) == F(1)            # (1-2)^2 + (1-1)^2
    assert O.cost(pr, x, 1.0) == F(146)


def test_constraint_evaluation_order_matches_the_reference_expressions():
    # -x[1] + 2 * x[2] - 1  and  (x[1]^2) / 4 + x[2]^2 - 1 in FP32, Julia's order
    pr = ref_problem()
    rng = np.random.default_rng(3)
    for _ in range(2000):
        x = rng.uniform(0, 2, 2).astype(F)
        g1 = (-x[0] + F(2) * x[1]) - F(1)
        g2 = (x[0] - F(2) * x[1]) + F(1)
        g3 = ((x[0] * x[0]) / F(4) + x[1] * x[1]) - F(1)
        got = [O.constraint(pr, k, x) for k in range(3)]
        assert got[0] == g1 and got[1] == g2 and got[2] == g3


def test_h_is_sqrt_of_iter():
    pr = one_dim(-0.5)
    x = np.array([1.0], F)                          # q = 0.5 -> 150
    for it in (1, 2, 3, 7, 1000, 2 ** 24 + 3):
        assert O.penalty(pr, x, float(F(it))) == np.sqrt(F(it), dtype=F) * F(150)


def _stationary_swarm(n, seed=1):
    # every particle at x* = 1 with v = 0 and best = global best = x*: a step
    # leaves it in place, so its new cost shows h's argument at the call site
    pr = one_dim(-0.5)
    s = O.Swarm(pr, n, seed)
    xs = np.ones((1, n), F)
    s.set_state(xs, np.zeros((1, n), F), xs, np.full(n, 1e30, F), 0)
    s.set_gbest(np.ones(1, F), 1e30)
    return s


def test_sync_step_passes_the_particle_index():
    n = 37
    s = _stationary_swarm(n)
    s.run_sync(1)
    i = np.arange(n)
    expected = F(1) + np.sqrt((i + 2).astype(F), dtype=F) * F(150)
    assert np.array_equal(s.x, np.ones((1, n), F))
    assert np.array_equal(s.cost, expected)


def test_qlock_step_passes_the_particle_index():
    n = 37
    s = _stationary_swarm(n)
    s.run_qlock(1)
    expected = F(1) + np.sqrt((np.arange(n) + 2).astype(F), dtype=F) * F(150)
    assert np.array_equal(s.cost, expected)


def test_async_passes_the_launch_iteration():
    n, T = 9, 6
    s = _stationary_swarm(n)
    s.run_async(T)
    # last iteration k = T: h(T + 1); the best is the first iteration's h(2)
    assert np.array_equal(s.cost, np.full(n, F(1) + np.sqrt(F(T + 1), dtype=F) * F(150)))
    assert np.array_equal(s.best_cost, np.full(n, F(1) + np.sqrt(F(2), dtype=F) * F(150)))
    # a second launch restarts the count (the reference's `for iter in 1:maxiters`)
    s.set_state(s.x, s.v, s.x, np.full(n, 1e30, F), s.iter)
    s.run_async(1)
    assert np.array_equal(s.cost, np.full(n, F(1) + np.sqrt(F(2), dtype=F) * F(150)))


def test_init_uses_h_of_one():
    pr = ref_problem()
    s = O.Swarm(pr, 64, 5)
    x = s.x
    for i in range(64):
        xi = x[:, i].copy()
        assert s.cost[i] == O.objective(pr, xi) + O.penalty(pr, xi, 1.0)
        assert O.penalty(pr, xi, 1.0) != O.penalty(pr, xi, 2.0) or O.penalty(pr, xi, 1.0) == 0
