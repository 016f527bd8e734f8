"""The two-lanes-per-particle ODE kernel writes both Lotka-Volterra
derivatives in one form, f = X * own + (Y * u1) * u2, with the sign of the
reference's subtracted term folded into Y (even lane: X = p1, Y = -p2; odd
lane: X = -p3, Y = p4).  This pins, in IEEE FP32 arithmetic on the host, that
the folded form equals the reference form bit for bit:

    p1 u1 - (p2 u1) u2  ==  p1 u1 + ((-p2) u1) u2

(negating a factor negates the rounded product exactly; a - b is a + (-b)),
over random values spanning many binades, signed zeros, infinities and values
near overflow.  NaN results must agree as NaN (their sign bit is not
observable in the losses: a NaN cost never wins a comparison).  A planted
regrouping, p1 u1 - p2 (u1 u2), must fail the same check.
"""
import numpy as np

F = np.float32


def _samples(rng, n):
    mag = F(10.0) ** rng.uniform(-30, 30, n).astype(F)
    vals = (mag * rng.choice([F(-1), F(1)], n)).astype(F)
    special = np.array([0.0, -0.0, np.inf, -np.inf, 3.0e38, -3.0e38, 1e-45, -1e-45, 1.0, -1.0],
                       F)
    return np.concatenate([vals, np.resize(special, n // 8)])


def _same(a, b):
    a = np.asarray(a, F)
    b = np.asarray(b, F)
    nan = np.isnan(a)
    assert np.array_equal(nan, np.isnan(b))
    assert np.array_equal(a[~nan].view(np.uint32), b[~nan].view(np.uint32))


def _reference_f1(p1, p2, u1, u2):
    return (p1 * u1) - ((p2 * u1) * u2)


def _folded(X, Y, own, u1, u2):
    return (X * own) + ((Y * u1) * u2)


def test_lotka_volterra_even_lane_form_is_exact():
    rng = np.random.default_rng(7)
    n = 200_000
    p1, p2, u1, u2 = (_samples(rng, n) for _ in range(4))
    m = min(len(p1), len(p2), len(u1), len(u2))
    p1, p2, u1, u2 = p1[:m], p2[:m], u1[:m], u2[:m]
    with np.errstate(all="ignore"):
        _same(_reference_f1(p1, p2, u1, u2), _folded(p1, -p2, u1, u1, u2))


def test_lotka_volterra_odd_lane_form_is_the_reference_expression():
    rng = np.random.default_rng(8)
    n = 200_000
    p3, p4, u1, u2 = (_samples(rng, n) for _ in range(4))
    m = min(len(p3), len(p4), len(u1), len(u2))
    p3, p4, u1, u2 = p3[:m], p4[:m], u1[:m], u2[:m]
    with np.errstate(all="ignore"):
        ref = ((-p3) * u2) + ((p4 * u1) * u2)
        _same(ref, _folded(-p3, p4, u2, u1, u2))


def test_planted_regrouping_is_caught():
    rng = np.random.default_rng(9)
    p1, p2, u1, u2 = (rng.uniform(0.1, 5.0, 10_000).astype(F) for _ in range(4))
    regrouped = (p1 * u1) - (p2 * (u1 * u2))
    ref = _reference_f1(p1, p2, u1, u2)
    assert not np.array_equal(ref.view(np.uint32), regrouped.view(np.uint32))
