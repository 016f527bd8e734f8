"""Independent pins of the oracle's async and queue-lock drivers.

`ora_async_run` and `ora_qlock_run` decide *which* attractor each step uses
and *when* a global best is published; the one-step move they call is pinned
separately (`test_move_against_fp64_formula`), and so is the objective.  Here
each driver is rebuilt in plain Python from those pinned pieces, following the
reference's control flow directly:

- async, `update_particle_states_async!` (src/kernels.jl:196-225): every
  particle starts from the launch's global best, steps `maxiters` times, and
  after each step replaces its private attractor by its own best when
  `best_cost < gbest.cost` (:215-218); `w *= wdamp` per step (:219).  The
  result is the argmin over particles of (best_position, best_cost)
  (src/lowerlevel_solve.jl:119-120).
- queue-lock, `update_particle_states!(..., ::ParallelPSOKernel, lock)`
  (src/kernels.jl:38-125): a block steps against the live global best
  (:69-80), queues the particles whose best cost beats the live global cost
  (:84-92), and only when more than one queued (`queue_num[1] > 1`, :97) takes
  the queue minimum (strict `<`, so the first queued of equal costs, :99-103)
  and publishes it if it still beats the global cost (:114-116).  Blocks are
  taken one after another in index order (the oracle's serialisation of the
  reference's concurrent blocks).

The brute force runs on tiny swarms (n = 4..8, d = 2..3, T <= 5) and must
agree with the oracle bit for bit; the scenario tests fix states by hand so
that exactly one or exactly two particles of a block improve.
"""
import numpy as np
import pytest


def _prob(O, obj, d, params=(), lb=-5.12, ub=5.12):
    return O.Problem(obj, d, np.full(d, lb, np.float32), np.full(d, ub, np.float32),
                     np.asarray(params, np.float32))


def _parts(s):
    return s.parts.copy()


def _step(O, P, part, gb, w, c1, c2, i, t, epoch, seed):
    """update_particle_state (src/kernels.jl:1-26) from the pinned move and
    objective: move, cost = f(x), and the particle's own best when improved."""
    d = P.d
    p = O.move(P, part, gb, float(w), float(c1), float(c2), i, t, epoch, seed)
    c = O.objective(P, p[:d])
    p[2 * d] = c
    if c < p[3 * d + 1]:
        p[2 * d + 1: 3 * d + 1] = p[:d]
        p[3 * d + 1] = c
    return p


def _argmin_best(parts, d):
    bc = parts[:, 3 * d + 1]
    i = int(np.argmin(bc))  # first occurrence: lowest index of equal costs
    return parts[i, 2 * d + 1: 3 * d + 1].copy(), np.float32(bc[i]), i


def brute_async(O, P, parts, gpos, gcost, T, w, wdamp, c1, c2, t0, epoch, seed):
    d = P.d
    parts = parts.copy()
    follows = 0
    for i in range(len(parts)):
        g, gc = gpos.copy(), np.float32(gcost)
        wk = np.float32(w)
        p = parts[i]
        for k in range(1, T + 1):
            p = _step(O, P, p, g, wk, c1, c2, i, t0 + k, epoch, seed)
            if p[3 * d + 1] < gc:
                g, gc = p[2 * d + 1: 3 * d + 1].copy(), np.float32(p[3 * d + 1])
                follows += 1
            wk = np.float32(wk * np.float32(wdamp))
        parts[i] = p
    return parts, _argmin_best(parts, d), follows


def brute_qlock(O, P, parts, gpos, gcost, gidx, T, w, wdamp, c1, c2, t0, epoch, seed, block,
                min_improvers=2):
    d = P.d
    n = len(parts)
    parts = parts.copy()
    gpos, gcost = gpos.copy(), np.float32(gcost)
    wk = np.float32(w)
    log = []  # (block start, queued count, published)
    for k in range(1, T + 1):
        for b0 in range(0, n, block):
            b1 = min(b0 + block, n)
            gb = gpos.copy()  # live global best as the block reads it
            for i in range(b0, b1):
                parts[i] = _step(O, P, parts[i], gb, wk, c1, c2, i, t0 + k, epoch, seed)
            queue = [i for i in range(b0, b1) if parts[i, 3 * d + 1] < gcost]
            published = False
            if len(queue) >= min_improvers:
                best = queue[0]
                for j in queue[1:]:
                    if parts[j, 3 * d + 1] < parts[best, 3 * d + 1]:
                        best = j
                if parts[best, 3 * d + 1] < gcost:
                    gpos = parts[best, 2 * d + 1: 3 * d + 1].copy()
                    gcost = np.float32(parts[best, 3 * d + 1])
                    gidx = best
                    published = True
            log.append((b0, len(queue), published))
        wk = np.float32(wk * np.float32(wdamp))
    return parts, (gpos, gcost, gidx), log


CASES = [  # (objective, d, params, n, T, seed, bad_gbest)
    ("ROSENBROCK", 2, (1, 100), 4, 3, 11, False),
    ("ROSENBROCK", 3, (1, 100), 6, 5, 12, True),
    ("RASTRIGIN", 2, (), 8, 4, 13, False),
    ("RASTRIGIN", 3, (), 5, 5, 14, True),
    ("SPHERE", 2, (), 7, 2, 15, True),
    ("ROSENBROCK", 2, (1, 100), 8, 5, 16, True),
]


def _setup(O, obj, d, params, n, seed, bad_gbest):
    P = _prob(O, getattr(O, obj), d, params)
    s = O.Swarm(P, n, seed, nthreads=1)
    if bad_gbest:
        # a launch attractor worse than every particle: the running best must
        # move off it at the first step
        corner = np.full(d, 5.12, np.float32)
        s.set_gbest(corner, O.objective(P, corner) + np.float32(1.0))
    return P, s


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("wdamp", [1.0, 0.97])
def test_async_driver_brute_force(oracle, case, wdamp):
    obj, d, params, n, T, seed, bad = case
    P, s = _setup(oracle, obj, d, params, n, seed, bad)
    gpos, gcost, _ = s.gbest()
    w, c1, c2 = 0.7298, 1.4962, 1.4962
    want, (wpos, wcost, widx), follows = brute_async(
        oracle, P, _parts(s), gpos, gcost, T, w, wdamp, c1, c2, s.iter, s.epoch, s.seed)
    s.run_async(T, w=w, wdamp=wdamp, c1=c1, c2=c2)
    np.testing.assert_array_equal(s.parts, want)
    pos, cost, idx = s.gbest()
    assert idx == widx and cost == wcost
    np.testing.assert_array_equal(pos, wpos)
    if bad:
        assert follows >= n  # every particle moved its attractor at least once


def test_async_attractor_is_followed(oracle):
    # Pure social pull (w = c1 = 0, c2 = 1) on a 1-D sphere with the launch
    # attractor at the upper bound: a particle that follows its own best is
    # pulled towards 0 after its first step; one that keeps the launch
    # attractor only ever moves up.
    P = _prob(oracle, oracle.SPHERE, 1, lb=-4.0, ub=4.0)
    s = oracle.Swarm(P, 6, 21, nthreads=1)
    x0 = np.array([[-3.0, -1.0, 0.5, 1.5, 2.5, 3.5]], np.float32)
    s.set_state(x0, np.zeros_like(x0), x0, oracle.evaluate(P, x0.T), 0)
    s.set_gbest(np.array([4.0], np.float32), np.float32(100.0))
    s.run_async(4, w=0.0, c1=0.0, c2=1.0)
    x = s.x[0]
    # step 1 moves every particle up (towards 4); from step 2 on each follows
    # its own best, which is its start (it only got worse), so it comes back
    # down: the end position lies below 4 and not above the step-1 position
    first = x0[0] + np.array([oracle.philox([i, 1, 0, 0], [21, 0])[1] >> 8
                              for i in range(6)], np.float64) * 2.0 ** -24 * (4.0 - x0[0])
    assert (x < np.minimum(first, 4.0) + 1e-6).all()
    assert (x < 4.0).all()
    # and the brute force agrees bit for bit
    s2 = oracle.Swarm(P, 6, 21, nthreads=1)
    s2.set_state(x0, np.zeros_like(x0), x0, oracle.evaluate(P, x0.T), 0)
    want, _, follows = brute_async(oracle, P, _parts(s2), np.array([4.0], np.float32),
                                   np.float32(100.0), 4, 0.0, 1.0, 0.0, 1.0, 0, 0, 21)
    np.testing.assert_array_equal(s.parts, want)
    assert follows >= 6


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("block,min_imp", [(1024, 2), (1024, 1), (2, 2), (3, 2), (4, 1)])
def test_qlock_driver_brute_force(oracle, case, block, min_imp):
    obj, d, params, n, T, seed, bad = case
    P, s = _setup(oracle, obj, d, params, n, seed, bad)
    gpos, gcost, gidx = s.gbest()
    w, wdamp, c1, c2 = 0.7298, 0.99, 1.4962, 1.4962
    want, (wpos, wcost, widx), log = brute_qlock(
        oracle, P, _parts(s), gpos, gcost, gidx, T, w, wdamp, c1, c2, s.iter, s.epoch, s.seed,
        block, min_imp)
    s.run_qlock(T, w=w, wdamp=wdamp, c1=c1, c2=c2, block=block, min_improvers=min_imp)
    np.testing.assert_array_equal(s.parts, want)
    pos, cost, idx = s.gbest()
    assert cost == wcost and idx == widx
    np.testing.assert_array_equal(pos, wpos)


def test_qlock_brute_force_covers_both_branches(oracle):
    # over the cases above, some blocks queue exactly one improver (and must
    # not publish) and some queue two or more (and publish)
    seen_one, seen_pub = 0, 0
    for case in CASES:
        obj, d, params, n, T, seed, bad = case
        P, s = _setup(oracle, obj, d, params, n, seed, bad)
        gpos, gcost, gidx = s.gbest()
        _, _, log = brute_qlock(oracle, P, _parts(s), gpos, gcost, gidx, T, 0.7298, 0.99,
                                1.4962, 1.4962, 0, 0, s.seed, 2)
        seen_one += sum(1 for _, q, p in log if q == 1 and not p)
        seen_pub += sum(1 for _, q, p in log if q >= 2 and p)
    assert seen_one > 0 and seen_pub > 0


# -- hand-built blocks: frozen swarm (w = c1 = c2 = 0), chosen best costs ------
def _frozen(oracle, xs, gcost, gpos=None):
    """Sphere swarm whose particles sit at `xs` (n x 2) with best cost +Inf, so
    one frozen step makes every best cost f(x); the global cost is `gcost`."""
    xs = np.asarray(xs, np.float32)
    n = len(xs)
    P = _prob(oracle, oracle.SPHERE, 2, lb=-8.0, ub=8.0)
    s = oracle.Swarm(P, n, 3, nthreads=1)
    x = xs.T.copy()
    s.set_state(x, np.zeros_like(x), x, np.full(n, np.inf, np.float32), 0)
    s.set_gbest(np.zeros(2, np.float32) if gpos is None else gpos, np.float32(gcost))
    s.gb_idx.value = 77  # a marker no particle index equals
    return s


def _qlock_frozen(s, block=1024, min_improvers=2):
    s.run_qlock(1, w=0.0, c1=0.0, c2=0.0, block=block, min_improvers=min_improvers)
    return s.gbest()


def test_qlock_one_improver_does_not_publish(oracle):
    # costs 0.25, 4, 9, 16 against a global cost of 1: one improver
    s = _frozen(oracle, [[0.5, 0], [2, 0], [3, 0], [4, 0]], 1.0)
    pos, cost, idx = _qlock_frozen(s)
    assert cost == np.float32(1.0) and idx == 77
    np.testing.assert_array_equal(s.best_cost, [0.25, 4, 9, 16])


def test_qlock_two_improvers_publish_the_minimum(oracle):
    # costs 4, 0.25, 9, 16 against 5: two improvers, the smaller is published
    s = _frozen(oracle, [[2, 0], [0.5, 0], [3, 0], [4, 0]], 5.0)
    pos, cost, idx = _qlock_frozen(s)
    assert cost == np.float32(0.25) and idx == 1
    np.testing.assert_array_equal(pos, [0.5, 0])


def test_qlock_single_improver_publishes_only_with_threshold_one(oracle):
    s = _frozen(oracle, [[0.5, 0], [2, 0], [3, 0], [4, 0]], 1.0)
    pos, cost, idx = _qlock_frozen(s, min_improvers=1)
    assert cost == np.float32(0.25) and idx == 0


def test_qlock_equal_costs_publish_the_first_queued(oracle):
    # (1, 0) and (0, -1) both cost 1 < 5: the first in particle order wins
    s = _frozen(oracle, [[3, 0], [1, 0], [0, -1], [4, 0]], 5.0)
    pos, cost, idx = _qlock_frozen(s)
    assert cost == np.float32(1.0) and idx == 1
    np.testing.assert_array_equal(pos, [1, 0])


def test_qlock_blocks_read_the_live_global_best(oracle):
    # blocks of 2, costs 0.25, 0.5 | 0.2, 0.4 against 5.  Block 0 publishes
    # 0.25; block 1 then sees the live 0.25, so only 0.2 queues: one improver,
    # no publication.  (Against the launch's 5 it would have queued two and
    # published 0.2.)
    s = _frozen(oracle, [[0.5, 0], [0, np.sqrt(0.5)], [0, np.sqrt(0.2)], [0, np.sqrt(0.4)]],
                5.0)
    pos, cost, idx = _qlock_frozen(s, block=2)
    assert idx == 0 and cost == np.float32(0.25)
    # the same particles as one block: all four queue, 0.2 is published
    s = _frozen(oracle, [[0.5, 0], [0, np.sqrt(0.5)], [0, np.sqrt(0.2)], [0, np.sqrt(0.4)]],
                5.0)
    pos, cost, idx = _qlock_frozen(s, block=1024)
    assert idx == 2


def test_qlock_later_block_steps_towards_the_published_best(oracle):
    # Blocks of 2 under a pure social pull (w = c1 = 0, c2 = 1) from a launch
    # attractor in the far corner: block 0 publishes, and block 1 must step
    # towards what block 0 published (the live global best), not towards the
    # launch attractor.  The brute force with the live read agrees bit for
    # bit; the same loop handing every block the launch attractor does not.
    P = _prob(oracle, oracle.SPHERE, 2, lb=-8.0, ub=8.0)
    x = np.array([[0.5, 2.0, 6.0, 7.0], [0.0, 0.0, 6.0, 7.0]], np.float32)
    far = np.array([-8.0, -8.0], np.float32)

    def fresh():
        s = oracle.Swarm(P, 4, 9, nthreads=1)
        s.set_state(x, np.zeros_like(x), x, oracle.evaluate(P, x.T), 0)
        s.set_gbest(far, np.float32(500.0))
        return s

    s = fresh()
    live, (gpos, gcost, gidx), log = brute_qlock(
        oracle, P, _parts(s), far, np.float32(500.0), 0, 1, 0.0, 1.0, 0.0, 1.0, 0, 0, 9, 2)
    assert log[0][2]  # block 0 published
    s.run_qlock(1, w=0.0, c1=0.0, c2=1.0, block=2)
    np.testing.assert_array_equal(s.parts, live)
    # block 0 alone is the same either way; block 1 is not
    snap = fresh().parts.copy()
    for i in range(4):
        snap[i] = _step(oracle, P, snap[i], far, 0.0, 0.0, 1.0, i, 1, 0, 9)
    np.testing.assert_array_equal(snap[:2], live[:2])
    assert not np.array_equal(snap[2:, :2], live[2:, :2])
