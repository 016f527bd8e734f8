"""Parity of the HIP kernels with the CPU oracle, through the C ABI.

The bar is bit-exact: positions, velocities, best positions and best costs of
every particle, and the global-best index (ties to the lowest index), equal
the oracle's after every checked iteration.  The update is FP32 in a fixed
operation order without contraction on both sides, so no tolerance is needed;
a failure means the kernel computes something else.
"""
import numpy as np
import pytest

import pps_oracle as O
from parity_util import assert_same_gbest, assert_same_state, make_pair, note
from ppswarm import _lib as L

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("obj,d,n", [(O.ROSENBROCK, 3, 1000), (O.RASTRIGIN, 10, 5000),
                                     (O.SPHERE, 7, 333), (O.ROSENBROCK, 32, 700),
                                     (O.RASTRIGIN, 45, 300), (O.BEALE, 2, 100)])
def test_init_matches_oracle(obj, d, n):
    note(f"init obj={obj} d={d} n={n}")
    g, o = make_pair(obj, d, n, seed=1000 + d)
    assert_same_state(g, o, "after init")
    assert_same_gbest(g, o, "after init")
    g.reinit()
    o.reinit()
    assert_same_state(g, o, "after reinit")
    assert_same_gbest(g, o, "after reinit")


@pytest.mark.parametrize("gbest_from", [L.GBEST_PBEST, L.GBEST_CURRENT])
def test_sync_c0_workgroup_trajectory(gbest_from):
    # 3-D Rosenbrock, lb=-1, ub=10, 1000 particles, 500 iterations
    note(f"sync C0 workgroup gbest_from={gbest_from}")
    g, o = make_pair(O.ROSENBROCK, 3, 1000, seed=42, lb=-1.0, ub=10.0, gbest_from=gbest_from)
    done = 0
    for chunk in (1, 1, 8, 40, 450):
        g.run(chunk)
        o.run_sync(chunk, gbest_from=gbest_from)
        done += chunk
        assert_same_state(g, o, f"after {done} iterations")
        assert_same_gbest(g, o, f"after {done} iterations")


@pytest.mark.parametrize("launch", [L.LAUNCH_EAGER, L.LAUNCH_GRAPH])
@pytest.mark.parametrize("obj,d,n", [(O.RASTRIGIN, 10, 5000), (O.ROSENBROCK, 10, 3001),
                                     (O.ROSENBROCK, 5, 70000),
                                     # draws ahead of the state (d <= 16), winner held in
                                     # registers (d <= 32), and the 64-wide runtime path
                                     (O.SPHERE, 16, 2000), (O.RASTRIGIN, 32, 1500),
                                     (O.ROSENBROCK, 45, 1100)])
def test_sync_multiblock_trajectory(launch, obj, d, n):
    note(f"sync multiblock launch={launch} obj={obj} d={d} n={n}")
    g, o = make_pair(obj, d, n, seed=7 + n, launch=launch)
    done = 0
    for chunk in (1, 4, 15):
        g.run(chunk, 0.7298, 0.99)
        o.run_sync(chunk, 0.7298, 0.99)
        done += chunk
        assert_same_state(g, o, f"after {done} iterations")
        assert_same_gbest(g, o, f"after {done} iterations")


@pytest.mark.parametrize("obj,d,n", [(O.RASTRIGIN, 10, 5000), (O.ROSENBROCK, 3, 70000)])
def test_graph_replay_after_reinit(obj, d, n):
    # init -> solve! -> reinit! -> solve! with the iterations in replays of a
    # captured chunk graph (40 = a chunk of 32 + 8 launched one by one):
    # every run after a reinit must draw from the new epoch's random stream
    note(f"graph replay after reinit obj={obj} d={d} n={n}")
    g, o = make_pair(obj, d, n, seed=31 + n, launch=L.LAUNCH_GRAPH)
    for rnd in range(3):
        g.run(40)
        o.run_sync(40)
        assert_same_state(g, o, f"round {rnd}")
        assert_same_gbest(g, o, f"round {rnd}")
        g.reinit()
        o.reinit()
        assert_same_state(g, o, f"reinit {rnd}")
    g.run(40)
    o.run_sync(40)
    assert_same_state(g, o, "after the last reinit")
    assert_same_gbest(g, o, "after the last reinit")


@pytest.mark.parametrize("launch", [L.LAUNCH_AUTO, L.LAUNCH_GRAPH])
def test_chunk_graph_runs_match_oracle(launch):
    # graph launches replay a captured chunk of 32 iterations whose inertia
    # weights come from the run's table and launch the rest one by one; the
    # automatic mode captures from 128 iterations on (or once captured).  Runs
    # shorter than a chunk, exact multiples, remainders, new weights, a table
    # that outgrows its 1024 entries (the chunk is captured again) and a
    # reinit must all equal the oracle
    note(f"chunk graph runs launch={launch}")
    g, o = make_pair(O.RASTRIGIN, 10, 5000, seed=77, launch=launch)
    runs = [(5, 0.7298, 1.0), (40, 0.7298, 0.99), (70, 0.9, 0.995), (32, 0.7298, 1.0),
            (3, 0.5, 1.0), (130, 0.7298, 0.97), (33, 0.7298, 0.99), (1500, 0.8, 0.999),
            (64, 0.7298, 1.0), (100, 0.6, 0.98)]
    for rnd, (T, w, wd) in enumerate(runs):
        g.run(T, w, wd)
        o.run_sync(T, w, wd)
        assert_same_state(g, o, f"run {rnd} (T={T})")
        assert_same_gbest(g, o, f"run {rnd} (T={T})")
        if rnd == 5:
            g.reinit()
            o.reinit()


def test_chunk_graph_ode_pair_kernel():
    # the ODE estimation's two-lanes-per-particle step in replayed chunks
    rng = np.random.default_rng(4)
    obs = (rng.uniform(0.5, 3, (101, 2))).astype(np.float32)
    note("chunk graph LV pair kernel")
    g, o = make_pair(O.LOTKA_VOLTERRA, 4, 300, seed=23, lb=0.0, ub=5.0, obs=obs,
                     launch=L.LAUNCH_GRAPH)
    g.run(40, 0.7298, 0.99)
    o.run_sync(40, 0.7298, 0.99)
    assert_same_state(g, o, "after 40 iterations")
    assert_same_gbest(g, o)


def test_sync_current_mode_multiblock():
    note("sync multiblock gbest_from=CURRENT")
    g, o = make_pair(O.ROSENBROCK, 4, 4000, seed=99, lb=-1.0, ub=10.0,
                     gbest_from=L.GBEST_CURRENT, launch=L.LAUNCH_GRAPH)
    g.run(25)
    o.run_sync(25, gbest_from=O.GBEST_CURRENT)
    assert_same_state(g, o)
    assert_same_gbest(g, o)


@pytest.mark.parametrize("n", [500, 5000])
def test_one_step_from_identical_state(n):
    # arbitrary states, including non-finite velocities and best costs, fed to
    # both sides, then one step against the same attractor
    d = 6
    note(f"one step from state n={n}")
    g, o = make_pair(O.RASTRIGIN, d, n, seed=5, launch=L.LAUNCH_EAGER)
    rng = np.random.default_rng(n)
    x = rng.uniform(-5.12, 5.12, (d, n)).astype(np.float32)
    v = rng.normal(0, 3, (d, n)).astype(np.float32)
    v[0, :5] = [np.inf, -np.inf, np.nan, 1e30, -0.0]
    b = rng.uniform(-5.12, 5.12, (d, n)).astype(np.float32)
    bc = rng.uniform(0, 200, n).astype(np.float32)
    bc[5:8] = [np.nan, np.inf, 0.0]
    g.set_state(x, v, b, bc, 17)
    o.set_state(x, v, b, bc, 17)
    gb = rng.uniform(-1, 1, d).astype(np.float32)
    g.set_gbest(gb, 1.0)
    o.set_gbest(gb, 1.0)
    g.run(1, 0.5, 1.0)
    o.run_sync(1, 0.5, 1.0)
    assert_same_state(g, o, "after one step")
    assert_same_gbest(g, o, "after one step")


@pytest.mark.parametrize("obj,d,n,T", [(O.RASTRIGIN, 10, 3000, 60), (O.ROSENBROCK, 3, 1000, 200),
                                       (O.ROSENBROCK, 16, 2500, 30), (O.SPHERE, 20, 600, 40)])
def test_async_matches_oracle(obj, d, n, T):
    note(f"async obj={obj} d={d} n={n} T={T}")
    g, o = make_pair(obj, d, n, seed=11 * d, mode=L.MODE_ASYNC)
    g.run(T, 0.7298, 0.995)
    o.run_async(T, 0.7298, 0.995)
    assert_same_state(g, o, f"after {T} iterations")
    assert_same_gbest(g, o)
    # a second launch starts from the stored global best
    g.run(7)
    o.run_async(7)
    assert_same_state(g, o, "after the second launch")
    assert_same_gbest(g, o)


@pytest.mark.parametrize("min_improvers", [2, 1])
def test_qlock_single_block_matches_oracle(min_improvers):
    # one workgroup (n <= 1024): the live global best is deterministic
    note(f"qlock single block min_improvers={min_improvers}")
    g, o = make_pair(O.ROSENBROCK, 3, 1000, seed=3, mode=L.MODE_QLOCK, lb=-1.0, ub=10.0,
                     min_improvers=min_improvers)
    for chunk in (1, 9, 90):
        g.run(chunk)
        o.run_qlock(chunk, min_improvers=min_improvers)
        assert_same_state(g, o)
        assert_same_gbest(g, o)


ONE_WG = ([(L.MODE_SYNC, O.ROSENBROCK, 10, n) for n in (1, 8, 37, 64, 65, 130, 1024)]
          + [(L.MODE_SYNC, O.RASTRIGIN, 3, 50), (L.MODE_SYNC, O.SPHERE, 32, 64),
             (L.MODE_SYNC, O.RASTRIGIN, 45, 40), (L.MODE_SYNC, O.SPHERE, 8, 300)]
          + [(L.MODE_QLOCK, O.ROSENBROCK, 10, n) for n in (1, 8, 64, 65, 200)]
          + [(L.MODE_QLOCK, O.SPHERE, 32, 60), (L.MODE_QLOCK, O.RASTRIGIN, 45, 33),
             (L.MODE_QLOCK, O.RASTRIGIN, 45, 300), (L.MODE_QLOCK, O.SPHERE, 16, 700)])


@pytest.mark.parametrize("variant", [0, 1])
@pytest.mark.parametrize("mode,obj,d,n", ONE_WG)
def test_one_workgroup_sizes_match_oracle(mode, obj, d, n, variant):
    # a swarm of at most 1024 particles runs all iterations in one workgroup
    # of ceil(n / 64) wavefronts; one wavefront takes each iteration's winner
    # from its lane, more exchange keys (and counts) through LDS.  variant:
    # the global best's source (sync) or the publication threshold (queue-lock)
    qlock = mode == L.MODE_QLOCK
    gf = (L.GBEST_PBEST, L.GBEST_CURRENT)[variant]
    mi = (2, 1)[variant]
    note(f"one workgroup mode={mode} obj={obj} d={d} n={n} variant={variant}")
    g, o = make_pair(obj, d, n, seed=31 + n + d, mode=mode, gbest_from=0 if qlock else gf,
                     min_improvers=mi,
                     launch=L.LAUNCH_AUTO if qlock else L.LAUNCH_WORKGROUP)
    done = 0
    for chunk in (1, 7, 52):
        g.run(chunk, 0.7298, 0.995)
        if qlock:
            o.run_qlock(chunk, 0.7298, 0.995, min_improvers=mi)
        else:
            o.run_sync(chunk, 0.7298, 0.995, gbest_from=gf)
        done += chunk
        assert_same_state(g, o, f"after {done} iterations")
        assert_same_gbest(g, o, f"after {done} iterations")


@pytest.mark.parametrize("mode", [L.MODE_SYNC, L.MODE_QLOCK])
@pytest.mark.parametrize("n", [37, 64, 200])
def test_one_workgroup_in_loop_argmin_ties(mode, n):
    # the winner an iteration hands to the next: ties go to the lowest index
    # (also +0.0 against -0.0); Sphere costs are >= 0, so with these best
    # costs no particle improves and the tie decides both the next
    # iteration's attractor and the final record
    qlock = mode == L.MODE_QLOCK
    d = 3
    note(f"one workgroup ties mode={mode} n={n}")
    g, o = make_pair(O.SPHERE, d, n, seed=11, mode=mode, lb=-1.0, ub=1.0,
                     launch=L.LAUNCH_AUTO if qlock else L.LAUNCH_WORKGROUP)
    rng = np.random.default_rng(n)
    x = rng.uniform(-1, 1, (d, n)).astype(np.float32)
    v = rng.normal(0, 0.3, (d, n)).astype(np.float32)
    b = rng.uniform(-1, 1, (d, n)).astype(np.float32)
    for lo, hi, clo, chi, rest in [(5, 33, -1.0, -1.0, -0.5), (0, n - 1, -1.0, -1.0, -0.5),
                                   (n // 2, n - 1, 0.0, -0.0, 1e-30),
                                   (n - 2, n - 1, -0.0, 0.0, 1e-30)]:
        bc = np.full(n, rest, np.float32)
        bc[lo] = clo
        bc[hi] = chi
        g.set_state(x, v, b, bc, 3)
        o.set_state(x, v, b, bc, 3)
        gb = rng.uniform(-1, 1, d).astype(np.float32)
        g.set_gbest(gb, 1.0)
        o.set_gbest(gb, 1.0)
        g.run(2, 0.5, 1.0)
        if qlock:
            o.run_qlock(2, 0.5, 1.0)
        else:
            o.run_sync(2, 0.5, 1.0)
        assert_same_state(g, o, f"ties at {lo}, {hi}")
        assert_same_gbest(g, o, f"ties at {lo}, {hi}")
        assert g.gbest().index == lo


def test_qlock_multiblock_invariants():
    # several workgroups race on the live global best, as in the reference:
    # check what every schedule guarantees
    d, n = 4, 20000
    note("qlock multiblock invariants")
    g, o = make_pair(O.ROSENBROCK, d, n, seed=8, mode=L.MODE_QLOCK, lb=-1.0, ub=10.0)
    prev = g.gbest().cost
    for _ in range(20):
        g.run(5)
        gb = g.gbest()
        x, v, b, bc, it = g.state()
        assert gb.cost <= prev
        assert bc.min() <= gb.cost
        assert (x >= -1).all() and (x <= 10).all()
        # the published global best is a real best position of its particle
        # (or an earlier one of the same particle with a higher cost)
        assert bc[gb.index] <= gb.cost
        prev = gb.cost


def test_lotka_volterra_sync_with_lexicographic_clamp():
    rng = np.random.default_rng(0)
    obs = (rng.uniform(0.5, 3, (101, 2))).astype(np.float32)
    note("LV sync lexicographic clamp")
    g, o = make_pair(O.LOTKA_VOLTERRA, 4, 300, seed=21, lb=0.0, ub=5.0, obs=obs,
                     clamp=L.CLAMP_LEXICOGRAPHIC, launch=L.LAUNCH_EAGER)
    assert_same_state(g, o, "after init")
    g.run(3)
    o.run_sync(3)
    assert_same_state(g, o, "after 3 iterations")
    assert_same_gbest(g, o)


def test_shifted_objective_parity():
    d, n = 8, 2048
    shift = np.linspace(-1, 1, d).astype(np.float32)
    note("shifted rastrigin parity")
    g, o = make_pair(O.RASTRIGIN, d, n, seed=77, shift=shift, launch=L.LAUNCH_GRAPH)
    g.run(10)
    o.run_sync(10)
    assert_same_state(g, o)
    assert_same_gbest(g, o)


TUNED = [(O.RASTRIGIN, 10, 300000, ppw, g) for ppw in (256, 512, 1024) for g in (1, 3, 8)] + [
    (O.ROSENBROCK, 32, 50000, 512, 2), (O.ROSENBROCK, 32, 50000, 256, 6),
    (O.SPHERE, 3, 70000, 1024, 4), (O.ROSENBROCK, 16, 40000, 1024, 2),
    (O.RASTRIGIN, 2, 200000, 512, 8)]


@pytest.mark.parametrize("obj,d,n,ppw,grid_per_cu", TUNED)
def test_tuned_launch_does_not_change_results(obj, d, n, ppw, grid_per_cu):
    # every choice the autotuner can make (particles per workgroup, and the
    # grid, which sets the tiles each thread steps) gives the same swarm
    note(f"tuned launch obj={obj} d={d} n={n} ppw={ppw} grid={grid_per_cu}")
    import ppswarm as P

    lb, ub = np.full(d, -5.12, np.float32), np.full(d, 5.12, np.float32)
    f = {O.RASTRIGIN: P.Rastrigin, O.ROSENBROCK: P.Rosenbrock, O.SPHERE: P.Sphere}[obj]()
    par = [1.0, 100.0] if obj == O.ROSENBROCK else None
    prob = P.OptimizationProblem(f, np.zeros(d, np.float32), par, lb=lb, ub=ub)
    g = P.Swarm(P.context(), prob, P.swarm_params(n, L.MODE_SYNC, seed=9, launch=L.LAUNCH_GRAPH,
                                                  grid_per_cu=grid_per_cu,
                                                  particles_per_workgroup=ppw))
    o = O.Swarm(O.Problem(obj, d, lb, ub, np.asarray(par or [], np.float32)), n, 9)
    assert_same_state(g, o, "init")
    g.run(6, 0.7298, 0.99)
    o.run_sync(6, 0.7298, 0.99)
    assert_same_state(g, o)
    assert_same_gbest(g, o)
    g.reinit()
    o.reinit()
    g.run(3)
    o.run_sync(3)
    assert_same_state(g, o, "after reinit")
    assert_same_gbest(g, o, "after reinit")


def test_workgroup_sizes_outside_the_compiled_set_are_refused():
    note("particles_per_workgroup refusals")
    import ppswarm as P

    def prob(f, d, par=None):
        return P.OptimizationProblem(f, np.zeros(d, np.float32), par, lb=np.full(d, -1.0),
                                     ub=np.full(d, 1.0))
    for pr, ppw in ((prob(P.Rastrigin(), 10), 384),        # not a compiled size
                    (prob(P.Rastrigin(), 10), 64),         # the ODE objectives' size
                    (prob(P.Rosenbrock(), 32, [1, 100]), 1024),  # d = 32: up to 512
                    (prob(P.Rastrigin(), 5), 512),          # runtime-dimension kernels
                    (prob(P.Beale(), 2), 512)):             # two-dimensional test functions
        with pytest.raises(P.PPSError):
            P.Swarm(P.context(), pr, P.swarm_params(5000, L.MODE_SYNC, seed=1,
                                                    launch=L.LAUNCH_GRAPH,
                                                    particles_per_workgroup=ppw))


def test_autotune_records_a_choice(tmp_path, monkeypatch):
    note("autotune")
    import ppswarm as P

    d, n = 8, 1 << 16
    prob = P.OptimizationProblem(P.Sphere(), np.zeros(d, np.float32), lb=np.full(d, -1.0),
                                 ub=np.full(d, 1.0))
    cache = tmp_path / "tune.txt"
    monkeypatch.setenv("PPS_TUNE_CACHE", str(cache))
    g, b, us = P.autotune_sync(prob, n, iters=5)
    assert g in (1, 2, 3, 4, 6, 8) and b in (256, 512, 1024) and us > 0
    assert P.tuned_launch(P.Sphere(), d, n) == (g, b)
    assert P.tuned_launch(P.Sphere(), d, 4 * n) is None
    # the choice is appended to the cache file: objective d ceil(log2 n) grid block
    line = cache.read_text().split()
    assert [int(v) for v in line] == [L.OBJ_SPHERE, d, 16, g, b]
    # a later swarm of that (objective, d, size bucket) runs with it, same results
    sw = P.Swarm(P.context(), prob, P.swarm_params(n, L.MODE_SYNC, seed=4, launch=L.LAUNCH_GRAPH))
    o = O.Swarm(O.Problem(O.SPHERE, d, np.full(d, -1.0, np.float32), np.full(d, 1.0, np.float32)),
                n, 4)
    sw.run(4)
    o.run_sync(4)
    assert_same_state(sw, o)


def test_autotune_ode_objective_keeps_its_workgroups():
    # the ODE objectives run 64-particle workgroups (two lanes per particle);
    # only their grid is tuned
    note("autotune ode")
    import json
    import os
    import ppswarm as P

    doc = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "lv_obs.json")))
    f = P.lotka_volterra_sse(np.array(doc["obs"], np.float32))
    prob = P.OptimizationProblem(f, np.zeros(4, np.float32), lb=np.zeros(4), ub=np.full(4, 5.0))
    g, b, us = P.autotune_sync(prob, 1 << 14, iters=3)
    assert b == 64 and g in (1, 2, 3, 4, 6, 8) and us > 0
    assert P.tuned_launch(f, 4, 1 << 14) == (g, b)


@pytest.mark.parametrize("n,launch", [(5000, L.LAUNCH_GRAPH), (1000, L.LAUNCH_WORKGROUP)])
def test_argmin_at_every_lane(n, launch):
    # the global-best argmin (wavefront DPP reduction, then across waves and
    # workgroups) must find a single minimum wherever it sits: every lane of
    # the first wavefront, lanes of later wavefronts and workgroups, the last
    # particle; ties go to the lowest index (also +0.0 against -0.0), and
    # NaN costs never win
    note(f"argmin at every lane n={n} launch={launch}")
    d = 2
    g, _ = make_pair(O.SPHERE, d, n, seed=3, launch=launch)
    rng = np.random.default_rng(n)
    x = rng.uniform(-1, 1, (d, n)).astype(np.float32)
    v = np.zeros((d, n), np.float32)
    b = rng.uniform(-1, 1, (d, n)).astype(np.float32)
    positions = list(range(64)) + [64 * k + (7 * k) % 64 for k in range(1, n // 64)] + [n - 1]
    for p in positions:
        bc = np.ones(n, np.float32)
        bc[p] = 0.5
        g.set_state(x, v, b, bc, 3)
        gb = g.gbest()
        assert gb.index == p and gb.cost == 0.5, (p, gb.index, gb.cost)
        np.testing.assert_array_equal(gb.position, b[:, p])
    for a, c, nan_below in [(5, 40, 3), (17, 63, 0), (63, 64, 0), (100, n - 1, 64), (0, n - 1, 0),
                            (1, 2, 2)]:
        bc = np.full(n, 2.0, np.float32)
        bc[c] = 0.25
        bc[a] = 0.25
        bc[:nan_below] = np.nan
        g.set_state(x, v, b, bc, 3)
        gb = g.gbest()
        assert gb.index == int(np.nanargmin(bc)), (a, c, nan_below, gb.index)
    bc = np.full(n, 1.0, np.float32)
    bc[n - 1] = 0.0
    bc[n // 2] = -0.0
    g.set_state(x, v, b, bc, 3)
    assert g.gbest().index == n // 2


@pytest.mark.parametrize("c1,c2", [(0.0, 1.4962), (2.0 ** -102, -2.5), (3.0, 1e-20),
                                   (-1.25, 7.0)])
@pytest.mark.parametrize("launch", [L.LAUNCH_WORKGROUP, L.LAUNCH_GRAPH])
def test_step_coefficients(c1, c2, launch):
    # the kernels fold the draws' 2^-24 scale into c1 and c2; bit for bit the
    # oracle's c * u01(u) for every coefficient the library accepts
    note(f"coefficients c1={c1} c2={c2} launch={launch}")
    import ppswarm as P

    d, n = 5, 900 if launch == L.LAUNCH_WORKGROUP else 3000
    lb, ub = np.full(d, -5.12, np.float32), np.full(d, 5.12, np.float32)
    prob = P.OptimizationProblem(P.Rastrigin(), np.zeros(d, np.float32), lb=lb, ub=ub)
    g = P.Swarm(P.context(), prob, P.swarm_params(n, L.MODE_SYNC, seed=41, launch=launch,
                                                  c1=c1, c2=c2))
    o = O.Swarm(O.Problem(O.RASTRIGIN, d, lb, ub), n, 41)
    for chunk in (1, 6):
        g.run(chunk)
        o.run_sync(chunk, c1=c1, c2=c2)
        assert_same_state(g, o, f"after {chunk}")
        assert_same_gbest(g, o, f"after {chunk}")


def test_tiny_step_coefficients_are_refused():
    note("coefficient refusals")
    import ctypes as C
    import ppswarm as P

    d = 3
    prob = P.OptimizationProblem(P.Rastrigin(), np.zeros(d, np.float32), lb=np.full(d, -1.0),
                                 ub=np.full(d, 1.0))
    for c1, c2 in ((1e-35, 1.0), (1.0, -3e-32)):
        with pytest.raises(P.PPSError):
            P.Swarm(P.context(), prob, P.swarm_params(2000, L.MODE_SYNC, seed=1, c1=c1, c2=c2))
    sv = P.Server()
    try:
        pr, keep = prob.c()
        sp = P.swarm_params(2000, L.MODE_SYNC, seed=1, c1=1e-35)
        t = C.c_uint64()
        rc = L.lib().pps_server_submit(sv.h, C.byref(pr), C.byref(sp), 10, 0.7298, 1.0,
                                       C.byref(t))
        assert rc == L.E_INVALID
    finally:
        sv.close()


def test_served_step_coefficients_equal_the_plain_path():
    # a served problem with other coefficients ends where pps_swarm_init +
    # pps_swarm_run end with them
    note("served coefficients")
    import ctypes as C
    import ppswarm as P

    d, n, T = 6, 4000, 40
    prob = P.OptimizationProblem(P.Rosenbrock(), np.zeros(d, np.float32), [1.0, 100.0],
                                 lb=np.full(d, -2.0), ub=np.full(d, 2.0))
    sp = P.swarm_params(n, L.MODE_SYNC, seed=9, c1=0.75, c2=2.0 ** -102)
    plain = P.Swarm(P.context(), prob, sp)
    plain.run(T)
    ref = plain.gbest()
    plain.free()
    sv = P.Server()
    try:
        pr, keep = prob.c()
        t = C.c_uint64()
        L.check(L.lib().pps_server_submit(sv.h, C.byref(pr), C.byref(sp), T, 0.7298, 1.0,
                                          C.byref(t)))
        res = sv.wait(int(t.value))
    finally:
        sv.close()
    assert res.index == ref.index
    assert np.float32(res.cost_f32) == np.float32(ref.cost)
    np.testing.assert_array_equal(np.asarray(res.u, np.float32), ref.position)
