"""Helpers that build the same swarm on the GPU (libpps via its C ABI) and in
the CPU oracle, and compare their state."""
import sys

import numpy as np

import pps_oracle as O
import ppswarm as P
from ppswarm import _lib as L

PARAMS = {O.ROSENBROCK: (1.0, 100.0), O.ROSEN2: (1.0, 100.0), O.QUADRATIC: (1000.0,)}


def note(msg):
    """Name the case on stdout before a GPU step (flushed)."""
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def objective_marker(obj, obs=None, u0=(1.0, 1.0), dt=0.01, spo=10):
    m = {O.ROSENBROCK: P.Rosenbrock, O.RASTRIGIN: P.Rastrigin, O.SPHERE: P.Sphere,
         O.BEALE: P.Beale, O.BOOTH: P.Booth, O.HIMMELBLAU: P.Himmelblau, O.QUADRATIC: P.Quadratic,
         O.BOUNDARY_QUADRATIC: P.BoundaryQuadratic, O.ROSEN2: P.Rosen2}
    if obj == O.LOTKA_VOLTERRA:
        return P.lotka_volterra_sse(obs, u0=u0, dt=dt, steps_per_obs=spo)
    return m[obj]()


def make_pair(obj, d, n, seed, mode=L.MODE_SYNC, *, lb=-5.12, ub=5.12, gbest_from=0,
              launch=L.LAUNCH_AUTO, shift=None, clamp=0, obs=None, u0=(1.0, 1.0), dt=0.01,
              spo=10, min_improvers=2, nthreads=0, cons=None, grid_per_cu=0):
    lbv = np.broadcast_to(np.asarray(lb, np.float32), (d,)).copy()
    ubv = np.broadcast_to(np.asarray(ub, np.float32), (d,)).copy()
    params = PARAMS.get(obj, ())
    if obj == O.LOTKA_VOLTERRA:
        params = (u0[0], u0[1], dt, spo)
    f = objective_marker(obj, obs, u0, dt, spo)
    prob = P.OptimizationProblem(f, np.zeros(d, np.float32), params if params else None,
                                 lb=lbv, ub=ubv, shift=shift, cons=cons)
    ctx = P.context()
    sp = P.swarm_params(n, mode, seed=seed, gbest_from=gbest_from, clamp=clamp, launch=launch,
                        qlock_min_improvers=min_improvers, grid_per_cu=grid_per_cu)
    g = P.Swarm(ctx, prob, sp)
    oprob = O.Problem(obj, d, lbv, ubv, np.asarray(params, np.float32), shift, clamp,
                      None if obs is None else np.asarray(obs, np.float32),
                      None if cons is None else cons.rows(d))
    o = O.Swarm(oprob, n, seed, nthreads=nthreads)
    return g, o


def assert_same_state(g, o, what=""):
    x, v, b, bc, it = g.state()
    np.testing.assert_array_equal(x, o.x, err_msg=f"x {what}")
    np.testing.assert_array_equal(v, o.v, err_msg=f"v {what}")
    np.testing.assert_array_equal(b, o.best, err_msg=f"best {what}")
    np.testing.assert_array_equal(bc, o.best_cost, err_msg=f"best_cost {what}")
    assert it == o.iter, what


def assert_same_gbest(g, o, what="", check_pos=True):
    gb = g.gbest()
    pos, cost, idx = o.gbest()
    assert gb.index == idx, f"gbest index {gb.index} != {idx} {what}"
    np.testing.assert_array_equal(np.float32(gb.cost), cost, err_msg=f"gbest cost {what}")
    if check_pos:
        np.testing.assert_array_equal(gb.position, pos, err_msg=f"gbest pos {what}")
