"""ODE parameter estimation on seeded random settings against the oracle, bit
for bit: Lotka-Volterra or FitzHugh-Nagumo, each of the five solvers (RK4,
Tsit5 fixed and adaptive, and both with samples read from the interpolant),
step sizes, steps per sample, tolerances, the step limit, the number of
samples, sync (any launch shape and global-best source) or async, both clamps
and two boxes of rates."""
import json
import os

import numpy as np
import pytest

import pps_oracle as O
import ppswarm as P
from parity_util import assert_same_gbest, assert_same_state, note
from ppswarm import _lib as L

pytestmark = pytest.mark.gpu

DOC = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "lv_obs.json")))
OBS = np.array(DOC["obs"], np.float32)
SOLVERS = ("rk4", "tsit5", "atsit5", "tsit5_saveat", "atsit5_saveat")


def _cases(count=30, seed=90210):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        model = str(rng.choice(["lotka_volterra", "fitzhugh_nagumo"]))
        solver = str(rng.choice(SOLVERS))
        kw = {"ode_alg": "rk4" if solver == "rk4" else "tsit5",
              "adaptive": solver.startswith("a")}
        if solver.endswith("saveat"):
            kw["saveat"] = float(rng.choice([0.1, 0.05]))
            kw["dt"] = float(rng.choice([0.03, 0.1])) if not kw["adaptive"] else 0.01
        else:
            kw["dt"] = float(rng.choice([0.01, 0.02]))
            kw["steps_per_obs"] = int(rng.choice([5, 10]))
        if kw["adaptive"]:
            kw["abstol"] = float(rng.choice([1e-6, 1e-5]))
            kw["reltol"] = float(rng.choice([1e-3, 1e-2]))
            kw["maxiters"] = int(rng.choice([10000, 300]))
        n_obs = int(rng.choice([11, 51, 101]))
        mode = int(rng.choice([L.MODE_SYNC, L.MODE_SYNC, L.MODE_ASYNC]))
        n = int(rng.integers(2, 1500))
        launch = int(rng.choice([L.LAUNCH_AUTO, L.LAUNCH_GRAPH, L.LAUNCH_EAGER]))
        grid = int(rng.choice([0, 1, 2, 4]))
        gbest_from = int(rng.choice([L.GBEST_PBEST, L.GBEST_CURRENT]))
        clamp = int(rng.choice([L.CLAMP_LEXICOGRAPHIC, L.CLAMP_COMPONENT]))
        box = [(0.0, 5.0), (0.2, 3.0)][int(rng.integers(0, 2))]
        chunks = (1, int(rng.integers(1, 3)))
        out.append((k, model, solver, kw, n_obs, mode, n, launch, grid, gbest_from, clamp, box,
                    chunks, int(rng.integers(1, 1 << 30))))
    return out


@pytest.mark.parametrize("case", _cases(), ids=lambda c: f"case{c[0]}")
def test_random_ode_estimation(case):
    (k, model, solver, kw, n_obs, mode, n, launch, grid, gbest_from, clamp, box, chunks,
     seed) = case
    note(f"random ode {k}: {model} {solver} {kw} n_obs={n_obs} mode={mode} n={n} "
         f"launch={launch} grid={grid} gbest_from={gbest_from} clamp={clamp} box={box} "
         f"chunks={chunks} seed={seed}")
    obs = OBS[:n_obs]
    f = P.ode_sse(model, obs, **kw)
    lb, ub = np.full(4, box[0], np.float32), np.full(4, box[1], np.float32)
    prob = P.OptimizationProblem(f, np.zeros(4), lb=lb, ub=ub)
    sync = mode == L.MODE_SYNC
    g = P.Swarm(P.context(), prob,
                P.swarm_params(n, mode, seed=seed, clamp=clamp,
                               gbest_from=gbest_from if sync else L.GBEST_PBEST,
                               launch=launch if sync else L.LAUNCH_AUTO,
                               grid_per_cu=grid if sync else 0))
    o = O.Swarm(O.Problem(f.code, 4, lb, ub, f.params(), clamp=clamp, obs=obs), n, seed)
    assert_same_state(g, o, "init")
    for chunk in chunks:
        g.run(chunk)
        if sync:
            o.run_sync(chunk, gbest_from=gbest_from)
        else:
            o.run_async(chunk)
        assert_same_state(g, o, f"after {chunk}")
        assert_same_gbest(g, o, f"after {chunk}")
    g.free()
