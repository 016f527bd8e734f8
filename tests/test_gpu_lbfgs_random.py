"""The local search on seeded random settings against the oracle, bit for bit
(every particle's position and value, the argmin): objective (Rosenbrock,
Rastrigin, sphere at d = 1..64, and the fixed two-dimensional test functions),
history m = 1..16 (both sides of the kernels' history capacities 10 and 16),
iteration count, Armijo (c, rho, backtracks) or Strong-Wolfe (c1, c2, first
and largest trial step, bracketing and zoom iterations), abstol, reltol,
bounds and a shifted optimum."""
import numpy as np
import pytest

import pps_oracle as O
import ppswarm as P
from parity_util import PARAMS, note
from ppswarm import _lib as L

pytestmark = pytest.mark.gpu

FIXED_2D = (O.BEALE, O.BOOTH, O.HIMMELBLAU, O.QUADRATIC, O.ROSEN2)


def _cases(count=60, seed=4242):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        if rng.random() < 0.2:
            obj, d = int(rng.choice(FIXED_2D)), 2
        else:
            obj = int(rng.choice([O.ROSENBROCK, O.RASTRIGIN, O.SPHERE]))
            d = int(rng.choice([1, 2, 3, 5, 8, 11, 16, 23, 32, 40, 64]))
            if obj == O.ROSENBROCK and d == 1:
                d = 2
        n = int(rng.integers(1, 1500))
        history = int(rng.integers(1, 17))
        maxiters = int(rng.integers(1, 41))
        wolfe = bool(rng.integers(0, 2))
        if wolfe:
            ls = dict(c1=float(rng.choice([1e-4, 1e-3])), c2=float(rng.choice([0.9, 0.5, 0.1])),
                      alpha_init=float(rng.choice([1.0, 0.5, 2.0])),
                      alpha_max=float(rng.choice([65536.0, 8.0])),
                      maxiters=int(rng.integers(2, 21)), zoom_maxiters=int(rng.integers(1, 31)))
        else:
            ls = dict(c=float(rng.choice([1e-4, 1e-2])), rho=float(rng.choice([0.5, 0.3])),
                      max_backtracks=int(rng.integers(1, 21)))
        abstol = float(rng.choice([0.0, 1e-5, 1e-2]))
        reltol = float(rng.choice([0.0, 0.0, 1e-6, 1e-3]))
        half = float(rng.choice([1.0, 3.0, 5.0]))
        shifted = obj not in FIXED_2D and bool(rng.integers(0, 2))
        out.append((k, obj, d, n, history, maxiters, wolfe, ls, abstol, reltol, half, shifted,
                    int(rng.integers(1, 1 << 30))))
    return out


@pytest.mark.parametrize("case", _cases(), ids=lambda c: f"case{c[0]}")
def test_random_local_search(case):
    k, obj, d, n, history, maxiters, wolfe, ls, abstol, reltol, half, shifted, seed = case
    note(f"random lbfgs {k}: obj={obj} d={d} n={n} m={history} iters={maxiters} "
         f"wolfe={wolfe} ls={ls} abstol={abstol} reltol={reltol} half={half} "
         f"shift={shifted} seed={seed}")
    lb = np.full(d, -half, np.float32)
    ub = np.full(d, half, np.float32)
    shift = (np.random.default_rng(seed).uniform(-0.5, 0.5, d).astype(np.float32)
             if shifted else None)
    params = PARAMS.get(obj, ())
    markers = {O.ROSENBROCK: P.Rosenbrock, O.RASTRIGIN: P.Rastrigin, O.SPHERE: P.Sphere,
               O.BEALE: P.Beale, O.BOOTH: P.Booth, O.HIMMELBLAU: P.Himmelblau,
               O.QUADRATIC: P.Quadratic, O.ROSEN2: P.Rosen2}
    prob = P.OptimizationProblem(markers[obj](), np.zeros(d, np.float32),
                                 params if params else None, lb=lb, ub=ub, shift=shift)
    g = P.Swarm(P.context(), prob, P.swarm_params(n, L.MODE_SYNC, seed=seed))
    o = O.Swarm(O.Problem(obj, d, lb, ub, np.asarray(params, np.float32), shift), n, seed)
    g.run(3)
    o.run_sync(3)
    if wolfe:
        marker = P.StrongWolfeLineSearch(**ls)
        op = O.lbfgs_params(maxiters, history, abstol, armijo_c=ls["c1"],
                            linesearch=O.LS_STRONG_WOLFE, wolfe_c2=ls["c2"],
                            alpha_init=ls["alpha_init"], alpha_max=ls["alpha_max"],
                            ls_maxiters=ls["maxiters"], zoom_maxiters=ls["zoom_maxiters"],
                            reltol=reltol)
    else:
        marker = P.ArmijoBacktracking(**ls)
        op = O.lbfgs_params(maxiters, history, abstol, armijo_c=ls["c"], armijo_rho=ls["rho"],
                            max_backtracks=ls["max_backtracks"], reltol=reltol)
    gx, gf, bx, bf, bi = g.lbfgs(maxiters=maxiters, history=history, abstol=abstol,
                                 linesearch=marker, reltol=reltol)
    ox, of = O.lbfgs_swarm(o, op)
    np.testing.assert_array_equal(gf, of)
    np.testing.assert_array_equal(gx, ox)
    i = int(np.argmin(of))
    assert bi == i and bf == of[i]
    np.testing.assert_array_equal(bx, ox[i])
    g.free()
