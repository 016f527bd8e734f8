"""ctypes wrapper of the CPU oracle (oracle/_build/libpps_oracle.so).

TEST INFRASTRUCTURE ONLY.  The oracle restates ParallelParticleSwarms.jl's
particle-swarm path on the CPU (see pps_oracle.c for the file:line map); the
parity tests and the smoke check compare the kernels' results with it, and the
bench's CPU-baseline leg times it.  The product never imports this module.

Particle state crosses this wrapper in the kernels' SoA convention
(x, v, best: d x n float32; best_cost: n float32) and is converted to and from
the oracle's AoS SPSOParticle layout here.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libpps_oracle.so")

MAXD = 64
MAXPAR = 8

ROSENBROCK, RASTRIGIN, SPHERE, BEALE, BOOTH, HIMMELBLAU, QUADRATIC, BOUNDARY_QUADRATIC, ROSEN2 = (
    1, 2, 3, 4, 5, 6, 7, 8, 9)
LOTKA_VOLTERRA = 16
LOTKA_VOLTERRA_TSIT5, LOTKA_VOLTERRA_ATSIT5 = 17, 18
LOTKA_VOLTERRA_TSIT5_SAVEAT, LOTKA_VOLTERRA_ATSIT5_SAVEAT = 19, 20
(FITZHUGH_NAGUMO, FITZHUGH_NAGUMO_TSIT5, FITZHUGH_NAGUMO_ATSIT5, FITZHUGH_NAGUMO_TSIT5_SAVEAT,
 FITZHUGH_NAGUMO_ATSIT5_SAVEAT) = 21, 22, 23, 24, 25
GBEST_PBEST, GBEST_CURRENT = 0, 1
SAMPLER_UNIFORM, SAMPLER_UNBOUNDED, SAMPLER_LATIN_HYPERCUBE = 0, 1, 2
CLAMP_COMPONENT, CLAMP_LEXICOGRAPHIC = 0, 1


class _Prob(C.Structure):
    _fields_ = [
        ("obj", C.c_int32), ("d", C.c_int32), ("has_shift", C.c_int32), ("clamp", C.c_int32),
        ("lb", C.c_float * MAXD), ("ub", C.c_float * MAXD), ("shift", C.c_float * MAXD),
        ("par", C.c_float * MAXPAR), ("obs", C.POINTER(C.c_float)), ("n_obs", C.c_int32),
        ("steps_per_obs", C.c_int32), ("cons", C.POINTER(C.c_float)), ("n_cons", C.c_int32),
        ("u0", C.c_float * MAXD),
    ]


LS_ARMIJO, LS_STRONG_WOLFE = 0, 1


class _LbfgsParams(C.Structure):
    _fields_ = [("maxiters", C.c_int32), ("history", C.c_int32), ("abstol", C.c_float),
                ("armijo_c", C.c_float), ("armijo_rho", C.c_float),
                ("max_backtracks", C.c_int32), ("linesearch", C.c_int32),
                ("wolfe_c2", C.c_float), ("alpha_init", C.c_float), ("alpha_max", C.c_float),
                ("ls_maxiters", C.c_int32), ("zoom_maxiters", C.c_int32),
                ("reltol", C.c_float)]


def lbfgs_params(maxiters=100, history=10, abstol=0.0, armijo_c=1e-4, armijo_rho=0.5,
                 max_backtracks=20, linesearch=LS_ARMIJO, wolfe_c2=0.9, alpha_init=1.0,
                 alpha_max=65536.0, ls_maxiters=20, zoom_maxiters=30, reltol=0.0):
    """Local-search settings.  linesearch LS_ARMIJO (armijo_c, armijo_rho,
    max_backtracks) or LS_STRONG_WOLFE (c1 = armijo_c, c2 = wolfe_c2,
    alpha_init, alpha_max, ls_maxiters, zoom_maxiters); reltol > 0 also stops
    after a step with |f - f_new| <= reltol |f|."""
    return _LbfgsParams(int(maxiters), int(history), float(abstol), float(armijo_c),
                        float(armijo_rho), int(max_backtracks), int(linesearch),
                        float(wolfe_c2), float(alpha_init), float(alpha_max), int(ls_maxiters),
                        int(zoom_maxiters), float(reltol))


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run `make -C oracle`)")
        L = C.CDLL(LIB_PATH)
        fp = C.POINTER(C.c_float)
        i64p = C.POINTER(C.c_int64)
        L.ora_philox.argtypes = [C.POINTER(C.c_uint32)] * 3
        L.ora_cos2pi.argtypes = [C.c_float]
        L.ora_cos2pi.restype = C.c_float
        L.ora_sin2pi.argtypes = [C.c_float]
        L.ora_sin2pi.restype = C.c_float
        L.ora_objective.argtypes = [C.POINTER(_Prob), fp]
        L.ora_objective.restype = C.c_float
        L.ora_constraint.argtypes = [C.POINTER(_Prob), C.c_int, fp]
        L.ora_constraint.restype = C.c_float
        L.ora_penalty.argtypes = [C.POINTER(_Prob), fp, C.c_float]
        L.ora_penalty.restype = C.c_float
        L.ora_cost.argtypes = [C.POINTER(_Prob), fp, C.c_float]
        L.ora_cost.restype = C.c_float
        L.ora_ode_sse_rk4.argtypes = [fp, C.POINTER(_Prob)]
        L.ora_root5.argtypes = [C.c_float]
        L.ora_root5.restype = C.c_float
        L.ora_tsit5_tableau.argtypes = [fp, fp]
        L.ora_tsit5_interp_weights.argtypes = [C.c_float, fp]
        L.ora_ode_sse_rk4.restype = C.c_float
        L.ora_eval.argtypes = [C.POINTER(_Prob), C.c_int64, fp, fp]
        L.ora_argmin.argtypes = [fp, C.c_int64, C.c_int]
        L.ora_argmin.restype = C.c_int64
        L.ora_init.argtypes = [C.POINTER(_Prob), C.c_int64, C.c_uint64, C.c_uint32, fp, fp, fp,
                               i64p, C.c_int, C.c_int32, fp]
        L.ora_lhs_permute.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64]
        L.ora_lhs_permute.restype = C.c_uint32
        L.ora_sync_run.argtypes = [C.POINTER(_Prob), C.c_int64, C.c_uint64, C.c_uint32,
                                   C.c_uint32, C.c_int32, C.c_float, C.c_float, C.c_float,
                                   C.c_float, C.c_int32, fp, fp, fp, i64p, C.c_int]
        L.ora_async_run.argtypes = [C.POINTER(_Prob), C.c_int64, C.c_uint64, C.c_uint32,
                                    C.c_uint32, C.c_int32, C.c_float, C.c_float, C.c_float,
                                    C.c_float, fp, fp, fp, i64p, C.c_int]
        L.ora_qlock_run.argtypes = [C.POINTER(_Prob), C.c_int64, C.c_uint64, C.c_uint32,
                                    C.c_uint32, C.c_int32, C.c_float, C.c_float, C.c_float,
                                    C.c_float, C.c_int32, C.c_int32, fp, fp, fp, i64p]
        L.ora_move.argtypes = [C.POINTER(_Prob), fp, fp, C.c_float, C.c_float, C.c_float,
                               C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64]
        L.ora_num_threads.restype = C.c_int
        L.ora_lbfgs_one.argtypes = [C.POINTER(_Prob), C.POINTER(_LbfgsParams), fp]
        L.ora_lbfgs_one.restype = C.c_float
        L.ora_lbfgs_run.argtypes = [C.POINTER(_Prob), C.POINTER(_LbfgsParams), C.c_int64, fp, fp,
                                    fp, C.c_int]
        L.ora_wolfe_search.argtypes = [C.POINTER(_Prob), C.POINTER(_LbfgsParams), C.c_int, fp,
                                       C.c_float, fp, fp, fp, fp, fp, fp, fp, fp]
        L.ora_wolfe_search.restype = C.c_int
        L.ora_value_grad.argtypes = [C.POINTER(_Prob), fp, fp, fp]
        _lib = L
    return _lib


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


@dataclass
class Problem:
    """Objective + bounds, mirroring the product's pps_problem."""

    obj: int
    d: int
    lb: np.ndarray
    ub: np.ndarray
    params: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))
    shift: np.ndarray | None = None
    clamp: int = CLAMP_COMPONENT
    obs: np.ndarray | None = None  # Lotka-Volterra observations, (n_obs, 2)
    cons: np.ndarray | None = None  # constraints, (n_cons, 2d + 1) rows [a, q, c]
    u0: np.ndarray | None = None  # centre of the unbounded sampler

    def c(self) -> _Prob:
        p = _Prob()
        p.obj, p.d, p.clamp = self.obj, self.d, self.clamp
        lb = np.asarray(self.lb, np.float32).reshape(-1)
        ub = np.asarray(self.ub, np.float32).reshape(-1)
        for j in range(self.d):
            p.lb[j], p.ub[j] = float(lb[j]), float(ub[j])
        par = np.asarray(self.params, np.float32).reshape(-1)
        for k in range(len(par)):
            p.par[k] = float(par[k])
        if self.shift is not None:
            p.has_shift = 1
            sh = np.asarray(self.shift, np.float32).reshape(-1)
            for j in range(self.d):
                p.shift[j] = float(sh[j])
        if self.obs is not None:
            self._obs_keep = np.ascontiguousarray(self.obs, np.float32).reshape(-1)
            p.obs = _fp(self._obs_keep)
            p.n_obs = len(self._obs_keep) // 2
            p.steps_per_obs = int(par[3])
        if self.u0 is not None:
            u0 = np.asarray(self.u0, np.float32).reshape(-1)
            for j in range(self.d):
                p.u0[j] = float(u0[j])
        if self.cons is not None:
            self._cons_keep = np.ascontiguousarray(self.cons, np.float32).reshape(-1)
            p.cons = _fp(self._cons_keep)
            p.n_cons = len(self._cons_keep) // (2 * self.d + 1)
        return p


def constraint(prob: Problem, k: int, x) -> np.float32:
    x = np.ascontiguousarray(x, np.float32)
    return np.float32(lib().ora_constraint(C.byref(prob.c()), int(k), _fp(x)))


def penalty(prob: Problem, x, iter_arg: float) -> np.float32:
    x = np.ascontiguousarray(x, np.float32)
    return np.float32(lib().ora_penalty(C.byref(prob.c()), _fp(x), float(iter_arg)))


def cost(prob: Problem, x, iter_arg: float) -> np.float32:
    x = np.ascontiguousarray(x, np.float32)
    return np.float32(lib().ora_cost(C.byref(prob.c()), _fp(x), float(iter_arg)))


def root5(x: float) -> np.float32:
    return np.float32(lib().ora_root5(float(np.float32(x))))


def tsit5_tableau():
    """(a 7x6, btilde 7) as the oracle stores them (float32)."""
    a = np.zeros((7, 6), np.float32)
    e = np.zeros(7, np.float32)
    lib().ora_tsit5_tableau(_fp(a), _fp(e))
    return a, e


def tsit5_interp_weights(theta: float) -> np.ndarray:
    """b_i(theta) of Tsit5's continuous extension as the oracle evaluates them."""
    b = np.zeros(7, np.float32)
    lib().ora_tsit5_interp_weights(float(np.float32(theta)), _fp(b))
    return b


def lhs_permute(i: int, n: int, j: int, epoch: int, seed: int) -> int:
    return int(lib().ora_lhs_permute(int(i), int(n), int(j), int(epoch), int(seed)))


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*[int(v) & 0xFFFFFFFF for v in ctr])
    k = (C.c_uint32 * 2)(*[int(v) & 0xFFFFFFFF for v in key])
    o = (C.c_uint32 * 4)()
    lib().ora_philox(c, k, o)
    return [o[i] for i in range(4)]


def cos2pi(x: float) -> float:
    return float(lib().ora_cos2pi(float(x)))


def sin2pi(x: float) -> float:
    return float(lib().ora_sin2pi(float(x)))


def objective(prob: Problem, x) -> np.float32:
    x = np.ascontiguousarray(x, np.float32)
    p = prob.c()
    return np.float32(lib().ora_objective(C.byref(p), _fp(x)))


def evaluate(prob: Problem, xs) -> np.ndarray:
    xs = np.ascontiguousarray(xs, np.float32).reshape(-1, prob.d)
    out = np.zeros(len(xs), np.float32)
    p = prob.c()
    lib().ora_eval(C.byref(p), len(xs), _fp(xs), _fp(out))
    return out


class Swarm:
    """Oracle swarm in AoS SPSOParticle layout ((3d+2) float32 per particle)."""

    def __init__(self, prob: Problem, n: int, seed: int, nthreads: int = 0,
                 sampler: int = SAMPLER_UNIFORM, positions=None):
        self.prob, self.n, self.seed = prob, int(n), int(seed)
        self.d = prob.d
        self.nthreads = nthreads or lib().ora_num_threads()
        self.parts = np.zeros((self.n, 3 * self.d + 2), np.float32)
        self.gb_pos = np.zeros(MAXD, np.float32)
        self.gb_cost = C.c_float(0.0)
        self.gb_idx = C.c_int64(0)
        self.epoch = 0
        self.iter = 0
        self._p = prob.c()
        self.sampler = int(sampler)
        self.init(positions)

    # -- views in the kernels' SoA convention
    @property
    def x(self):
        return self.parts[:, : self.d].T.copy()

    @property
    def v(self):
        return self.parts[:, self.d: 2 * self.d].T.copy()

    @property
    def cost(self):
        return self.parts[:, 2 * self.d].copy()

    @property
    def best(self):
        return self.parts[:, 2 * self.d + 1: 3 * self.d + 1].T.copy()

    @property
    def best_cost(self):
        return self.parts[:, 3 * self.d + 1].copy()

    def set_state(self, x, v, best, best_cost, it: int):
        d = self.d
        self.parts[:, :d] = np.asarray(x, np.float32).T
        self.parts[:, d: 2 * d] = np.asarray(v, np.float32).T
        self.parts[:, 2 * d + 1: 3 * d + 1] = np.asarray(best, np.float32).T
        self.parts[:, 3 * d + 1] = np.asarray(best_cost, np.float32)
        self.iter = int(it)

    def gbest(self):
        return self.gb_pos[: self.d].copy(), np.float32(self.gb_cost.value), int(self.gb_idx.value)

    def set_gbest(self, pos, cost):
        self.gb_pos[: self.d] = np.asarray(pos, np.float32)
        self.gb_cost.value = float(cost)

    def init(self, positions=None):
        pos = None
        if positions is not None:
            pos = np.ascontiguousarray(positions, np.float32).reshape(self.n, self.d)
        lib().ora_init(C.byref(self._p), self.n, self.seed, self.epoch, _fp(self.parts),
                       _fp(self.gb_pos), C.byref(self.gb_cost), C.byref(self.gb_idx),
                       self.nthreads, self.sampler, None if pos is None else _fp(pos))
        self.iter = 0

    def reinit(self, positions=None):
        self.epoch += 1
        self.init(positions)

    def run_sync(self, iters, w=0.7298, wdamp=1.0, c1=1.4962, c2=1.4962, gbest_from=GBEST_PBEST):
        lib().ora_sync_run(C.byref(self._p), self.n, self.seed, self.epoch, self.iter, int(iters),
                           w, wdamp, c1, c2, gbest_from, _fp(self.parts), _fp(self.gb_pos),
                           C.byref(self.gb_cost), C.byref(self.gb_idx), self.nthreads)
        self.iter += int(iters)

    def run_async(self, iters, w=0.7298, wdamp=1.0, c1=1.4962, c2=1.4962):
        lib().ora_async_run(C.byref(self._p), self.n, self.seed, self.epoch, self.iter,
                            int(iters), w, wdamp, c1, c2, _fp(self.parts), _fp(self.gb_pos),
                            C.byref(self.gb_cost), C.byref(self.gb_idx), self.nthreads)
        self.iter += int(iters)

    def run_qlock(self, iters, w=0.7298, wdamp=1.0, c1=1.4962, c2=1.4962, block=1024,
                  min_improvers=2):
        lib().ora_qlock_run(C.byref(self._p), self.n, self.seed, self.epoch, self.iter,
                            int(iters), w, wdamp, c1, c2, block, min_improvers,
                            _fp(self.parts), _fp(self.gb_pos), C.byref(self.gb_cost),
                            C.byref(self.gb_idx))
        self.iter += int(iters)


def move(prob: Problem, particle, gb, w, c1, c2, i, t, epoch, seed):
    """One velocity/position move of a single AoS particle (no evaluation)."""
    p = np.ascontiguousarray(particle, np.float32).copy()
    g = np.zeros(MAXD, np.float32)
    g[: prob.d] = np.asarray(gb, np.float32)
    pc = prob.c()
    lib().ora_move(C.byref(pc), _fp(p), _fp(g), w, c1, c2, i, t, epoch, seed)
    return p


def lbfgs_one(prob: Problem, x0, params=None):
    """Projected L-BFGS from x0; returns (x, f)."""
    x = np.zeros(MAXD, np.float32)
    x[: prob.d] = np.asarray(x0, np.float32)
    p = prob.c()
    lp = params or lbfgs_params()
    f = lib().ora_lbfgs_one(C.byref(p), C.byref(lp), _fp(x))
    return x[: prob.d].copy(), np.float32(f)


def group_size(d: int) -> int:
    G = 1
    while G < d:
        G <<= 1
    return G


def wolfe_search(prob: Problem, x, p, params=None):
    """One Strong-Wolfe search from x along p (d floats each) inside the
    problem's bounds: (ok, alpha, xn, fn, gn, f, g) with f, g at x."""
    d = prob.d
    G = group_size(d)
    lp = params or lbfgs_params(linesearch=LS_STRONG_WOLFE)
    xs = np.zeros(MAXD, np.float32)
    ps = np.zeros(MAXD, np.float32)
    lo = np.zeros(MAXD, np.float32)
    hi = np.zeros(MAXD, np.float32)
    xs[:d], ps[:d] = np.asarray(x, np.float32), np.asarray(p, np.float32)
    lo[:d], hi[:d] = np.asarray(prob.lb, np.float32), np.asarray(prob.ub, np.float32)
    f, g = value_grad(prob, xs[:d])
    gs = np.zeros(MAXD, np.float32)
    gs[:d] = g
    xn = np.zeros(MAXD, np.float32)
    gn = np.zeros(MAXD, np.float32)
    fn = C.c_float()
    a = C.c_float()
    pc = prob.c()
    ok = lib().ora_wolfe_search(C.byref(pc), C.byref(lp), G, _fp(xs), float(f), _fp(gs), _fp(ps),
                                _fp(lo), _fp(hi), _fp(xn), C.byref(fn), _fp(gn), C.byref(a))
    return bool(ok), np.float32(a.value), xn[:d].copy(), np.float32(fn.value), gn[:d].copy(), f, g


def value_grad(prob: Problem, x):
    """FP32 value and analytic gradient the local search uses at x."""
    d = prob.d
    xs = np.zeros(MAXD, np.float32)
    xs[:d] = np.asarray(x, np.float32)
    g = np.zeros(MAXD, np.float32)
    f = C.c_float()
    pc = prob.c()
    lib().ora_value_grad(C.byref(pc), _fp(xs), C.byref(f), _fp(g))
    return np.float32(f.value), g[:d].copy()


def lbfgs_swarm(swarm: "Swarm", params=None):
    """L-BFGS from every particle's best position: (x n x d, f n)."""
    n, d = swarm.n, swarm.d
    ox = np.zeros((n, d), np.float32)
    of = np.zeros(n, np.float32)
    lp = params or lbfgs_params()
    lib().ora_lbfgs_run(C.byref(swarm._p), C.byref(lp), n, _fp(swarm.parts), _fp(ox), _fp(of),
                        swarm.nthreads)
    return ox, of
