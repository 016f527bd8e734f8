"""Host-side mirror of ParallelParticleSwarms.jl's public interface over libpps.

Names and argument meanings follow the reference (v1.4.0):

* algorithm descriptors ``ParallelPSOKernel(num_particles; global_update,
  backend)``, ``ParallelSyncPSOKernel(num_particles; backend)``,
  ``HybridPSO(; backend, pso, local_opt)``, ``SimpleLBFGS()``
  (src/algorithms.jl:78-121,192-249);
* ``init`` / ``solve_`` (Julia ``solve!``) / ``reinit_`` (``reinit!``) /
  ``solve`` (src/init.jl:56-184, src/solve.jl:11-67);
* the solution fields ``u``, ``objective``, ``original``, ``stats.time``,
  ``retcode`` (src/solve.jl:28-33,53-58).

The backend is ``MI355XBackend()``: the swarm lives in HBM on one GPU and every
iteration runs in the HIP kernels of libpps.  Objectives are a closed set of
markers (``Rosenbrock()``, ``Rastrigin()``, ...) because an arbitrary Python or
Julia callable cannot cross the C ABI into a kernel; anything else raises
``ValueError`` (the reference compiles any closure into its kernel,
src/kernels.jl:31-33).
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field
from typing import Any, Optional

import numpy as np

from . import _lib as L

# ---------------------------------------------------------------------------
# random seeding (stands in for Julia's Random.seed! + task-local RNG)
_seed_state = {"base": 0x5EED_0F_5EED, "count": 0}


def seed_(s: int) -> None:
    """Reset the seed sequence that ``init`` draws swarm seeds from."""
    _seed_state["base"] = int(s) & 0xFFFFFFFFFFFFFFFF
    _seed_state["count"] = 0


def _splitmix64(z: int) -> int:
    z = (z + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return z ^ (z >> 31)


def _next_seed() -> int:
    s = _splitmix64(_seed_state["base"] + _seed_state["count"])
    _seed_state["count"] += 1
    return s


# ---------------------------------------------------------------------------
# backend and contexts
@dataclass(frozen=True)
class MI355XBackend:
    """One MI355X (gfx950) device, addressed by HIP device ordinal."""

    device: int = 0


_contexts: dict[int, "Context"] = {}


class Context:
    def __init__(self, device: int = 0):
        h = C.c_void_p()
        L.check(L.lib().pps_ctx_create(device, C.byref(h)))
        self.h = h
        self.device = device

    def info(self) -> dict:
        cu, clk = C.c_int32(), C.c_int32()
        tot, fr = C.c_int64(), C.c_int64()
        L.check(L.lib().pps_ctx_info(self.h, C.byref(cu), C.byref(clk), C.byref(tot), C.byref(fr)))
        return {"num_cus": cu.value, "clock_khz": clk.value, "total_mem": tot.value,
                "free_mem": fr.value}

    def close(self):
        if self.h:
            L.lib().pps_ctx_destroy(self.h)
            self.h = None


def context(backend: MI355XBackend = MI355XBackend()) -> Context:
    c = _contexts.get(backend.device)
    if c is None:
        c = _contexts[backend.device] = Context(backend.device)
    return c


# ---------------------------------------------------------------------------
# objectives
@dataclass(frozen=True)
class Objective:
    code: int
    name: str
    fixed_dim: Optional[int] = None
    n_params: int = 0

    def __call__(self, x, p=None):
        """FP64 host evaluation (the reference's ``prob.f(u, p)``)."""
        x = np.asarray(x, np.float64).reshape(1, -1)
        lb = np.full(x.shape[1], -np.inf)
        prob = OptimizationProblem(self, np.zeros(x.shape[1], np.float32), p, lb=lb, ub=-lb,
                                   _validate=False)
        return float(eval_f64(prob, x)[0])


def Rosenbrock():
    """Σ_{i<d} p[2](x[i+1]-x[i]^2)^2 + (p[1]-x[i])^2 (test/gpu.jl:14-20)."""
    return Objective(L.OBJ_ROSENBROCK, "rosenbrock", None, 2)


def Rastrigin():
    """10 d + Σ (x_i^2 - 10 cos(2π x_i))."""
    return Objective(L.OBJ_RASTRIGIN, "rastrigin")


def Sphere():
    return Objective(L.OBJ_SPHERE, "sphere")


def Beale():
    return Objective(L.OBJ_BEALE, "beale", 2)


def Booth():
    return Objective(L.OBJ_BOOTH, "booth", 2)


def Himmelblau():
    return Objective(L.OBJ_HIMMELBLAU, "himmelblau", 2)


def Quadratic():
    """p[1] (x1 - 1)^2 + (x2 + 2)^2 (test/lbfgs.jl:32)."""
    return Objective(L.OBJ_QUADRATIC, "quadratic", 2, 1)


def BoundaryQuadratic():
    """(x1 + 1)^2 (test/lbfgs.jl:33)."""
    return Objective(L.OBJ_BOUNDARY_QUADRATIC, "boundary_quadratic", 1)


def Rosen2():
    """(p[1] - x1)^2 + p[2] (x2 - x1^2)^2 (test/lbfgs.jl:34)."""
    return Objective(L.OBJ_ROSEN2, "rosen2", 2, 2)


@dataclass(frozen=True)
class OdeSSE(Objective):
    """Sum of squared errors of an FP32 ODE trajectory against observations
    (the loss of src/ode_pso.jl:47-52) for a two-state model with four rates:
    Lotka-Volterra (α, β, γ, δ; benchmarks/Lotka_Volterra/lotka_volterra.jl:9-13)
    or FitzHugh-Nagumo (a, b, τ⁻¹, l; benchmarks/Fitzhugh_Nagumo/ode_estimation.jl:
    6-12).  ``obs`` is (n_obs, 2), sample k taken at t = k * dt *
    steps_per_obs starting from ``u0`` at sample 0.  Solved in the kernel with
    fixed-step RK4 (``ode_alg="rk4"``), fixed-step Tsit5 (``"tsit5"``, the
    reference's GPUTsit5 with Val(false)) or adaptive Tsit5 (``"tsit5",
    adaptive=True``, Val(true): ``dt`` is the initial step, ``abstol`` /
    ``reltol`` the tolerances, ``maxiters`` the step limit of one evaluation,
    beyond which the loss is +Inf).  With ``saveat`` (Tsit5 only) sample k is
    at t = k * saveat and steps of ``dt`` run independently of the samples,
    which are read from Tsit5's continuous extension, as GPUTsit5 fills
    ``saveat`` (benchmarks/Lotka_Volterra/lotka_volterra.jl:124-127: dt = 0.1,
    samples every 0.01); ``steps_per_obs`` is then unused."""

    obs: Any = None
    u0: tuple = (1.0, 1.0)
    dt: float = 0.01
    steps_per_obs: int = 10
    adaptive: bool = False
    abstol: float = 1e-6
    reltol: float = 1e-3
    maxiters: int = 10000
    saveat: Optional[float] = None
    model: str = "lotka_volterra"

    def params(self) -> np.ndarray:
        if self.saveat is not None:
            return np.array([self.u0[0], self.u0[1], self.saveat, self.dt, self.abstol,
                             self.reltol, self.maxiters], np.float32)
        if self.adaptive:
            obs_dt = np.float32(np.float64(self.dt) * self.steps_per_obs)
            return np.array([self.u0[0], self.u0[1], obs_dt, self.dt, self.abstol, self.reltol,
                             self.maxiters], np.float32)
        return np.array([self.u0[0], self.u0[1], self.dt, self.steps_per_obs], np.float32)


LotkaVolterraSSE = OdeSSE


def ode_sse(model, obs, u0=(1.0, 1.0), dt=0.01, steps_per_obs=10, *, ode_alg="rk4",
            adaptive=False, abstol=1e-6, reltol=1e-3, maxiters=10000, saveat=None) -> OdeSSE:
    """The SSE objective of ``model`` ("lotka_volterra" or "fitzhugh_nagumo")."""
    if model not in L.ODE_MODEL_BASE:
        raise ValueError(f"model is one of {sorted(L.ODE_MODEL_BASE)}, got {model!r}")
    obs = np.ascontiguousarray(obs, np.float32).reshape(-1, 2)
    if ode_alg not in ("rk4", "tsit5"):
        raise ValueError(f"ode_alg is 'rk4' or 'tsit5', got {ode_alg!r}")
    if adaptive and ode_alg != "tsit5":
        raise ValueError("adaptive stepping uses Tsit5 (ode_alg='tsit5')")
    if saveat is not None and ode_alg != "tsit5":
        raise ValueError("saveat between steps uses Tsit5's interpolant (ode_alg='tsit5')")
    if saveat is not None:
        solver = L.ODE_ATSIT5_SAVEAT if adaptive else L.ODE_TSIT5_SAVEAT
        saveat = float(saveat)
    else:
        solver = (L.ODE_RK4 if ode_alg == "rk4" else
                  L.ODE_ATSIT5 if adaptive else L.ODE_TSIT5)
    return OdeSSE(L.ODE_MODEL_BASE[model] + solver, model, 4, 6 if adaptive else 4, obs,
                  tuple(u0), float(dt), int(steps_per_obs), bool(adaptive), float(abstol),
                  float(reltol), int(maxiters), saveat, model)


def lotka_volterra_sse(obs, u0=(1.0, 1.0), dt=0.01, steps_per_obs=10, **kw) -> OdeSSE:
    return ode_sse("lotka_volterra", obs, u0, dt, steps_per_obs, **kw)


def fitzhugh_nagumo_sse(obs, u0=(1.0, 1.0), dt=0.01, steps_per_obs=10, **kw) -> OdeSSE:
    return ode_sse("fitzhugh_nagumo", obs, u0, dt, steps_per_obs, **kw)


# ---------------------------------------------------------------------------
# problem
MAXCONS = 8


class QuadraticConstraints:
    """Inequality constraints g_k(x) <= 0 for ``OptimizationFunction(f; cons)``,
    in the separable quadratic family

        g_k(x) = sum_j linear[k, j] x_j + sum_j quadratic[k, j] x_j^2 + const[k]

    (k < 8).  The reference's test problem (test/constraints.jl:11-14),
    ``[-x1 + 2 x2 - 1, x1 - 2 x2 + 1, x1^2/4 + x2^2 - 1]``, is
    ``QuadraticConstraints(linear=[[-1, 2], [1, -2], [0, 0]],
    quadratic=[[0, 0], [0, 0], [0.25, 1]], const=[-1, 1, -1])``.  The swarm
    minimises f + the reference's penalty (src/kernels.jl:28-36,
    src/utils.jl:320-372); the solution's objective is f alone, as the
    reference's (src/solve.jl:29-31)."""

    def __init__(self, *, linear=None, quadratic=None, const):
        c = np.asarray(const, np.float32).reshape(-1)
        m = c.size
        if m < 1 or m > MAXCONS:
            raise ValueError(f"1..{MAXCONS} constraints, got {m}")
        self.const = c
        self.linear = None if linear is None else np.asarray(linear, np.float32).reshape(m, -1)
        self.quadratic = (None if quadratic is None
                          else np.asarray(quadratic, np.float32).reshape(m, -1))

    def rows(self, d: int) -> np.ndarray:
        m = self.const.size
        out = np.zeros((m, 2 * d + 1), np.float32)
        for name, block, col in (("linear", self.linear, 0), ("quadratic", self.quadratic, d)):
            if block is not None:
                if block.shape[1] != d:
                    raise ValueError(f"{name} coefficients must have {d} columns")
                out[:, col:col + d] = block
        out[:, 2 * d] = self.const
        return out

    def __call__(self, x) -> np.ndarray:
        """g(x) in float64 (for checks and reporting)."""
        x = np.asarray(x, np.float64).reshape(-1)
        r = self.rows(x.size).astype(np.float64)
        d = x.size
        return r[:, :d] @ x + r[:, d:2 * d] @ (x * x) + r[:, 2 * d]


class GPUUniformSampler:
    """The default initialiser: x = r * (ub - lb) + lb (src/utils.jl:57-84)."""

    code = L.SAMPLER_UNIFORM


class GPUUnboundedSampler:
    """x_j = u0_j + r |u0_j| if u0_j != 0, else r (src/utils.jl:23-55); chosen
    automatically when lb/ub are absent or infinite in every coordinate."""

    code = L.SAMPLER_UNBOUNDED


class LatinHypercubeSample:
    """Latin hypercube design drawn on the device (one particle per stratum
    of each coordinate, keyed Feistel permutation + jitter, re-drawn by
    reinit!).  For a host-side design (QuasiMonteCarlo's samplers in the
    reference, src/init.jl:15-36) pass any object with a
    ``sample(n, lb, ub) -> (d, n) array`` method as the sampler instead."""

    code = L.SAMPLER_LATIN_HYPERCUBE


class OptimizationProblem:
    """``OptimizationProblem(f, u0, p; lb, ub)``.  Without lb/ub (or with
    infinite bounds) the swarm starts from the unbounded sampler around u0, as
    the reference's init does.  ``cons`` takes QuadraticConstraints (the
    reference's ``OptimizationFunction(f; cons)``); ``lcons``/``ucons`` are
    accepted and, as in the reference's penalty, not used."""

    def __init__(self, f, u0, p=None, *, lb=None, ub=None, shift=None, cons=None, lcons=None,
                 ucons=None, _validate=True):
        if not isinstance(f, Objective):
            raise ValueError(
                "objective must be one of the ppswarm objective markers (Rosenbrock(), "
                f"Rastrigin(), ...); got {type(f).__name__}")
        self.f = f
        self.u0 = np.asarray(u0, np.float32).reshape(-1)
        self.d = int(self.u0.size)
        if f.fixed_dim is not None and self.d != f.fixed_dim:
            raise ValueError(f"{f.name} is {f.fixed_dim}-dimensional, u0 has {self.d}")
        if isinstance(f, OdeSSE):
            self.p = f.params()
        elif p is None:
            self.p = np.zeros(0, np.float32)
        else:
            self.p = np.asarray(p, np.float32).reshape(-1)
        if _validate and self.p.size < f.n_params:
            raise ValueError(f"{f.name} needs {f.n_params} parameters, got {self.p.size}")
        if (lb is None) != (ub is None):
            raise ValueError("give both lb and ub, or neither (unbounded)")
        self.lb = None if lb is None else np.asarray(lb, np.float32).reshape(-1)
        self.ub = None if ub is None else np.asarray(ub, np.float32).reshape(-1)
        if _validate and lb is not None and (self.lb.size != self.d or self.ub.size != self.d):
            raise ValueError("lb/ub must have the dimension of u0")
        self.shift = None if shift is None else np.asarray(shift, np.float32).reshape(-1)
        if cons is not None and not isinstance(cons, QuadraticConstraints):
            raise ValueError("cons must be a QuadraticConstraints marker")
        self.cons = cons
        self.lcons, self.ucons = lcons, ucons

    def c(self):
        """pps_problem view; the returned tuple keeps the arrays alive."""
        keep = []

        def ptr(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, np.float32)
            keep.append(a)
            return a.ctypes.data_as(L.fp)

        pr = L.pps_problem()
        pr.objective = self.f.code
        pr.d = self.d
        pr.lb = ptr(self.lb)
        pr.ub = ptr(self.ub)
        pr.params = ptr(self.p) if self.p.size else None
        pr.n_params = int(self.p.size)
        pr.shift = ptr(self.shift)
        pr.u0 = ptr(self.u0)
        if isinstance(self.f, OdeSSE):
            pr.obs = ptr(self.f.obs)
            pr.n_obs = int(np.asarray(self.f.obs).size // 2)
        if self.cons is not None:
            pr.cons = ptr(self.cons.rows(self.d))
            pr.n_cons = int(self.cons.const.size)
        return pr, keep


def eval_f64(prob: OptimizationProblem, xs) -> np.ndarray:
    """FP64 host evaluation of the objective at the rows of xs (n x d)."""
    xs = np.ascontiguousarray(xs, np.float64).reshape(-1, prob.d)
    out = np.zeros(len(xs), np.float64)
    pr, keep = prob.c()
    L.check(L.lib().pps_eval_f64(C.byref(pr), len(xs), xs.ctypes.data_as(L.dp),
                                 out.ctypes.data_as(L.dp)))
    return out


# ---------------------------------------------------------------------------
# algorithms (src/algorithms.jl)
@dataclass
class ParallelPSOKernel:
    num_particles: int
    global_update: bool = True
    backend: MI355XBackend = field(default_factory=MI355XBackend)


@dataclass
class ParallelSyncPSOKernel:
    num_particles: int
    backend: MI355XBackend = field(default_factory=MI355XBackend)


@dataclass
class StrongWolfeLineSearch:
    """Strong-Wolfe line search on the projected path (Nocedal & Wright
    Alg. 3.5/3.6): the fields the reference forwards from LineSearch.jl's
    StrongWolfeLineSearch to the local search (src/hybrid.jl:131-136)."""

    c1: float = 1e-4
    c2: float = 0.9
    alpha_init: float = 1.0
    alpha_max: float = 65536.0
    maxiters: int = 20
    zoom_maxiters: int = 30


@dataclass
class ArmijoBacktracking:
    """Armijo backtracking: the first a = rho^k (k <= max_backtracks) with
    f(P(x + a p)) <= f + c g'(P(x + a p) - x).  Not a reference line search;
    the HybridPSO benchmark configuration names it."""

    c: float = 1e-4
    rho: float = 0.5
    max_backtracks: int = 20


@dataclass
class SimpleLBFGS:
    """Projected L-BFGS local search: history ``threshold`` (the reference's
    history size, SimpleOptimization.__get_threshold) and ``linesearch``,
    StrongWolfeLineSearch() by default as in the reference, or
    ArmijoBacktracking()."""

    threshold: int = 10
    linesearch: Any = field(default_factory=StrongWolfeLineSearch)


def lbfgs_params(maxiters: int, history: int, abstol: float, linesearch,
                 reltol: float = 0.0) -> "L.pps_lbfgs_params":
    """pps_lbfgs_params for a line-search marker; anything else is refused
    (the reference throws ArgumentError for a line search other than
    StrongWolfeLineSearch, src/hybrid.jl:129-130)."""
    lp = L.pps_lbfgs_params()
    L.lib().pps_lbfgs_params_default(C.byref(lp))
    lp.maxiters, lp.history, lp.abstol = int(maxiters), int(history), float(abstol)
    lp.reltol = float(reltol)
    if isinstance(linesearch, StrongWolfeLineSearch):
        lp.linesearch = L.LINESEARCH_STRONG_WOLFE
        lp.armijo_c, lp.wolfe_c2 = float(linesearch.c1), float(linesearch.c2)
        lp.alpha_init, lp.alpha_max = float(linesearch.alpha_init), float(linesearch.alpha_max)
        lp.ls_maxiters, lp.zoom_maxiters = int(linesearch.maxiters), int(linesearch.zoom_maxiters)
    elif isinstance(linesearch, ArmijoBacktracking):
        lp.linesearch = L.LINESEARCH_ARMIJO
        lp.armijo_c, lp.armijo_rho = float(linesearch.c), float(linesearch.rho)
        lp.max_backtracks = int(linesearch.max_backtracks)
    else:
        raise ValueError("the MI355X local search takes linesearch=StrongWolfeLineSearch(...) or "
                         f"ArmijoBacktracking(...), got {type(linesearch).__name__}")
    return lp


@dataclass
class HybridPSO:
    backend: MI355XBackend = field(default_factory=MI355XBackend)
    pso: Any = None
    local_opt: Any = field(default_factory=SimpleLBFGS)

    def __post_init__(self):
        if self.pso is None:
            self.pso = ParallelPSOKernel(100, global_update=False, backend=self.backend)


# ---------------------------------------------------------------------------
@dataclass
class GBest:
    position: np.ndarray
    cost: float
    index: int = -1


@dataclass
class Stats:
    time: float
    device_ms: float = 0.0


@dataclass
class OptimizationSolution:
    u: np.ndarray
    objective: float          # FP64 host evaluation of f(u)
    objective_f32: float      # FP32 cost the device computed for u
    original: Optional[np.ndarray]
    stats: Stats
    retcode: str = "Default"
    index: int = -1


class Swarm:
    """Owning wrapper of a pps_swarm handle."""

    def __init__(self, ctx: Context, prob: OptimizationProblem, params: L.pps_swarm_params,
                 positions=None):
        """positions: optional (n, d) initial positions (the QMC path)."""
        self.ctx, self.prob, self.params = ctx, prob, params
        self.d, self.n = prob.d, int(params.n)
        pr, keep = prob.c()
        h = C.c_void_p()
        if positions is None:
            L.check(L.lib().pps_swarm_init(ctx.h, C.byref(pr), C.byref(params), C.byref(h)))
        else:
            pos = self._positions(positions)
            L.check(L.lib().pps_swarm_init_positions(ctx.h, C.byref(pr), C.byref(params),
                                                     pos.ctypes.data_as(L.fp), C.byref(h)))
        self.h = h

    def _positions(self, positions) -> np.ndarray:
        pos = np.ascontiguousarray(positions, np.float32)
        if pos.shape != (self.n, self.d):
            raise ValueError(f"positions must be ({self.n}, {self.d}), got {pos.shape}")
        return pos

    def __del__(self):
        self.free()

    def free(self):
        if getattr(self, "h", None):
            L.lib().pps_swarm_free(self.h)
            self.h = None

    def run(self, maxiters: int, w: float = 0.7298, wdamp: float = 1.0) -> None:
        L.check(L.lib().pps_swarm_run(self.h, int(maxiters), float(np.float32(w)),
                                      float(np.float32(wdamp))))

    def reinit(self, positions=None):
        if positions is None:
            L.check(L.lib().pps_swarm_reinit(self.h))
        else:
            pos = self._positions(positions)
            L.check(L.lib().pps_swarm_reinit_positions(self.h, pos.ctypes.data_as(L.fp)))

    def gbest(self) -> GBest:
        pos = np.zeros(self.d, np.float32)
        cost = C.c_float()
        idx = C.c_int64()
        L.check(L.lib().pps_swarm_get_gbest(self.h, pos.ctypes.data_as(L.fp), C.byref(cost),
                                            C.byref(idx)))
        return GBest(pos, float(cost.value), int(idx.value))

    def set_gbest(self, pos, cost) -> None:
        pos = np.ascontiguousarray(pos, np.float32)
        L.check(L.lib().pps_swarm_set_gbest(self.h, pos.ctypes.data_as(L.fp), float(cost)))

    def positions(self) -> np.ndarray:
        out = np.zeros((self.n, self.d), np.float32)
        L.check(L.lib().pps_swarm_get_positions(self.h, out.ctypes.data_as(L.fp)))
        return out

    def state(self):
        x = np.zeros((self.d, self.n), np.float32)
        v = np.zeros((self.d, self.n), np.float32)
        b = np.zeros((self.d, self.n), np.float32)
        bc = np.zeros(self.n, np.float32)
        it = C.c_uint32()
        L.check(L.lib().pps_swarm_get_state(self.h, x.ctypes.data_as(L.fp), v.ctypes.data_as(L.fp),
                                            b.ctypes.data_as(L.fp), bc.ctypes.data_as(L.fp),
                                            C.byref(it)))
        return x, v, b, bc, int(it.value)

    def set_state(self, x, v, best, best_cost, it: int):
        a = [np.ascontiguousarray(q, np.float32) for q in (x, v, best, best_cost)]
        L.check(L.lib().pps_swarm_set_state(self.h, *[q.ctypes.data_as(L.fp) for q in a],
                                            int(it)))

    def lbfgs(self, maxiters=100, history=10, abstol=0.0, armijo_c=1e-4, armijo_rho=0.5,
              max_backtracks=20, want_all=True, linesearch=None, reltol=0.0):
        """Projected L-BFGS from every particle's best position, with
        ``linesearch`` (a StrongWolfeLineSearch or ArmijoBacktracking marker;
        None: Armijo backtracking with armijo_c, armijo_rho, max_backtracks).
        reltol > 0 also stops a particle after a step with |f - f_new| <=
        reltol |f|.  Returns (x n x d or None, f n or None, best_x, best_f,
        best_idx)."""
        if linesearch is None:
            linesearch = ArmijoBacktracking(armijo_c, armijo_rho, max_backtracks)
        lp = lbfgs_params(maxiters, history, abstol, linesearch, reltol)
        ox = np.zeros((self.n, self.d), np.float32) if want_all else None
        of = np.zeros(self.n, np.float32) if want_all else None
        bx = np.zeros(self.d, np.float32)
        bf = C.c_float()
        bi = C.c_int64()
        L.check(L.lib().pps_swarm_lbfgs(
            self.h, C.byref(lp), ox.ctypes.data_as(L.fp) if want_all else None,
            of.ctypes.data_as(L.fp) if want_all else None, bx.ctypes.data_as(L.fp), C.byref(bf),
            C.byref(bi)))
        return ox, of, bx, float(bf.value), int(bi.value)

    def elapsed_ms(self) -> float:
        ms = C.c_float()
        L.check(L.lib().pps_swarm_elapsed_ms(self.h, C.byref(ms)))
        return float(ms.value)


def swarm_params(n: int, mode: int = L.MODE_SYNC, *, seed: int = 0,
                 gbest_from: int = L.GBEST_PBEST, clamp: int = L.CLAMP_COMPONENT,
                 launch: int = L.LAUNCH_AUTO, qlock_min_improvers: int = 2,
                 c1: float = 1.4962, c2: float = 1.4962,
                 grid_per_cu: int = 0, sampler: int = L.SAMPLER_UNIFORM,
                 particles_per_workgroup: int = 0) -> L.pps_swarm_params:
    p = L.pps_swarm_params()
    L.lib().pps_swarm_params_default(C.byref(p))
    p.n, p.mode, p.seed = int(n), int(mode), int(seed) & 0xFFFFFFFFFFFFFFFF
    p.grid_per_cu = int(grid_per_cu)
    p.sampler = int(sampler)
    p.particles_per_workgroup = int(particles_per_workgroup)
    p.gbest_from, p.clamp, p.launch = int(gbest_from), int(clamp), int(launch)
    p.qlock_min_improvers = int(qlock_min_improvers)
    p.c1, p.c2 = float(c1), float(c2)
    return p


def autotune_sync(prob: OptimizationProblem, n: int, iters: int = 20,
                  backend: MI355XBackend = MI355XBackend()):
    """Time the sync step's workgroup sizes (particles per workgroup) and grid
    sizes (tiles per thread) for this (objective, d, n) and remember the
    fastest pair for later swarms; returns (grid_per_cu,
    particles_per_workgroup, us_per_iteration)."""
    pr, keep = prob.c()
    g = C.c_int32()
    b = C.c_int32()
    us = C.c_float()
    L.check(L.lib().pps_autotune_sync(context(backend).h, C.byref(pr), int(n), int(iters),
                                      C.byref(g), C.byref(b), C.byref(us)))
    return int(g.value), int(b.value), float(us.value)


def tuned_launch(objective: Objective, d: int, n: int):
    """The cached (grid_per_cu, particles_per_workgroup) for (objective, d,
    n), or None."""
    g = C.c_int32()
    b = C.c_int32()
    rc = L.lib().pps_tuned_launch(int(objective.code), int(d), int(n), C.byref(g), C.byref(b))
    if rc == L.E_NOTFOUND:
        return None
    L.check(rc)
    return int(g.value), int(b.value)


# ---------------------------------------------------------------------------
# init / solve! / reinit! (src/init.jl, src/solve.jl)
@dataclass
class PSOCache:
    prob: OptimizationProblem
    alg: Any
    swarm: Swarm
    gbest: GBest
    seed: int
    sampler: Any = None


def _mode_of(alg) -> int:
    if isinstance(alg, ParallelSyncPSOKernel):
        return L.MODE_SYNC
    if isinstance(alg, ParallelPSOKernel):
        return L.MODE_QLOCK if alg.global_update else L.MODE_ASYNC
    raise ValueError(f"unsupported algorithm {type(alg).__name__}")


def _host_sample(sampler, prob: OptimizationProblem, n: int) -> np.ndarray:
    """QuasiMonteCarlo.sample(n, lb, ub, sampler) -> d x n; returned as n x d."""
    x = np.asarray(sampler.sample(n, prob.lb, prob.ub), np.float32)
    if x.shape != (prob.d, n):
        raise ValueError(f"sampler.sample must return a ({prob.d}, {n}) array, got {x.shape}")
    return np.ascontiguousarray(x.T)


class HybridPSOCache:
    """``init(prob, HybridPSO)`` (src/init.jl:143-159): the PSO stage's cache
    and the algorithm; other attributes read through to the PSO cache, as the
    reference's getproperty does (src/init.jl:170-176)."""

    def __init__(self, pso_cache: PSOCache, alg: "HybridPSO"):
        self.pso_cache = pso_cache
        self.alg = alg

    def __getattr__(self, name):
        if name in ("pso_cache", "alg"):
            raise AttributeError(name)
        return getattr(self.pso_cache, name)


def init(prob: OptimizationProblem, alg, *, seed: Optional[int] = None, sampler=None,
         **kw):
    """``init(prob, alg; sampler)`` (src/init.jl:56-115): sampler is
    GPUUniformSampler() (default), GPUUnboundedSampler(), LatinHypercubeSample()
    or a host sampler with ``sample(n, lb, ub)``.  For HybridPSO: a
    HybridPSOCache around the PSO stage's cache."""
    if isinstance(alg, HybridPSO):
        if not isinstance(alg.local_opt, SimpleLBFGS):
            raise ValueError("HybridPSO on the MI355X backend supports local_opt=SimpleLBFGS() "
                             f"(got {type(alg.local_opt).__name__})")
        return HybridPSOCache(init(prob, alg.pso, seed=seed, sampler=sampler, **kw), alg)
    seed = _next_seed() if seed is None else int(seed)
    ctx = context(alg.backend)
    code = getattr(sampler, "code", L.SAMPLER_UNIFORM)
    params = swarm_params(alg.num_particles, _mode_of(alg), seed=seed, sampler=code, **kw)
    host = sampler is not None and not hasattr(sampler, "code") and prob.lb is not None
    pos = _host_sample(sampler, prob, alg.num_particles) if host else None
    sw = Swarm(ctx, prob, params, positions=pos)
    return PSOCache(prob, alg, sw, sw.gbest(), seed, sampler if host else None)


def reinit_(cache) -> None:
    """``reinit!``: a fresh initial swarm from the cache's sampler (a
    HybridPSOCache reinitialises its PSO stage, src/init.jl:160-168)."""
    if isinstance(cache, HybridPSOCache):
        reinit_(cache.pso_cache)
        return
    if cache.sampler is not None:
        cache.swarm.reinit(_host_sample(cache.sampler, cache.prob, cache.swarm.n))
    else:
        cache.swarm.reinit()
    cache.gbest = cache.swarm.gbest()


def solve_(cache, maxiters: int = 100, *, w: float = 0.7298, wdamp: float = 1.0,
           original: bool = True, **kw) -> OptimizationSolution:
    """``solve!``.  For ParallelSyncPSOKernel the attractor of the first
    iteration is the cache's initial global best on every call, as in the
    reference (its solve! does not store the result back, src/solve.jl:36-59).
    A HybridPSOCache runs the PSO stage, then the local phase
    (local_maxiters, abstol, reltol, linesearch; src/hybrid.jl:100-158)."""
    if isinstance(cache, HybridPSOCache):
        from .hybrid import solve_hybrid_cache

        return solve_hybrid_cache(cache, maxiters, w=w, wdamp=wdamp, **kw)
    if kw:
        raise TypeError(f"unexpected keyword arguments {sorted(kw)}")
    sw = cache.swarm
    if isinstance(cache.alg, ParallelSyncPSOKernel):
        sw.set_gbest(cache.gbest.position, cache.gbest.cost)
    t0 = time.perf_counter()
    sw.run(maxiters, w, wdamp)
    t1 = time.perf_counter()
    g = sw.gbest()
    if not isinstance(cache.alg, ParallelSyncPSOKernel):
        cache.gbest = g
    obj = float(eval_f64(cache.prob, g.position[None, :])[0])
    return OptimizationSolution(g.position, obj, g.cost, sw.positions() if original else None,
                                Stats(t1 - t0, sw.elapsed_ms()), "Default", g.index)


def solve(prob: OptimizationProblem, alg, maxiters: int = 100, *, seed: Optional[int] = None,
          w: float = 0.7298, wdamp: float = 1.0, **kw) -> OptimizationSolution:
    if isinstance(alg, HybridPSO):
        return solve_(init(prob, alg, seed=seed), maxiters, w=w, wdamp=wdamp, **kw)
    cache = init(prob, alg, seed=seed)
    return solve_(cache, maxiters, w=w, wdamp=wdamp)
