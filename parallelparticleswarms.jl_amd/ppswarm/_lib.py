"""ctypes binding of libpps (include/pps.h).

The shared library is built in-tree into ``parallelparticleswarms.jl_amd/lib``.
Loading fails loudly when it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import re

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.environ.get("PPS_LIBPPS") or os.path.join(_PKG, "lib", "libpps.so")
HEADER = os.path.join(os.path.dirname(_PKG), "include", "pps.h")

# status codes
OK, E_INVALID, E_HIP, E_NOMEM, E_NOTFOUND, E_PENDING, E_CANCELLED, E_TIMEOUT, E_STATE = (
    0, -1, -2, -3, -4, -5, -6, -7, -8)

# objectives
OBJ_ROSENBROCK, OBJ_RASTRIGIN, OBJ_SPHERE, OBJ_BEALE, OBJ_BOOTH, OBJ_HIMMELBLAU = 1, 2, 3, 4, 5, 6
OBJ_QUADRATIC, OBJ_BOUNDARY_QUADRATIC, OBJ_ROSEN2, OBJ_LOTKA_VOLTERRA = 7, 8, 9, 16
OBJ_LOTKA_VOLTERRA_TSIT5, OBJ_LOTKA_VOLTERRA_ATSIT5 = 17, 18
OBJ_LOTKA_VOLTERRA_TSIT5_SAVEAT, OBJ_LOTKA_VOLTERRA_ATSIT5_SAVEAT = 19, 20
# FitzHugh-Nagumo: the same five solvers at 21..25
OBJ_FITZHUGH_NAGUMO = 21
ODE_MODEL_BASE = {"lotka_volterra": OBJ_LOTKA_VOLTERRA, "fitzhugh_nagumo": OBJ_FITZHUGH_NAGUMO}
ODE_RK4, ODE_TSIT5, ODE_ATSIT5, ODE_TSIT5_SAVEAT, ODE_ATSIT5_SAVEAT = 0, 1, 2, 3, 4

MODE_SYNC, MODE_ASYNC, MODE_QLOCK = 0, 1, 2
GBEST_PBEST, GBEST_CURRENT = 0, 1
CLAMP_COMPONENT, CLAMP_LEXICOGRAPHIC = 0, 1
SAMPLER_UNIFORM, SAMPLER_UNBOUNDED, SAMPLER_LATIN_HYPERCUBE = 0, 1, 2
LAUNCH_AUTO, LAUNCH_EAGER, LAUNCH_GRAPH, LAUNCH_WORKGROUP = 0, 1, 2, 3

fp = C.POINTER(C.c_float)
dp = C.POINTER(C.c_double)
i64p = C.POINTER(C.c_int64)
u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)


class PPSError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"libpps error {code}: {msg}")
        self.code = code


ABI_VERSION = 3  # include/pps.h PPS_ABI_VERSION mirrored by the structures below


class pps_problem(C.Structure):
    _fields_ = [
        ("objective", C.c_int32), ("d", C.c_int32), ("lb", fp), ("ub", fp), ("params", fp),
        ("n_params", C.c_int32), ("shift", fp), ("obs", fp), ("n_obs", C.c_int32),
        ("cons", fp), ("n_cons", C.c_int32), ("u0", fp),
    ]


class pps_swarm_params(C.Structure):
    _fields_ = [
        ("n", C.c_int64), ("mode", C.c_int32), ("gbest_from", C.c_int32), ("clamp", C.c_int32),
        ("launch", C.c_int32), ("seed", C.c_uint64), ("c1", C.c_float), ("c2", C.c_float),
        ("qlock_min_improvers", C.c_int32), ("grid_per_cu", C.c_int32),
        ("sampler", C.c_int32), ("particles_per_workgroup", C.c_int32),
    ]


LINESEARCH_ARMIJO, LINESEARCH_STRONG_WOLFE = 0, 1


class pps_lbfgs_params(C.Structure):
    _fields_ = [("maxiters", C.c_int32), ("history", C.c_int32), ("abstol", C.c_float),
                ("armijo_c", C.c_float), ("armijo_rho", C.c_float),
                ("max_backtracks", C.c_int32), ("linesearch", C.c_int32),
                ("wolfe_c2", C.c_float), ("alpha_init", C.c_float), ("alpha_max", C.c_float),
                ("ls_maxiters", C.c_int32), ("zoom_maxiters", C.c_int32),
                ("reltol", C.c_float)]


class pps_server_config(C.Structure):
    _fields_ = [("hbm_budget", C.c_int64), ("max_active", C.c_int32),
                ("slice_us", C.c_int32), ("reserved", C.c_int32)]


class pps_result(C.Structure):
    _fields_ = [("gbest_cost", C.c_float), ("gbest_cost_f64", C.c_double),
                ("gbest_idx", C.c_int64), ("d", C.c_int32), ("gbest_pos", C.c_float * 64),
                ("t_submit_s", C.c_double), ("t_admit_s", C.c_double), ("t_done_s", C.c_double)]


_lib = None


def _sig(L, name, args, res=C.c_int):
    f = getattr(L, name)
    f.argtypes = args
    f.restype = res


def lib():
    """Load libpps.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"libpps.so not found at {LIB_PATH}; build it with "
            "`make -C parallelparticleswarms.jl_amd/csrc` (or __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    _sig(L, "pps_last_error", [], C.c_char_p)
    _sig(L, "pps_abi_version", [])
    _sig(L, "pps_swarm_params_default", [C.POINTER(pps_swarm_params)], None)
    _sig(L, "pps_ctx_create", [C.c_int, C.POINTER(vp)])
    _sig(L, "pps_ctx_destroy", [vp])
    _sig(L, "pps_ctx_info", [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32), i64p, i64p])
    _sig(L, "pps_swarm_init", [vp, C.POINTER(pps_problem), C.POINTER(pps_swarm_params),
                               C.POINTER(vp)])
    _sig(L, "pps_swarm_reinit", [vp])
    _sig(L, "pps_swarm_init_positions", [vp, C.POINTER(pps_problem), C.POINTER(pps_swarm_params),
                                         fp, C.POINTER(vp)])
    _sig(L, "pps_swarm_reinit_positions", [vp, fp])
    _sig(L, "pps_swarm_run", [vp, C.c_int32, C.c_float, C.c_float])
    _sig(L, "pps_swarm_get_gbest", [vp, fp, fp, i64p])
    _sig(L, "pps_swarm_set_gbest", [vp, fp, C.c_float])
    _sig(L, "pps_swarm_get_positions", [vp, fp])
    _sig(L, "pps_swarm_get_state", [vp, fp, fp, fp, fp, u32p])
    _sig(L, "pps_swarm_set_state", [vp, fp, fp, fp, fp, C.c_uint32])
    _sig(L, "pps_swarm_elapsed_ms", [vp, fp])
    _sig(L, "pps_swarm_free", [vp])
    _sig(L, "pps_eval_f64", [C.POINTER(pps_problem), C.c_int64, dp, dp])
    _sig(L, "pps_lbfgs_params_default", [C.POINTER(pps_lbfgs_params)], None)
    _sig(L, "pps_autotune_sync", [vp, C.POINTER(pps_problem), C.c_int64, C.c_int32,
                                  C.POINTER(C.c_int32), C.POINTER(C.c_int32), fp])
    _sig(L, "pps_tuned_launch", [C.c_int32, C.c_int32, C.c_int64, C.POINTER(C.c_int32),
                                 C.POINTER(C.c_int32)])
    _sig(L, "pps_swarm_lbfgs", [vp, C.POINTER(pps_lbfgs_params), fp, fp, fp, fp, i64p])
    _sig(L, "pps_swarm_lbfgs_noalloc", [vp, C.POINTER(pps_lbfgs_params), fp, fp, fp, fp, i64p])
    _sig(L, "pps_server_create", [vp, C.POINTER(pps_server_config), C.POINTER(vp)])
    _sig(L, "pps_server_submit", [vp, C.POINTER(pps_problem), C.POINTER(pps_swarm_params),
                                  C.c_int32, C.c_float, C.c_float, u64p])
    _sig(L, "pps_server_poll", [vp, C.c_uint64, C.POINTER(pps_result)])
    _sig(L, "pps_server_wait", [vp, C.c_uint64, C.c_double, C.POINTER(pps_result)])
    _sig(L, "pps_server_cancel", [vp, C.c_uint64])
    _sig(L, "pps_server_stats", [vp, i64p, i64p, i64p, i64p])
    _sig(L, "pps_server_destroy", [vp])
    _sig(L, "pps_batch_solve", [vp, C.c_int32, C.POINTER(pps_problem),
                                C.POINTER(pps_swarm_params), C.POINTER(C.c_int32), fp, fp,
                                C.POINTER(pps_result), fp])
    if L.pps_abi_version() != ABI_VERSION:
        raise RuntimeError(f"{LIB_PATH} has ABI version {L.pps_abi_version()}, this binding "
                           f"mirrors version {ABI_VERSION}; rebuild the library")
    _lib = L
    return L


def check(rc: int) -> None:
    if rc != OK:
        msg = lib().pps_last_error()
        raise PPSError(rc, msg.decode() if msg else "")


def declared_functions(header: str = HEADER) -> list[str]:
    """Names of the functions include/pps.h declares."""
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pps_[a-z0-9_]+)\s*\(", text)))
