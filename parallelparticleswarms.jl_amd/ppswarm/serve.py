"""Served / batched solving of many independent problems (libpps server).

Each request is an independent sync-PSO solve (ParallelSyncPSOKernel semantics)
of a Rosenbrock or Rastrigin problem at d in {2, 4, 8, 16, 32}.  The server
admits requests as descriptor slots and HBM free up, packs every admitted group
into one launch (one workgroup per problem, the iteration loop inside it) and
retires each problem as soon as its workgroup finishes.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _lib as L
from . import api


@dataclass
class Result:
    u: np.ndarray
    cost_f32: float
    cost_f64: float
    index: int
    t_submit: float
    t_admit: float
    t_done: float

    @property
    def latency(self) -> float:
        return self.t_done - self.t_submit


def _result(r: L.pps_result) -> Result:
    return Result(np.array(r.gbest_pos[: r.d], np.float32), float(r.gbest_cost),
                  float(r.gbest_cost_f64), int(r.gbest_idx), r.t_submit_s, r.t_admit_s,
                  r.t_done_s)


@dataclass
class Request:
    prob: api.OptimizationProblem
    n: int
    maxiters: int
    seed: int
    w: float = 0.7298
    wdamp: float = 1.0


class Server:
    def __init__(self, ctx: Optional[api.Context] = None, hbm_budget: int = 0,
                 max_active: int = 0, slice_us: int = 0):
        self.ctx = ctx or api.context()
        cfg = L.pps_server_config(int(hbm_budget), int(max_active), int(slice_us))
        h = C.c_void_p()
        L.check(L.lib().pps_server_create(self.ctx.h, C.byref(cfg), C.byref(h)))
        self.h = h

    def submit(self, req: Request) -> int:
        pr, keep = req.prob.c()
        sp = api.swarm_params(req.n, L.MODE_SYNC, seed=req.seed)
        t = C.c_uint64()
        L.check(L.lib().pps_server_submit(self.h, C.byref(pr), C.byref(sp), int(req.maxiters),
                                          float(req.w), float(req.wdamp), C.byref(t)))
        return int(t.value)

    def poll(self, ticket: int) -> Optional[Result]:
        r = L.pps_result()
        rc = L.lib().pps_server_poll(self.h, ticket, C.byref(r))
        if rc == L.E_PENDING:
            return None
        L.check(rc)
        return _result(r)

    def wait(self, ticket: int, timeout: float = 0.0) -> Result:
        r = L.pps_result()
        L.check(L.lib().pps_server_wait(self.h, ticket, float(timeout), C.byref(r)))
        return _result(r)

    def cancel(self, ticket: int) -> None:
        L.check(L.lib().pps_server_cancel(self.h, ticket))

    def stats(self) -> dict:
        v = [C.c_int64() for _ in range(4)]
        L.check(L.lib().pps_server_stats(self.h, *[C.byref(x) for x in v]))
        return dict(zip(("submitted", "completed", "launches", "peak_bytes"), [x.value for x in v]))

    def close(self) -> None:
        if self.h:
            L.lib().pps_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_solve(reqs: list[Request], ctx: Optional[api.Context] = None):
    """Solve all requests through one temporary server; returns (results,
    elapsed_ms)."""
    ctx = ctx or api.context()
    n = len(reqs)
    probs = (L.pps_problem * n)()
    params = (L.pps_swarm_params * n)()
    keep = []
    for i, q in enumerate(reqs):
        pr, k = q.prob.c()
        keep.append(k)
        probs[i] = pr
        params[i] = api.swarm_params(q.n, L.MODE_SYNC, seed=q.seed)
    iters = (C.c_int32 * n)(*[int(q.maxiters) for q in reqs])
    ws = np.array([q.w for q in reqs], np.float32)
    wd = np.array([q.wdamp for q in reqs], np.float32)
    out = (L.pps_result * n)()
    ms = C.c_float()
    L.check(L.lib().pps_batch_solve(ctx.h, n, probs, params, iters, ws.ctypes.data_as(L.fp),
                                    wd.ctypes.data_as(L.fp), out, C.byref(ms)))
    return [_result(out[i]) for i in range(n)], float(ms.value)


def solve_alone(req: Request, launch: int = L.LAUNCH_AUTO, ctx=None) -> Result:
    """The plain per-problem path: init, `maxiters` iterations with the
    library's launch choice (one workgroup up to 1024 particles, else one
    launch per iteration for a one-shot run), read the best."""
    ctx = ctx or api.context()
    sw = api.Swarm(ctx, req.prob, api.swarm_params(req.n, L.MODE_SYNC, seed=req.seed,
                                                   launch=launch))
    sw.run(req.maxiters, req.w, req.wdamp)
    g = sw.gbest()
    sw.free()
    f64 = float(api.eval_f64(req.prob, g.position[None, :])[0])
    return Result(g.position, g.cost, f64, g.index, 0.0, 0.0, 0.0)
