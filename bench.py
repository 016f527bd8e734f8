#!/usr/bin/env python3
"""Benchmark of the MI355X particle-swarm backend (libpps).

Headline workload (one step): the plain ParallelSyncPSOKernel path at the
single-GPU benchmark shape — 10-D Rastrigin and 10-D Rosenbrock (p = [1, 100]),
bounds [-5.12, 5.12], 2^20 particles each, 1000 iterations each, FP32 SoA state
resident in HBM, the launch shape chosen by the library's autotuner (untimed,
before the warm-up) and the iterations launched as the library chooses at this
size (one launch each; replayed chunk graphs below 2^17 particles).  `value`
is particle-iterations per second over the timed steps (init excluded, as in
the reference's benchmarks/CPU_vs_GPU/benchmark.jl:107-121).

Also reported on the same line:
  roofline      HBM roofline of the dominant kernel (the fused sync step):
                algorithmic bytes per launch / average launch time (hipEvents
                on the swarm's stream over the timed region).  At 2^20 x 10-D
                the 130 MB state sits inside the 256 MB Infinity Cache.
  roofline_hbm  the same kernel with the state past the cache (2^22
                particles x 10-D, 520 MB), both objectives.
  async         ParallelPSOKernel(global_update=false) at the same shape.
  qlock         ParallelPSOKernel(global_update=true), the reference's default:
                one launch per iteration against the live global best, at the
                same shape.
  c0            3-D Rosenbrock, 1000 particles, 500 iterations (one
                resident workgroup): time per solve and per iteration.
  batched       the 4096-problem served mix solved in packed launches
                (all submitted at once): solves/s and particle-iters/s, next
                to the plain per-problem launch loop (init + one launch per
                iteration, problem after problem) timed on the first 256
                problems of the same mix, whose results must equal the packed
                results bit for bit.
  served        the same mix arriving as a Poisson stream at a fixed rate
                below saturation: solves/s and p50/p99 request latency.
  hybrid        HybridPSO: sync PSO, 16-D Rastrigin, 2^16 particles, 200
                iterations, then projected L-BFGS (m = 10, 100 iterations,
                Armijo) from every particle; final minimum re-evaluated in FP64.
  ode           Lotka-Volterra parameter estimation: 2^14 particles, 300
                iterations, in-kernel FP32 RK4 (dt = 0.01, 101 observations);
                the same with Tsit5 at fixed and adaptive steps.  The
                adaptive leg is timed as the library runs it (a particle's
                loss stops once it reaches the particle's best cost) and
                again with every loss summed in full; both must give the
                same solution.
  cpu_baseline  the CPU oracle (reference-faithful AoS restatement, OpenMP
                over particles + serial argmin per iteration) on a bounded
                sample of the same workload on this host; `c0` adds the
                reference's CPU-backend case (3-D Rosenbrock, 1000 particles,
                500 iterations) per solve, all threads and one thread.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "parallelparticleswarms.jl_amd"))

import numpy as np  # noqa: E402

METRIC = "particle-iters/sec plain vs batched; served solves/sec + p99 latency"
HBM_PEAK = 8.0e12
N = 1 << 20
D = 10
T = 1000
BYTES_PER_PARTICLE_ITER = 4 * (5 * D + 1)  # read x,v,best (3d) + best_cost, write x,v (2d)


# Set by --eager: every multi-workgroup sync swarm of every leg launches its
# iterations one by one (the same kernels; used for profiler runs).
EAGER = False


def _launch(L, default):
    return L.LAUNCH_EAGER if EAGER else default


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_swarm(P, obj, n, d, seed, mode, launch=None, shape=(0, 0)):
    lb = np.full(d, -5.12, np.float32)
    ub = np.full(d, 5.12, np.float32)
    f = P.Rastrigin() if obj == "rastrigin" else P.Rosenbrock()
    params = None if obj == "rastrigin" else [1.0, 100.0]
    prob = P.OptimizationProblem(f, np.zeros(d, np.float32), params, lb=lb, ub=ub)
    from ppswarm import _lib as L

    sp = P.swarm_params(n, mode, seed=seed,
                        launch=L.LAUNCH_AUTO if launch is None else launch,
                        particles_per_workgroup=shape[0], grid_per_cu=shape[1])
    return P.Swarm(P.context(), prob, sp)


def traffic_from_profiles():
    """HBM bytes per sync-step launch measured with rocprofv3 --pmc, if the
    summary is committed (profiles/sync_step_c1_traffic.md)."""
    path = os.path.join(ROOT, "profiles", "sync_step_c1_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            try:
                return float(line.split(":", 1)[1].strip())
            except ValueError:
                return None
    return None


def cpu_baseline(sample_iters=200):
    import pps_oracle as O

    O.lib()
    threads = O.lib().ora_num_threads()
    total_pi, total_s = 0, 0.0
    for obj, params in ((O.RASTRIGIN, ()), (O.ROSENBROCK, (1.0, 100.0))):
        prob = O.Problem(obj, D, np.full(D, -5.12, np.float32), np.full(D, 5.12, np.float32),
                         np.asarray(params, np.float32))
        s = O.Swarm(prob, N, seed=7, nthreads=threads)
        s.run_sync(1)  # warm-up
        t0 = time.perf_counter()
        s.run_sync(sample_iters)
        total_s += time.perf_counter() - t0
        total_pi += N * sample_iters
    # the reference's own CPU-backend case: 3-D Rosenbrock, lb = -1, ub = 10, 1000
    # particles, 500 iterations, the global best taken from the argmin's current
    # position as the CPU backend does (src/lowerlevel_solve.jl:59-60)
    c0 = {}
    prob = O.Problem(O.ROSENBROCK, 3, np.full(3, -1.0, np.float32), np.full(3, 10.0, np.float32),
                     np.asarray((1.0, 100.0), np.float32))
    for key, nt in (("solve_ms", threads), ("solve_ms_1_thread", 1)):
        times = []
        for rep in range(6):
            s = O.Swarm(prob, 1000, seed=rep, nthreads=nt)
            t0 = time.perf_counter()
            s.run_sync(500, gbest_from=O.GBEST_CURRENT)
            times.append(time.perf_counter() - t0)
        c0[key] = float(np.median(times[1:]) * 1e3)
    # the reference's CPU_vs_GPU benchmark on the CPU backend (the restatement's
    # sync run, global best from the argmin's current position), up to 2^15
    # particles: a timed run after an untimed one on the same swarm
    ref = []
    prob = O.Problem(O.ROSENBROCK, 10, np.full(10, -1.0, np.float32),
                     np.full(10, 10.0, np.float32), np.asarray((1.0, 100.0), np.float32))
    for lg in REF_SWEEP_LOG2N:
        if lg > 15:
            break
        s = O.Swarm(prob, 1 << lg, seed=lg, nthreads=threads)
        s.run_sync(REF_SWEEP_ITERS, gbest_from=O.GBEST_CURRENT)
        t0 = time.perf_counter()
        s.run_sync(REF_SWEEP_ITERS, gbest_from=O.GBEST_CURRENT)
        ref.append({"particles": 1 << lg, "wall_ms": (time.perf_counter() - t0) * 1e3})
    return {"value": total_pi / total_s, "unit": "particle-iters/s", "cores": threads,
            "kind": "port",
            "reference_sweep": ref,
            "sample": f"sync PSO, 10-D Rastrigin + 10-D Rosenbrock, 2^20 particles each, "
                      f"{sample_iters} timed iterations each (after 1 warm-up), "
                      f"{total_s:.1f} s of CPU time",
            "c0": dict(c0, workload="3-D Rosenbrock, 1000 particles, 500 iterations, "
                                    "median of 5 solves (after 1)")}


def c4_mix(P, L, ctx, count=4096, plain_sample=256, served_qps=None):
    reqs, _ = P.served_mix(count, seed=2024)
    pi = sum(q.n * q.maxiters for q in reqs)
    P.batch_solve(reqs[:64])  # warm-up (module load, arena)
    t0 = time.perf_counter()
    res, ms = P.batch_solve(reqs)
    wall = time.perf_counter() - t0
    # plain per-problem launch loop on a sample of the same mix
    t0 = time.perf_counter()
    plain = [P.solve_alone(q, _launch(L, L.LAUNCH_AUTO)) for q in reqs[:plain_sample]]
    pwall = time.perf_counter() - t0
    mismatch = sum(1 for a, b in zip(res, plain)
                   if not (np.array_equal(a.u, b.u) and a.cost_f32 == b.cost_f32))
    max_df64 = max(abs(a.cost_f64 - b.cost_f64) for a, b in zip(res, plain))
    plain_rate = plain_sample / pwall
    batched = {"solves_per_s": count / wall, "particle_iters_per_s": pi / wall,
               "wall_s": wall, "problems": count,
               "plain_loop_solves_per_s": plain_rate,
               "plain_loop_sample": plain_sample,
               "speedup_vs_plain_loop": (count / wall) / plain_rate,
               "bitwise_mismatches_vs_plain": mismatch,
               "max_abs_f64_diff_vs_plain": max_df64,
               "max_abs_f32_vs_f64": max(abs(r.cost_f64 - r.cost_f32) for r in res),
               "max_rel_f32_vs_f64": max(abs(r.cost_f64 - r.cost_f32) / max(1.0, abs(r.cost_f64))
                                         for r in res),
               "objective_tolerance": "|f32(u) - f64(u)| <= 1e-4 * max(1, |f64(u)|) for every "
                                      "returned u; f64 of the packed u equals f64 of the plain "
                                      "u (max_abs_f64_diff_vs_plain)"}
    # served: Poisson arrivals at half the packed throughput
    qps = served_qps or 0.5 * count / wall
    reqs2, arr = P.served_mix(count, seed=2024, qps=qps)
    with P.Server(ctx) as sv:
        t0 = time.perf_counter()
        tickets = []
        for q, a in zip(reqs2, arr):
            d = a - (time.perf_counter() - t0)
            if d > 0.0005:
                time.sleep(d - 0.0003)
            while time.perf_counter() - t0 < a:
                pass
            tickets.append(sv.submit(q))
        sres = [sv.wait(t) for t in tickets]
        swall = time.perf_counter() - t0
        st = sv.stats()
    lat = np.array([r.latency for r in sres]) * 1e3
    served = {"qps_offered": qps, "solves_per_s": count / swall,
              "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
              "mean_ms": float(lat.mean()), "launches": st["launches"],
              "peak_resident_bytes": st["peak_bytes"],
              "bitwise_equal_to_batched": all(np.array_equal(a.u, b.u) and a.cost_f32 == b.cost_f32
                                              for a, b in zip(res, sres))}
    return batched, served


def hybrid_c2(P, L, ctx):
    d, n = 16, 1 << 16
    prob = P.OptimizationProblem(P.Rastrigin(), np.zeros(d, np.float32), lb=np.full(d, -5.12),
                                 ub=np.full(d, 5.12))
    sw = P.Swarm(ctx, prob, P.swarm_params(n, L.MODE_SYNC, seed=16,
                                           launch=_launch(L, L.LAUNCH_GRAPH)))
    # Warm-up outside the timed region: the 200-iteration graph is captured and
    # launched once (a new graph's first launch uploads it), and the local
    # search runs once (a kernel's first launch loads its code object).  The
    # timed run starts from a fresh swarm (the next epoch of the stream).
    sw.run(200)
    sw.lbfgs(maxiters=100, history=10, want_all=False)
    sw.reinit()
    t0 = time.perf_counter()
    sw.run(200)
    pso_ms = sw.elapsed_ms()
    g = sw.gbest()
    _, _, bx, bf, bi = sw.lbfgs(maxiters=100, history=10, want_all=False)
    wall = time.perf_counter() - t0
    lb_ms = sw.elapsed_ms()
    best = min(float(P.eval_f64(prob, g.position[None, :])[0]),
               float(P.eval_f64(prob, bx[None, :])[0]))
    sw.free()
    return {"wall_s": wall, "pso_device_ms": pso_ms, "lbfgs_device_ms": lb_ms,
            "pso_particle_iters_per_s": n * 200 / (pso_ms / 1e3),
            "lbfgs_particle_iters_per_s": n * 100 / (lb_ms / 1e3),
            "pso_best_f32": g.cost, "lbfgs_best_f32": bf, "final_objective_f64": best}


def _full_loss(run):
    """Run `run()` once more with every ODE loss summed in full
    (PPS_ODE_FULL_LOSS=1): returns (device ms, solution position)."""
    prev = os.environ.get("PPS_ODE_FULL_LOSS")
    os.environ["PPS_ODE_FULL_LOSS"] = "1"
    try:
        sol, sw = run()
        ms = sw.elapsed_ms()
        sw.free()
    finally:
        if prev is None:
            del os.environ["PPS_ODE_FULL_LOSS"]
        else:
            os.environ["PPS_ODE_FULL_LOSS"] = prev
    return ms, np.asarray(sol.u)


def _ode_leg(run, particles, iters, extra=None, compare_full=False):
    """One ODE estimation leg.  compare_full (adaptive solvers): also time it
    with every loss summed in full instead of stopped at each particle's best
    cost (the library default); the two solutions must be the same bit for
    bit."""
    sol, sw = run()
    ms = sw.elapsed_ms()
    sw.free()
    out = {"device_ms": ms, "particle_iters_per_s": particles * iters / (ms / 1e3),
           "theta": [float(v) for v in sol.u], "loss_f64": sol.objective}
    if compare_full:
        full_ms, full_u = _full_loss(run)
        out["full_loss_device_ms"] = full_ms
        out["same_solution_as_full_loss"] = bool(np.array_equal(np.asarray(sol.u), full_u))
    if extra:
        out.update(extra(sol))
    return out


def ode_c3(P, L, clock_khz):
    import json as _json
    lk = {"launch": _launch(L, L.LAUNCH_AUTO)}
    doc = _json.load(open(os.path.join(ROOT, "tests", "golden", "lv_obs.json")))
    obs = np.array(doc["obs"], np.float32)
    P.parameter_estim_ode(obs, np.zeros(4), np.full(4, 5.0), 1 << 14, 5, seed=1, **lk)  # warm-up
    t0 = time.perf_counter()
    sol, sw = P.parameter_estim_ode(obs, np.zeros(4), np.full(4, 5.0), 1 << 14, 300, seed=3, **lk)
    wall = time.perf_counter() - t0
    ms = sw.elapsed_ms()
    sw.free()
    rk4_steps = 1000  # per objective evaluation (100 intervals x 10 steps)
    out = {"wall_s": wall, "device_ms": ms,
           "particle_iters_per_s": (1 << 14) * 300 / (ms / 1e3),
           "rk4_steps_per_s": (1 << 14) * 300 * rk4_steps / (ms / 1e3),
           "theta": [float(v) for v in sol.u], "loss_f64": sol.objective}
    # Latency floor: a particle's loss is one chain of 1000 dependent RK4
    # steps, so an iteration cannot end before one workgroup's chain does.
    # The same kernel and iteration graph on 64 particles (one workgroup of
    # two waves, nothing else on the device) measures that floor, launch and
    # global-best hand-over included; the full swarm's time per iteration is
    # compared with it.
    _, sw1 = P.parameter_estim_ode(obs, np.zeros(4), np.full(4, 5.0), 64, 300, seed=3,
                                   launch=_launch(L, L.LAUNCH_GRAPH))
    sw1.free()
    _, sw1 = P.parameter_estim_ode(obs, np.zeros(4), np.full(4, 5.0), 64, 300, seed=3,
                                   launch=_launch(L, L.LAUNCH_GRAPH))
    floor_us = sw1.elapsed_ms() * 1e3 / 300
    sw1.free()
    it_us = ms * 1e3 / 300
    out["latency_bound"] = {
        "bound": "dependent chain", "floor_us_per_iter": floor_us, "us_per_iter": it_us,
        "frac": floor_us / it_us,
        "cycles_per_rk4_step_at_floor": floor_us * clock_khz / 1e3 / rk4_steps,
        "workload": "the same kernel on 64 particles (one workgroup), 300 iterations"}
    # the reference's default solver, GPUTsit5: fixed steps (dt = 0.01) and
    # adaptive steps (abstol 1e-6, reltol 1e-3, initial dt 0.01), same shape
    for key, adaptive in (("tsit5", False), ("tsit5_adaptive", True)):
        out[key] = _ode_leg(
            lambda: P.parameter_estim_ode(obs, np.zeros(4), np.full(4, 5.0), 1 << 14, 300, seed=3, **lk,
                                          ode_alg="tsit5", adaptive=adaptive), 1 << 14, 300,
            compare_full=adaptive)
    # the reference's own Lotka-Volterra benchmark shape (10000 particles, 100
    # iterations, rates in [0, 10], GPUTsit5 dt = 0.1 with saveat every 0.01
    # on [0, 3], noise-free data)
    doc = _json.load(open(os.path.join(ROOT, "tests", "golden", "lv_obs_benchmark.json")))
    obs_b = np.array(doc["obs"], np.float32)
    theta_true = np.asarray(doc["theta"])
    P.parameter_estim_ode(obs_b, np.zeros(4), np.full(4, 10.0), 10000, 5, seed=1, **lk, ode_alg="tsit5",
                          dt=0.1, saveat=0.01)  # warm-up
    out["reference_benchmark_shape"] = _ode_leg(
        lambda: P.parameter_estim_ode(obs_b, np.zeros(4), np.full(4, 10.0), 10000, 100, seed=1, **lk,
                                      ode_alg="tsit5", dt=0.1, saveat=0.01), 10000, 100,
        lambda sol: {"max_rel_err_theta": float(np.max(np.abs(np.asarray(sol.u) / theta_true - 1))),
                     "workload": "10000 particles x 100 iterations, Tsit5 dt 0.1, "
                                 "301 samples every 0.01"})
    # the reference's FitzHugh-Nagumo benchmark shape (same solver settings,
    # rates in [0, 5])
    doc = _json.load(open(os.path.join(ROOT, "tests", "golden", "fhn_obs_benchmark.json")))
    obs_f = np.array(doc["obs"], np.float32)
    kw = dict(ode_alg="tsit5", dt=0.1, saveat=0.01, model="fitzhugh_nagumo")
    P.parameter_estim_ode(obs_f, np.zeros(4), np.full(4, 5.0), 10000, 5, seed=0, **lk, **kw)  # warm-up
    out["fitzhugh_nagumo_reference_benchmark_shape"] = _ode_leg(
        lambda: P.parameter_estim_ode(obs_f, np.zeros(4), np.full(4, 5.0), 10000, 100, seed=0, **lk,
                                      **kw), 10000, 100,
        lambda sol: {"workload": "10000 particles x 100 iterations, Tsit5 dt 0.1, "
                                 "301 samples every 0.01"})
    return out


def n_sweep(P, L):
    """Sync step at 2^10..2^24 particles (10-D, both objectives): average launch
    time over a timed run (launched as the library's automatic mode chooses:
    replayed chunk graphs below 2^17 particles, one by one from there) after
    1000 warm-up iterations (the
    headline's timed range also starts after 1000 iterations).  The
    state is 124 B per particle, so it outgrows the 256 MB Infinity Cache
    from 2^22 particles on; below that the achieved rate includes cache hits."""
    rows = []
    for lg in range(10, 25, 2):
        n = 1 << lg
        iters = int(max(30, min(1000, 4e9 / (n * BYTES_PER_PARTICLE_ITER))))
        us = []
        for obj, seed in (("rastrigin", 11), ("rosenbrock", 12)):
            s = make_swarm(P, obj, n, D, seed, L.MODE_SYNC, L.LAUNCH_AUTO)
            s.run(1000)  # past the early phase (and the chunk graph's capture)
            s.run(iters)
            s.run(iters)
            us.append(s.elapsed_ms() * 1e3 / iters)
            s.free()
        avg = sum(us) / len(us)
        rate = n * BYTES_PER_PARTICLE_ITER / (avg * 1e-6)
        rows.append({"log2_n": lg, "iterations": iters, "us_per_launch": [round(u, 2) for u in us],
                     "state_mb": round(n * 4 * (3 * D + 1) / 1e6, 1),
                     "achieved_gbps": round(rate / 1e9, 1), "frac": round(rate / HBM_PEAK, 4),
                     "particle_iters_per_s": n / (avg * 1e-6)})
        log(f"[bench] sweep 2^{lg}: {avg:.2f} us/launch, {rate / 1e12:.2f} TB/s")
    return rows


# The reference's own benchmark (benchmarks/CPU_vs_GPU/benchmark.jl:10-23,
# 86-197, 200-222): 10-D Rosenbrock p = [1, 100], bounds [-1, 10], N = 2^3,
# 2^5, ..., 2^19 particles, 500 iterations; each algorithm runs once untimed and
# then once timed on the same swarm (vectorized_solve! twice on the same
# particles); HybridPSO = async PSO (global_update = false) + SimpleLBFGS with
# local_maxiters = 30 through solve, timed by its second solve's stats.time.
REF_SWEEP_LOG2N = tuple(range(3, 20, 2))
REF_SWEEP_ITERS = 500


def reference_problem_rosenbrock10(P):
    return P.OptimizationProblem(P.Rosenbrock(), np.zeros(10, np.float32), [1.0, 100.0],
                                 lb=np.full(10, -1.0), ub=np.full(10, 10.0))


def reference_sweep(P, L, ctx):
    prob = reference_problem_rosenbrock10(P)
    rows = []
    for lg in REF_SWEEP_LOG2N:
        n = 1 << lg
        row = {"particles": n}
        for name, mode in (("sync", L.MODE_SYNC), ("async", L.MODE_ASYNC),
                           ("queue_lock", L.MODE_QLOCK)):
            sw = P.Swarm(ctx, prob, P.swarm_params(n, mode, seed=lg))
            sw.run(REF_SWEEP_ITERS)
            t0 = time.perf_counter()
            sw.run(REF_SWEEP_ITERS)
            wall = time.perf_counter() - t0
            row[name] = {"wall_ms": wall * 1e3, "device_ms": sw.elapsed_ms(),
                         "best_f32": sw.gbest().cost}
            sw.free()
        alg = P.HybridPSO(pso=P.ParallelPSOKernel(n, global_update=False),
                          local_opt=P.SimpleLBFGS())
        P.solve(prob, alg, REF_SWEEP_ITERS, seed=lg, local_maxiters=30)
        sol = P.solve(prob, alg, REF_SWEEP_ITERS, seed=lg, local_maxiters=30)
        row["hybrid"] = {"wall_ms": sol.stats.time * 1e3, "objective": sol.objective}
        rows.append(row)
        log(f"[bench] reference sweep 2^{lg}: sync {row['sync']['wall_ms']:.3f} ms")
    return {"workload": "the reference's CPU_vs_GPU benchmark: 10-D Rosenbrock, bounds [-1, 10], "
                        "500 iterations, a timed run after an untimed one on the same swarm; "
                        "HybridPSO = async PSO + L-BFGS (30 iterations, Strong-Wolfe)",
            "rows": rows}


def beyond_cache(P, L, lg=22, iters=200):
    """The sync step with the state past the 256 MB Infinity Cache (2^22
    particles x 10-D: 520 MB of state), both objectives: average launch time
    over `iters` iterations after 1000 untimed ones (launched one by one, the
    library's choice at this size)."""
    n = 1 << lg
    us = []
    for obj, seed in (("rastrigin", 21), ("rosenbrock", 22)):
        s = make_swarm(P, obj, n, D, seed, L.MODE_SYNC)
        s.run(1000)
        s.run(iters)
        us.append(s.elapsed_ms() * 1e3 / iters)
        s.free()
    avg = sum(us) / len(us)
    rate = n * BYTES_PER_PARTICLE_ITER / (avg * 1e-6)
    return {"particles": n, "dim": D, "state_mb": round(n * 4 * (3 * D + 1) / 1e6, 1),
            "iterations_timed": iters, "us_per_launch": [round(u, 2) for u in us],
            "bytes_per_launch": n * BYTES_PER_PARTICLE_ITER, "achieved": rate / 1e9,
            "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": rate / HBM_PEAK,
            "particle_iters_per_s": n / (avg * 1e-6)}


# Problems with n * d at or above this run as several workgroups per
# iteration (PART items) in the packed kernel, smaller ones in one workgroup
# each (SMALL items); the library's split.
LARGE_WORK = 8192


def profile_mix(P, kind):
    """Only the packed solve of the 4096-problem served mix, or of its SMALL
    or PART problems, with no warm-up (for counter runs, whose per-dispatch
    figures are summed over every k_slice dispatch).  Algorithmic bytes: each
    problem's init writes x, v, best (3d) and best cost, each iteration reads
    x, v, best (3d) and best cost and writes x, v (2d)."""
    reqs, _ = P.served_mix(4096, seed=2024)
    if kind == "small":
        reqs = [q for q in reqs if q.n * q.prob.d < LARGE_WORK]
    elif kind == "large":
        reqs = [q for q in reqs if q.n * q.prob.d >= LARGE_WORK]
    t0 = time.perf_counter()
    res, ms = P.batch_solve(reqs)
    wall = time.perf_counter() - t0
    ab = sum(q.n * 4 * ((3 * q.prob.d + 1) + q.maxiters * (5 * q.prob.d + 1)) for q in reqs)
    return {"profile_mix": kind, "problems": len(reqs), "wall_s": wall,
            "particle_iters": sum(q.n * q.maxiters for q in reqs), "algorithmic_bytes": ab,
            "algorithmic_gbps": ab / wall / 1e9}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the autotuner before the headline (default launch shape)")
    ap.add_argument("--launch-shape", default=None,
                    help="PARTICLES_PER_WORKGROUP/GRID_PER_CU of the headline swarms instead of "
                         "the autotuner's choice (e.g. 512/2)")
    ap.add_argument("--eager", action="store_true",
                    help="force one launch per iteration for every multi-workgroup sync swarm "
                         "of every leg (the library's own choice for the headline's 2^20 "
                         "particles; the same kernels everywhere)")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--extras", default="hbm,async,qlock,c0,refsweep,c4,hybrid,ode",
                    help="comma list of the legs to run besides the headline")
    ap.add_argument("--profile-mix", choices=("all", "small", "large"),
                    help="run only the packed solve of the served mix (all, its SMALL or its "
                         "PART problems) and print its algorithmic bytes; for counter runs")
    ap.add_argument("--sweep", action="store_true",
                    help="also time the sync step at 2^10..2^24 particles (adds ~1 min)")
    args = ap.parse_args()
    global EAGER
    EAGER = args.eager
    legs = set() if args.no_extras else {x.strip() for x in args.extras.split(",") if x.strip()}
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}): "
              "the swarm path has no multi-GPU launch", file=sys.stderr)
        sys.exit(2)

    import ppswarm as P
    from ppswarm import _lib as L

    ctx = P.context()
    info = ctx.info()
    log(f"[bench] device: {info}")
    if args.profile_mix:
        print(json.dumps(profile_mix(P, args.profile_mix)), flush=True)
        return
    # The library's autotuner picks particles per workgroup and workgroups per
    # CU for each headline problem (cached for the swarms created after it);
    # untimed, before the warm-up.
    tuned = {}
    shape = (0, 0)
    if args.launch_shape:
        ppw, g = (int(v) for v in args.launch_shape.split("/"))
        shape = (ppw, g)
        tuned = {"forced": {"particles_per_workgroup": ppw, "grid_per_cu": g}}
    elif not args.no_tune:
        for obj in ("rastrigin", "rosenbrock"):
            f = P.Rastrigin() if obj == "rastrigin" else P.Rosenbrock()
            prob = P.OptimizationProblem(f, np.zeros(D, np.float32),
                                         None if obj == "rastrigin" else [1.0, 100.0],
                                         lb=np.full(D, -5.12), ub=np.full(D, 5.12))
            g, b, us = P.autotune_sync(prob, N, iters=200)
            tuned[obj] = {"grid_per_cu": g, "particles_per_workgroup": b, "us_per_iter": us}
        log(f"[bench] autotuned: {tuned}")
    hl = L.LAUNCH_EAGER if args.eager else L.LAUNCH_AUTO
    swarms = [make_swarm(P, "rastrigin", N, D, 101, L.MODE_SYNC, hl, shape),
              make_swarm(P, "rosenbrock", N, D, 202, L.MODE_SYNC, hl, shape)]

    per_swarm_ms = [0.0, 0.0]

    def step(timed=False):
        ms = 0.0
        for k, s in enumerate(swarms):
            s.run(T)
            ms += s.elapsed_ms()
            if timed:
                per_swarm_ms[k] += s.elapsed_ms()
        return ms

    for i in range(args.warmup):
        step()
        log(f"[bench] warmup {i + 1}/{args.warmup} done")
    dev_ms = 0.0
    t0 = time.perf_counter()
    for i in range(args.steps):
        dev_ms += step(timed=True)
        log(f"[bench] step {i + 1}/{args.steps} done")
    wall = time.perf_counter() - t0
    pi_per_step = 2 * N * T
    value = pi_per_step * args.steps / wall
    launches = 2 * T * args.steps
    avg_launch_s = dev_ms / 1e3 / launches
    bytes_per_launch = N * BYTES_PER_PARTICLE_ITER
    achieved = bytes_per_launch / avg_launch_s
    out = {
        "metric": METRIC, "value": value, "unit": "particle-iters/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "plain ParallelSyncPSOKernel, 10-D Rastrigin + 10-D Rosenbrock "
                               "(p=[1,100]), bounds [-5.12,5.12], 2^20 particles each, "
                               "1000 iterations each per step, FP32 SoA, "
                               + "one launch per iteration (the library's choice from 2^17 "
                                 "particles at d > 4; replayed chunk graphs below)",
                   "particles": N, "dim": D, "iterations": T, "problems_per_step": 2,
                   "launch": tuned or "default (256 particles per workgroup, 2 per CU)"},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK,
                     "traffic": traffic_from_profiles(), "kernel": "k_sync_step",
                     "bytes_per_launch": bytes_per_launch,
                     "avg_launch_us": avg_launch_s * 1e6,
                     "avg_launch_us_by_problem": {
                         name: per_swarm_ms[k] * 1e3 / (T * args.steps)
                         for k, name in enumerate(("rastrigin", "rosenbrock"))}},
    }
    for s in swarms:
        s.free()
    if "hbm" in legs:
        out["roofline_hbm"] = beyond_cache(P, L)
        log(f"[bench] beyond-cache leg done: {out['roofline_hbm']['frac']:.3f} of HBM peak")

    if "async" in legs:
        # async (ParallelPSOKernel(global_update=false)) at the same shape
        a = [make_swarm(P, "rastrigin", N, D, 303, L.MODE_ASYNC),
             make_swarm(P, "rosenbrock", N, D, 404, L.MODE_ASYNC)]
        for s in a:
            s.run(50)
        t0 = time.perf_counter()
        ams = 0.0
        for s in a:
            s.run(T)
            ams += s.elapsed_ms()
        aw = time.perf_counter() - t0
        out["async"] = {"value": 2 * N * T / aw, "unit": "particle-iters/s",
                        "device_ms": ams, "workload": "same shape, one launch per problem"}
        for s in a:
            s.free()
        log("[bench] async done")
    if "qlock" in legs:
        # queue-lock (ParallelPSOKernel(global_update=true)) at the same shape
        q = [make_swarm(P, "rastrigin", N, D, 505, L.MODE_QLOCK, L.LAUNCH_AUTO),
             make_swarm(P, "rosenbrock", N, D, 606, L.MODE_QLOCK, L.LAUNCH_AUTO)]
        for s in q:
            s.run(50)
        t0 = time.perf_counter()
        qms = 0.0
        for s in q:
            s.run(T)
            qms += s.elapsed_ms()
        qw = time.perf_counter() - t0
        out["qlock"] = {"value": 2 * N * T / qw, "unit": "particle-iters/s", "device_ms": qms,
                        "us_per_launch": qms * 1e3 / (2 * T),
                        "workload": "same shape, one launch per iteration"}
        for s in q:
            s.free()
        log("[bench] qlock done")
    if "c0" in legs:
        # C0: 3-D Rosenbrock, lb=-1, ub=10, 1000 particles, 500 iterations
        prob = P.OptimizationProblem(P.Rosenbrock(), np.zeros(3, np.float32), [1.0, 100.0],
                                     lb=np.full(3, -1.0), ub=np.full(3, 10.0))
        times, dms = [], []
        for rep in range(6):
            s = P.Swarm(ctx, prob, P.swarm_params(1000, L.MODE_SYNC, seed=rep))
            t0 = time.perf_counter()
            s.run(500)
            times.append(time.perf_counter() - t0)
            dms.append(s.elapsed_ms())
            s.free()
        out["c0"] = {"solve_ms_median": float(np.median(times[1:]) * 1e3),
                     "device_us_per_iter": float(np.median(dms[1:]) * 1e3 / 500),
                     "particle_iters_per_s": 1000 * 500 / float(np.median(times[1:]))}
        log("[bench] c0 done")

    if args.sweep:
        out["sweep"] = n_sweep(P, L)
        log("[bench] sweep done")
    if "refsweep" in legs:
        out["reference_sweep"] = reference_sweep(P, L, ctx)
        log("[bench] reference sweep done")

    if "c4" in legs:
        out["batched"], out["served"] = c4_mix(P, L, ctx)
        log("[bench] c4 done")
    if "hybrid" in legs:
        out["hybrid"] = hybrid_c2(P, L, ctx)
        log("[bench] hybrid done")
    if "ode" in legs:
        out["ode"] = ode_c3(P, L, info["clock_khz"])
        log("[bench] ode done")

    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        out["cpu_baseline"] = cpu_baseline()
        log("[bench] cpu baseline done")
        if "reference_sweep" in out:
            # CPU over GPU sync time per size, as the reference's benchmark
            # reports mean(cpu_times ./ gpu_sync_times)
            gpu = {r["particles"]: r["sync"]["wall_ms"] for r in out["reference_sweep"]["rows"]}
            ratios = [c["wall_ms"] / gpu[c["particles"]]
                      for c in out["cpu_baseline"]["reference_sweep"] if c["particles"] in gpu]
            out["reference_sweep"]["cpu_over_gpu_sync"] = ratios
            out["reference_sweep"]["mean_cpu_over_gpu_sync"] = (float(np.mean(ratios))
                                                                if ratios else None)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
