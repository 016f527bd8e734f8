/*
 * pps.h — C ABI of libpps, the MI355X (gfx950) particle-swarm backend.
 *
 * Drop-in boundary for the particle-swarm hot path of ParallelParticleSwarms.jl
 * (v1.4.0).  A Julia package extension `ccall`s these entry points in place of
 * the KernelAbstractions kernels and host drivers the reference launches; the
 * Python package `ppswarm` binds the same symbols with ctypes.
 *
 * Conventions
 *  - Every function returns int: PPS_OK (0) or a negative PPS_E_* code.  The
 *    message of the last failure on the calling thread is pps_last_error().
 *    No C++ exception crosses this boundary.
 *  - Handles (pps_ctx, pps_swarm, pps_server) are library-owned; free them with
 *    the matching _free / _destroy call.  Host buffers are caller-owned and are
 *    only read or written for the duration of the call.
 *  - A pps_ctx (one HIP device + one stream) is single-thread-affine.  A
 *    pps_server is thread-safe for submit/poll/wait.
 *  - Device arithmetic is FP32.  pps_eval_f64 evaluates objectives in FP64 on
 *    the host.
 *  - Positions exchanged as "AoS" are particle-major, N x d floats, the layout
 *    of a Vector{SVector{d,Float32}} (reference src/ParallelParticleSwarms.jl:34-40).
 *    State exchanged as "SoA" is dimension-major, d x N floats (row j holds
 *    coordinate j of every particle), the layout the kernels keep in HBM.
 */
#ifndef PPS_H
#define PPS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PPS_ABI_VERSION 3
#define PPS_MAXCONS 8  /* constraints per problem */

/* ---- status codes ------------------------------------------------------ */
enum {
  PPS_OK = 0,
  PPS_E_INVALID = -1,     /* bad argument / unsupported combination         */
  PPS_E_HIP = -2,         /* HIP runtime error                              */
  PPS_E_NOMEM = -3,       /* device allocation failed or admission refused  */
  PPS_E_NOTFOUND = -4,    /* unknown ticket                                 */
  PPS_E_PENDING = -5,     /* request not finished yet (poll)                */
  PPS_E_CANCELLED = -6,   /* request cancelled                              */
  PPS_E_TIMEOUT = -7,     /* wait timed out                                 */
  PPS_E_STATE = -8        /* wrong lifecycle state                          */
};

/* ---- objectives (closed set; the reference accepts any Julia closure,
 *      src/kernels.jl:31-33, which cannot cross a C ABI) ------------------ */
enum {
  /* Σ_{i<d} p2*(x_{i+1}-x_i^2)^2 + (p1-x_i)^2, p=[p1,p2]
   * (reference test/gpu.jl:14-20, README.md:20-22). */
  PPS_OBJ_ROSENBROCK = 1,
  /* 10 d + Σ (x_i^2 - 10 cos(2π x_i)). */
  PPS_OBJ_RASTRIGIN = 2,
  /* Σ x_i^2. */
  PPS_OBJ_SPHERE = 3,
  /* Two-dimensional known-answer functions of reference test/lbfgs.jl:29-35. */
  PPS_OBJ_BEALE = 4,
  PPS_OBJ_BOOTH = 5,
  PPS_OBJ_HIMMELBLAU = 6,
  PPS_OBJ_QUADRATIC = 7,          /* p1 (x1-1)^2 + (x2+2)^2              */
  PPS_OBJ_BOUNDARY_QUADRATIC = 8, /* (x1+1)^2                            */
  PPS_OBJ_ROSEN2 = 9,             /* (p1-x1)^2 + p2 (x2-x1^2)^2          */
  /* Lotka-Volterra parameter estimation: sum of squared errors of an FP32
   * trajectory against observations (reference src/ode_pso.jl:47-52), solved
   * in the kernel by one of three methods. */
  PPS_OBJ_LOTKA_VOLTERRA = 16,       /* fixed-step RK4                         */
  PPS_OBJ_LOTKA_VOLTERRA_TSIT5 = 17, /* fixed-step Tsit5 (GPUTsit5, Val(false)) */
  PPS_OBJ_LOTKA_VOLTERRA_ATSIT5 = 18, /* adaptive Tsit5 (GPUTsit5, Val(true))   */
  /* Tsit5 with saveat semantics: steps are not tied to the observation times;
   * each observation inside a step is read from Tsit5's continuous extension
   * (how GPUTsit5 fills `saveat`, e.g. dt = 0.1 with saveat every 0.01 in
   * benchmarks/Lotka_Volterra/lotka_volterra.jl:124-127). */
  PPS_OBJ_LOTKA_VOLTERRA_TSIT5_SAVEAT = 19,  /* fixed dt                     */
  PPS_OBJ_LOTKA_VOLTERRA_ATSIT5_SAVEAT = 20, /* adaptive                     */
  /* FitzHugh-Nagumo parameter estimation, rates (a, b, tau_inv, l):
   * dv = v - v^3/3 - w + l, dw = tau_inv (v + a - b w) (reference
   * benchmarks/Fitzhugh_Nagumo/ode_estimation.jl:6-12), with the same five
   * solvers, parameters and observation layout as Lotka-Volterra. */
  PPS_OBJ_FITZHUGH_NAGUMO = 21,
  PPS_OBJ_FITZHUGH_NAGUMO_TSIT5 = 22,
  PPS_OBJ_FITZHUGH_NAGUMO_ATSIT5 = 23,
  PPS_OBJ_FITZHUGH_NAGUMO_TSIT5_SAVEAT = 24,
  PPS_OBJ_FITZHUGH_NAGUMO_ATSIT5_SAVEAT = 25
};

/* Swarm update semantics. */
enum {
  PPS_MODE_SYNC = 0,  /* ParallelSyncPSOKernel: Jacobi step, argmin after each
                         iteration (src/kernels.jl:127-194,
                         src/lowerlevel_solve.jl:1-65)                        */
  PPS_MODE_ASYNC = 1, /* ParallelPSOKernel(global_update=false): one launch,
                         per-particle running best (src/kernels.jl:196-225)   */
  PPS_MODE_QLOCK = 2  /* ParallelPSOKernel(global_update=true): live global
                         best, block queue + lock (src/kernels.jl:38-125)     */
};

/* Which position becomes the global best after a sync iteration. */
enum {
  PPS_GBEST_PBEST = 0,   /* best_position of the argmin particle (GPU backend,
                            src/kernels.jl:153,171)                          */
  PPS_GBEST_CURRENT = 1  /* position of the argmin particle (CPU backend,
                            src/lowerlevel_solve.jl:59-60)                   */
};

/* Bound handling after x <- x + v. */
enum {
  PPS_CLAMP_COMPONENT = 0,    /* min.(max.(x, lb), ub)  src/kernels.jl:15-17 */
  PPS_CLAMP_LEXICOGRAPHIC = 1 /* min(max(x, lb), ub) on whole vectors,
                                 src/ode_pso.jl:18-19; the ODE estimation
                                 path's, so the ODE objectives only          */
};

/* Launch strategy for PPS_MODE_SYNC. */
enum {
  PPS_LAUNCH_AUTO = 0,       /* WORKGROUP up to 1024 particles, GRAPH below
                                2^17 (below 2^19 for d <= 4; a swarm's first
                                capture waits for a run of 128 iterations,
                                EAGER before), EAGER above                 */
  PPS_LAUNCH_EAGER = 1,      /* one launch per iteration, no graph            */
  PPS_LAUNCH_GRAPH = 2,      /* chunks of 32 iterations replayed from a graph
                                captured once per swarm (weights from a
                                per-run table), the rest one by one        */
  PPS_LAUNCH_WORKGROUP = 3   /* whole loop in one workgroup (small swarms)    */
};

/* Problem: objective, bounds and parameters (OptimizationProblem fields
 * f, lb, ub, p; reference src/init.jl:88-115, src/utils.jl:310-318). */
typedef struct pps_problem {
  int32_t objective;     /* PPS_OBJ_*                                        */
  int32_t d;             /* dimension, 1..64 (ODE objectives: 4)             */
  const float* lb;       /* d floats, lb <= ub; finite, or -Inf/+Inf in every */
  const float* ub;       /* coordinate; both NULL = unbounded (the unbounded
                            sampler around u0, src/init.jl:70-74)            */
  const float* params;   /* n_params floats (Rosenbrock: [p1, p2])           */
  int32_t n_params;
  const float* shift;    /* d floats or NULL: f is evaluated at x - shift    */
  /* ODE objectives (Lotka-Volterra, FitzHugh-Nagumo) only: observations
   * u(t_k), k = 0..n_obs-1,
   * interleaved [u1(t0), u2(t0), u1(t1), ...]; the trajectory starts at
   * params[0..1] = u0.  RK4 and fixed-step Tsit5: steps dt = params[2],
   * sampled every (int)params[3] steps.  Adaptive Tsit5: t_k = k * params[2],
   * initial step params[3], abstol params[4], reltol params[5], optional step
   * limit params[6] (default 10000; beyond it the loss is +Inf; controller
   * documented at ode_sse_tsit5 in the device sources).  The _SAVEAT forms:
   * t_k = k * params[2], step (initial step when adaptive) params[3], and
   * for the adaptive one params[4..6] as above; the last step ends on t_{n_obs-1}.
   * With adaptive steps, a particle's loss inside a swarm step is summed only
   * until it reaches the particle's best cost, since the sum of squares
   * cannot fall back below it and only the comparison is used: positions,
   * velocities, best positions, best costs and the global best are the same
   * as with full sums, and a particle whose trajectory turns stiff stops long
   * before the step limit.  Setting PPS_ODE_FULL_LOSS=1 in the environment
   * before a swarm is created makes every loss a full sum (for timing
   * comparisons). */
  const float* obs;
  int32_t n_obs;
  /* Inequality constraints g_k(x) <= 0, k = 0..n_cons-1 (n_cons <= PPS_MAXCONS),
   * the reference's `OptimizationFunction(f; cons)` for the separable quadratic
   * family
   *     g_k(x) = sum_j a_kj x_j + sum_j q_kj x_j^2 + c_k
   * stored per constraint as 2d+1 floats [a_k0..a_k(d-1), q_k0..q_k(d-1), c_k].
   * Handled as the reference does (src/kernels.jl:28-36, src/utils.jl:320-372):
   * cost = f(x) + h(iter) * (sum_k theta(q_k) + sum_k theta(q_k) q_k^gamma(q_k)),
   * q_k = max(g_k(x), 0), default theta/gamma schedules, h = sqrt, and `iter` as
   * the reference passes it (init: 1; sync and queue-lock: particle index + 1;
   * async: iteration of the launch + 1).  lcons/ucons are not used by the
   * reference's penalty and have no field here.  Rosenbrock, Rastrigin and
   * Sphere objectives; NULL / 0 for none. */
  const float* cons;
  int32_t n_cons;
  /* u0 (d floats, NULL = zeros): the unbounded sampler's centre
   * (src/utils.jl:23-55). */
  const float* u0;
} pps_problem;

/* Swarm/algorithm parameters (ParallelSyncPSOKernel / ParallelPSOKernel
 * fields, src/algorithms.jl:78-121,192-204). */
typedef struct pps_swarm_params {
  int64_t n;                 /* num_particles                                */
  int32_t mode;              /* PPS_MODE_*                                   */
  int32_t gbest_from;        /* PPS_GBEST_* (sync only)                      */
  int32_t clamp;             /* PPS_CLAMP_*                                  */
  int32_t launch;            /* PPS_LAUNCH_* (sync only)                     */
  uint64_t seed;             /* Philox key of this swarm's random stream     */
  float c1, c2;              /* 1.4962f0 in the reference; 0, not finite or
                                at least 2^-102 in magnitude (others are
                                refused)                                     */
  int32_t qlock_min_improvers; /* QLOCK: block publishes only when at least
                                  this many particles beat the global best;
                                  the reference uses 2 (src/kernels.jl:97)   */
  int32_t grid_per_cu;       /* workgroups per CU of the multi-workgroup sync
                                step; 0 = autotuned value if one is cached
                                for (objective, d, N), else 2. Results do
                                not depend on it.                            */
  int32_t sampler;           /* PPS_SAMPLER_*: initial positions (init and
                                reinit); bounds that are infinite in every
                                coordinate select PPS_SAMPLER_UNBOUNDED, as
                                the reference's init does (src/init.jl:70-74) */
  int32_t particles_per_workgroup; /* particles (= threads) per workgroup of
                                the multi-workgroup sync step: 256, 512 or
                                1024 (Rosenbrock, Rastrigin and Sphere at
                                d in {1,2,3,4,8,10,16}, 256 or 512 at d = 32,
                                without constraints, state arrays below
                                4 GiB), 256 otherwise (64 for the ODE
                                objectives, their only size); 0 =
                                autotuned value if one is cached for
                                (objective, d, N), else the objective's own
                                size.  With G
                                workgroups each thread steps ceil(tiles / G)
                                tiles of that many particles per iteration.
                                Results do not depend on it.               */
} pps_swarm_params;

/* Initial-position samplers (the reference's `sampler` keyword of init,
 * src/init.jl:15-54,56-115). */
enum {
  PPS_SAMPLER_UNIFORM = 0,        /* GPUUniformSampler: r*(ub-lb)+lb (default) */
  PPS_SAMPLER_UNBOUNDED = 1,      /* GPUUnboundedSampler: |u0_j| > 0 ?
                                     u0_j + r*|u0_j| : r  (src/utils.jl:23-55)  */
  PPS_SAMPLER_LATIN_HYPERCUBE = 2 /* Latin hypercube design on the device: in
                                     each coordinate, particle i takes stratum
                                     pi_j(i) of n equal strata of [lb, ub] plus
                                     a uniform offset inside it; pi_j is a keyed
                                     permutation (4-round Feistel network with a
                                     Philox round function, cycle-walked to
                                     [0, n)), re-keyed by every reinit.  The
                                     stratification of QuasiMonteCarlo's
                                     LatinHypercubeSample; any other design can
                                     be passed as positions, see below.        */
};

/* Fill p with the reference defaults (c1 = c2 = 1.4962, SYNC, PBEST,
 * per-component clamp, qlock_min_improvers = 2, auto launch). */
void pps_swarm_params_default(pps_swarm_params* p);

typedef struct pps_ctx pps_ctx;
typedef struct pps_swarm pps_swarm;
typedef struct pps_server pps_server;

/* ---- errors / context --------------------------------------------------- */
const char* pps_last_error(void);
int pps_abi_version(void);
/* Create a context on HIP device `device` with its own stream. */
int pps_ctx_create(int device, pps_ctx** out);
int pps_ctx_destroy(pps_ctx* ctx);
/* Device facts read from hipDeviceProp: CU count, clock (kHz), memory. */
int pps_ctx_info(pps_ctx* ctx, int32_t* num_cus, int32_t* clock_khz,
                 int64_t* total_mem, int64_t* free_mem);

/* ---- swarm lifecycle (init / solve! / reinit!, src/init.jl:56-141) ----- */
/* Allocate SoA state in HBM and run the uniform initialiser
 * (src/utils.jl:57-84, src/init.jl:38-54): x_j = r*(ub_j-lb_j)+lb_j, v = 0,
 * cost = f(x), best = x; gbest = argmin (ties: lowest index). */
int pps_swarm_init(pps_ctx* ctx, const pps_problem* prob,
                   const pps_swarm_params* params, pps_swarm** out);
/* Draw a fresh initial swarm (next epoch of the random stream). */
int pps_swarm_reinit(pps_swarm* s);
/* init / reinit! with host-sampled initial positions: the reference's path
 * for QuasiMonteCarlo samplers (src/init.jl:15-36: QuasiMonteCarlo.sample on
 * the host, then gpu_init_particles! over the samples, src/utils.jl:86-112).
 * positions: n x d floats, particle-major (the memory of the d x n matrix
 * QuasiMonteCarlo.sample returns), finite; used as given (not clamped).
 * params->sampler is ignored by these two calls. */
int pps_swarm_init_positions(pps_ctx* ctx, const pps_problem* prob,
                             const pps_swarm_params* params, const float* positions,
                             pps_swarm** out);
int pps_swarm_reinit_positions(pps_swarm* s, const float* positions);
/* Run `maxiters` iterations; w_k = w * wdamp^k as an iterated FP32 product
 * (src/lowerlevel_solve.jl:23-33).  Resumable: the random stream continues. */
int pps_swarm_run(pps_swarm* s, int32_t maxiters, float w, float wdamp);
/* Global best after the last run: position (d floats), FP32 cost, particle
 * index of the argmin (PPS_MODE_QLOCK: index of the last published winner). */
int pps_swarm_get_gbest(pps_swarm* s, float* pos, float* cost, int64_t* idx);
/* Set the attractor used by the next iteration (the reference's
 * cache.gbest, src/init.jl:88-115; ParallelSyncPSOKernel does not write the
 * result back into it, src/solve.jl:36-59). */
int pps_swarm_set_gbest(pps_swarm* s, const float* pos, float cost);
/* Current positions, AoS (N x d): the solution's `original` field
 * (src/solve.jl:53). */
int pps_swarm_get_positions(pps_swarm* s, float* out_aos);
/* Full SoA state (any pointer may be NULL): x, v, best (d x N each), best_cost
 * (N), and the iteration counter of the random stream. */
int pps_swarm_get_state(pps_swarm* s, float* x, float* v, float* best,
                        float* best_cost, uint32_t* iter);
int pps_swarm_set_state(pps_swarm* s, const float* x, const float* v,
                        const float* best, const float* best_cost,
                        uint32_t iter);
/* Device time of the last run (hipEvents on the swarm's stream around its
 * launches; the capture of a new iteration graph happens before the first). */
int pps_swarm_elapsed_ms(pps_swarm* s, float* ms);
int pps_swarm_free(pps_swarm* s);

/* ---- HybridPSO local phase (src/hybrid.jl:100-158, src/lbfgs.jl:9-21) --- */
/* Line searches of the local phase, on the projected path
 * phi(a) = f(P(x + a p)). */
enum {
  PPS_LINESEARCH_ARMIJO = 0,      /* first a = rho^k with
                                     f(P(x + a p)) <= f + c g'(P(x + a p) - x)  */
  PPS_LINESEARCH_STRONG_WOLFE = 1 /* phi(a) <= phi(0) + c1 a phi'(0) and
                                     |phi'(a)| <= c2 |phi'(0)| (Nocedal & Wright
                                     Alg. 3.5/3.6; the reference's only accepted
                                     search, LineSearch.jl's StrongWolfeLineSearch,
                                     src/hybrid.jl:129-136); phi' counts the
                                     coordinates the projection lets move      */
};
typedef struct pps_lbfgs_params {
  int32_t maxiters;        /* local_maxiters                                */
  int32_t history;         /* m, 1..16                                      */
  float abstol;            /* stop when max |P(x - grad f) - x| <= abstol   */
  float armijo_c;          /* sufficient-decrease constant c1 (1e-4)        */
  float armijo_rho;        /* Armijo: backtracking factor (0.5)             */
  int32_t max_backtracks;  /* Armijo: step reductions per search (20)       */
  int32_t linesearch;      /* PPS_LINESEARCH_*                              */
  float wolfe_c2;          /* Strong-Wolfe: curvature constant c2 (0.9),
                              armijo_c < wolfe_c2 < 1                      */
  float alpha_init;        /* Strong-Wolfe: first trial step (1)            */
  float alpha_max;         /* Strong-Wolfe: largest trial step (65536); trial
                              steps double from alpha_init up to it         */
  int32_t ls_maxiters;     /* Strong-Wolfe: trial steps before a bracket (20) */
  int32_t zoom_maxiters;   /* Strong-Wolfe: zoom iterations (30); zoom's trial
                              is the minimiser of the quadratic through
                              phi(lo), phi'(lo), phi(hi) when inside the
                              middle 80 % of the bracket, else its midpoint;
                              without a step meeting both conditions the
                              last step of sufficient decrease is taken     */
  float reltol;            /* also stop after a step with
                              |f - f_new| <= reltol |f| (0: off)            */
} pps_lbfgs_params;
/* The reference's defaults: Strong-Wolfe (c1 1e-4, c2 0.9), m = 10, 100
 * iterations; the Armijo fields at c 1e-4, rho 0.5, 20 reductions. */
void pps_lbfgs_params_default(pps_lbfgs_params* p);
/* Projected L-BFGS from every particle's best position (objectives with an
 * analytic gradient: Rosenbrock, Rastrigin, Sphere and the 2-D functions).
 * out_x: N x d AoS (may be NULL), out_f: N values (non-finite -> +Inf,
 * src/lbfgs.jl:20; may be NULL), best_x/best_f/best_idx: the argmin (lowest
 * index; any may be NULL).  The swarm state is not modified. */
int pps_swarm_lbfgs(pps_swarm* s, const pps_lbfgs_params* lp, float* out_x,
                    float* out_f, float* best_x, float* best_f,
                    int64_t* best_idx);

/* Allocation-free form of pps_swarm_lbfgs (same arguments, same results bit
 * for bit).  It has no counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Its device scratch (the local
 * search's SoA results, their argmin key and the AoS staging of out_x) is a
 * file-scope device array of PPS_NOALLOC_SCRATCH_BYTES per GPU, held until
 * the process ends; the call needs about 4 * n * (2d + 1) bytes of it
 * (4 * n * (d + 1) with out_x NULL) and returns PPS_E_INVALID beyond that.
 * It calls no device or host allocator.  Every call on one GPU shares that
 * GPU's array and leaves its contents there: calls that overlap on one GPU
 * (several threads, several contexts) race. */
#define PPS_NOALLOC_SCRATCH_BYTES (64 << 20)
int pps_swarm_lbfgs_noalloc(pps_swarm* s, const pps_lbfgs_params* lp, float* out_x,
                            float* out_f, float* best_x, float* best_f,
                            int64_t* best_idx);

/* ---- autotuning ------------------------------------------------------- */
/* Time the multi-workgroup sync step on a scratch swarm of `n` particles
 * (after max(iters, 1000) untimed iterations, `iters` timed ones, best of
 * three) for every particles_per_workgroup
 * the kernels are compiled for (256, 512, 1024; 256 only for other objectives
 * and dimensions) and grid_per_cu in {1, 2, 3, 4, 6, 8} (which sets the tiles
 * each thread steps per iteration), remember the fastest pair for
 * (objective, d, ceil(log2 n)) for later pps_swarm_init calls in this
 * process, and append the line
 *     objective d ceil(log2 n) grid_per_cu particles_per_workgroup
 * to the file named by $PPS_TUNE_CACHE if set (read back at the first lookup;
 * lines without the last field mean 256).  The ODE objectives keep their
 * 64-particle workgroups; only the grid is tuned for them.  Output pointers
 * may be NULL. */
int pps_autotune_sync(pps_ctx* ctx, const pps_problem* prob, int64_t n,
                      int32_t iters, int32_t* best_grid_per_cu,
                      int32_t* best_particles_per_workgroup,
                      float* best_us_per_iter);
/* The cached choice for (objective, d, n): PPS_OK with both values, or
 * PPS_E_NOTFOUND (outputs untouched). */
int pps_tuned_launch(int32_t objective, int32_t d, int64_t n,
                     int32_t* grid_per_cu, int32_t* particles_per_workgroup);

/* ---- host FP64 re-evaluation -------------------------------------------- */
int pps_eval_f64(const pps_problem* prob, int64_t n, const double* x_aos,
                 double* f);

/* ---- served / batched solving (no reference counterpart) --------------- */
/* Independent sync-PSO problems (ParallelSyncPSOKernel semantics with the
 * best-position global best) of Rosenbrock or Rastrigin at d in 1..32.  Each problem's result equals what pps_swarm_init +
 * pps_swarm_run compute for it alone, bit for bit. */
typedef struct pps_server_config {
  int64_t hbm_budget;      /* bytes of swarm state admitted at once; 0 = 80%
                              of free HBM at create time, at most 16 GiB    */
  int32_t max_active;      /* problems resident at once (0 = 4096)         */
  int32_t slice_us;        /* device time per launch slice (0 = 1000 us);
                              while a problem with n * d >= 8192 is
                              resident (one iteration per slice) at most
                              20 us                                         */
  int32_t reserved;
} pps_server_config;

typedef struct pps_result {
  float gbest_cost;        /* FP32 cost of the final global best            */
  double gbest_cost_f64;   /* same position re-evaluated in FP64            */
  int64_t gbest_idx;
  int32_t d;
  float gbest_pos[64];
  double t_submit_s, t_admit_s, t_done_s; /* host steady-clock seconds      */
} pps_result;

/* Start a server (its own host thread and streams) on ctx's device. */
int pps_server_create(pps_ctx* ctx, const pps_server_config* cfg,
                      pps_server** out);
/* Queue one problem; arrays are copied during the call.  Problems whose state
 * exceeds the whole HBM budget are refused with PPS_E_NOMEM. */
int pps_server_submit(pps_server* sv, const pps_problem* prob,
                      const pps_swarm_params* params, int32_t maxiters,
                      float w, float wdamp, uint64_t* ticket);
/* PPS_OK with the result, PPS_E_PENDING while queued or running.  A request's
 * final status (its result, PPS_E_CANCELLED or PPS_E_HIP) is returned once, by
 * poll or wait; the ticket is then released and unknown (PPS_E_NOTFOUND). */
int pps_server_poll(pps_server* sv, uint64_t ticket, pps_result* out);
/* Block until done (timeout_s <= 0: no timeout; PPS_E_TIMEOUT otherwise); the
 * final status is returned once, as for poll. */
int pps_server_wait(pps_server* sv, uint64_t ticket, double timeout_s,
                    pps_result* out);
/* Cancel a request that is still queued (PPS_E_STATE once admitted). */
int pps_server_cancel(pps_server* sv, uint64_t ticket);
/* Counters: submitted, completed, packed launches, peak admitted bytes. */
int pps_server_stats(pps_server* sv, int64_t* submitted, int64_t* completed,
                     int64_t* launches, int64_t* peak_bytes);
/* Finish the running problems, cancel queued ones, free everything. */
int pps_server_destroy(pps_server* sv);

/* Solve `count` independent problems through a temporary server (all
 * submitted at once); results in order; wall time in elapsed_ms. */
int pps_batch_solve(pps_ctx* ctx, int32_t count, const pps_problem* probs,
                    const pps_swarm_params* params, const int32_t* maxiters,
                    const float* w, const float* wdamp, pps_result* out,
                    float* elapsed_ms);

#ifdef __cplusplus
}
#endif
#endif /* PPS_H */
