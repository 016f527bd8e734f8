// pps_runtime.hip — host runtime behind the C ABI of include/pps.h:
// contexts, swarm ownership in HBM, the iteration drivers (eager launches, a
// captured hipGraph of the iteration loop, or one resident workgroup), result
// read-back and the FP64 host evaluator.
//
// Reference drivers this replaces: vectorized_solve! (src/lowerlevel_solve.jl:1-121),
// init/reinit! (src/init.jl:15-141), solve! result assembly (src/solve.jl:11-59).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <map>
#include <mutex>
#include <tuple>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "pps_kernels.h"
#include "pps_runtime.h"

using namespace pps;

namespace pps {

thread_local std::string g_err;

int check_step_coefs(float c1, float c2) {
  for (float c : {c1, c2})
    if (c != 0.0f && std::isfinite(c) && std::fabs(c) < 0x1.0p-102f)
      return fail(PPS_E_INVALID,
                  "c1 = %g, c2 = %g: nonzero coefficients below 2^-102 in magnitude are not "
                  "supported", (double)c1, (double)c2);
  return PPS_OK;
}

int fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

int hip_fail(hipError_t e, const char* what, const char* file, int line) {
  return fail(PPS_E_HIP, "%s: %s (%s:%d)", what, hipGetErrorString(e), file, line);
}

int fill_prob_args(const pps_problem* prob, ProbArgs* P) {
  if (!prob) return fail(PPS_E_INVALID, "problem is NULL");
  memset(P, 0, sizeof *P);
  // An adaptive-step ODE loss stops summing at the particle's best cost (the
  // swarm state is the same); PPS_ODE_FULL_LOSS=1 in the environment makes
  // every loss full.
  const char* full = getenv("PPS_ODE_FULL_LOSS");
  P->ode_stop = (full && full[0] && strcmp(full, "0") != 0) ? 0 : 1;
  const int d = prob->d;
  if (d < 1 || d > PPS_MAXD) return fail(PPS_E_INVALID, "dimension %d outside 1..%d", d, PPS_MAXD);
  if (!prob->lb != !prob->ub) return fail(PPS_E_INVALID, "lb and ub are both given or both NULL");
  P->obj = prob->objective;
  P->d = d;
  // No bounds = infinite bounds (check_init_bounds, src/utils.jl:310-318).
  // Bounds are either finite in every coordinate or -Inf/+Inf in every one
  // (the unbounded case, src/init.jl:70-74); a mix would make the uniform
  // initialiser produce non-finite positions and is refused.
  int n_inf = 0;
  for (int j = 0; j < d; ++j) {
    const float lo = prob->lb ? prob->lb[j] : -INFINITY, hi = prob->ub ? prob->ub[j] : INFINITY;
    if (std::isnan(lo) || std::isnan(hi)) return fail(PPS_E_INVALID, "bound %d is NaN", j);
    if (!(lo <= hi)) return fail(PPS_E_INVALID, "lb[%d] > ub[%d]", j, j);
    if (lo == -INFINITY && hi == INFINITY)
      ++n_inf;
    else if (!std::isfinite(lo) || !std::isfinite(hi))
      return fail(PPS_E_INVALID, "bounds must be finite, or -Inf/+Inf in every dimension "
                                 "(dimension %d)", j);
    P->lb[j] = lo;
    P->ub[j] = hi;
  }
  if (n_inf != 0 && n_inf != d)
    return fail(PPS_E_INVALID, "bounds must be finite, or -Inf/+Inf in every dimension");
  if (prob->n_params < 0 || prob->n_params > PPS_MAXPAR)
    return fail(PPS_E_INVALID, "n_params %d outside 0..%d", prob->n_params, PPS_MAXPAR);
  if (prob->n_params > 0 && !prob->params) return fail(PPS_E_INVALID, "params is NULL");
  for (int k = 0; k < prob->n_params; ++k) P->par[k] = prob->params[k];
  if (prob->shift) {
    P->has_shift = 1;
    for (int j = 0; j < d; ++j) P->shift[j] = prob->shift[j];
  }
  int need_par = 0;
  switch (prob->objective) {
    case PPS_OBJ_ROSENBROCK:
    case PPS_OBJ_ROSEN2: need_par = 2; break;
    case PPS_OBJ_QUADRATIC: need_par = 1; break;
    default: break;
  }
  if (is_ode_objective(prob->objective)) need_par = ode_adaptive(prob->objective) ? 6 : 4;
  if (prob->n_params < need_par)
    return fail(PPS_E_INVALID, "objective %d needs %d parameters, got %d", prob->objective,
                need_par, prob->n_params);
  if (is_ode_objective(prob->objective)) {
    if (d != 4) return fail(PPS_E_INVALID, "ODE estimation has 4 rates, d=%d", d);
    if (!prob->obs || prob->n_obs < 1) return fail(PPS_E_INVALID, "ODE estimation needs observations");
    P->n_obs = prob->n_obs;
    const int obj = prob->objective;
    const bool adaptive = ode_adaptive(obj);
    const bool saveat = ode_saveat(obj);
    if (adaptive || saveat) {
      const float* q = prob->params;
      if (!(q[2] > 0.0f && std::isfinite((float)(prob->n_obs - 1) * q[2])))
        return fail(PPS_E_INVALID, "observation interval must be positive (and n_obs times it finite)");
      if (!(q[3] > 0.0f && std::isfinite(q[3])))
        return fail(PPS_E_INVALID, adaptive ? "initial step must be positive" : "step must be positive");
      if (!adaptive && (double)(prob->n_obs - 1) * q[2] / q[3] > 1.0e7)
        return fail(PPS_E_INVALID, "more than 1e7 fixed steps per evaluation");
    }
    if (adaptive) {
      const float* q = prob->params;
      if (!(q[4] >= 0.0f && q[5] >= 0.0f && q[4] + q[5] > 0.0f && std::isfinite(q[4] + q[5])))
        return fail(PPS_E_INVALID, "abstol and reltol must be >= 0 with a positive sum");
      if (prob->n_params > 6 && !(q[6] >= 1.0f && q[6] <= 1.0e7f && q[6] == std::floor(q[6])))
        return fail(PPS_E_INVALID, "the step limit must be an integer in 1..1e7");
    } else if (!saveat) {
      const float spo = prob->params[3];
      if (!(spo >= 1.0f && spo <= 1.0e6f) || spo != std::floor(spo))
        return fail(PPS_E_INVALID, "steps per observation must be a positive integer");
      P->steps_per_obs = (int32_t)spo;
    }
  }
  if (prob->n_cons < 0 || prob->n_cons > PPS_MAXCONS)
    return fail(PPS_E_INVALID, "n_cons %d outside 0..%d", prob->n_cons, PPS_MAXCONS);
  if (prob->n_cons > 0) {
    if (!prob->cons) return fail(PPS_E_INVALID, "cons is NULL with n_cons = %d", prob->n_cons);
    if (prob->objective != PPS_OBJ_ROSENBROCK && prob->objective != PPS_OBJ_RASTRIGIN &&
        prob->objective != PPS_OBJ_SPHERE)
      return fail(PPS_E_INVALID,
                  "constraints are supported with the Rosenbrock, Rastrigin and Sphere "
                  "objectives (objective %d given)", prob->objective);
    for (int k = 0; k < prob->n_cons * (2 * d + 1); ++k)
      if (!std::isfinite(prob->cons[k]))
        return fail(PPS_E_INVALID, "constraint coefficient %d is not finite", k);
    P->n_cons = prob->n_cons;  // the coefficients are copied to device memory by the caller
  }
  return PPS_OK;
}

bool prob_unbounded(const ProbArgs& P) { return P.d > 0 && std::isinf(P.lb[0]); }

}  // namespace pps

// ---------------------------------------------------------------------------
struct pps_swarm {
  pps_ctx* ctx = nullptr;
  ProbArgs P{};
  pps_swarm_params prm{};
  const KernelSet* ks = nullptr;
  SwarmDev S{};
  void* mem = nullptr;
  size_t mem_bytes = 0;
  uint32_t iter = 0;  // completed iterations in this epoch
  int launch = PPS_LAUNCH_EAGER;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  float last_ms = 0.0f;
  // Captured chunk of kChunk sync launches (weights from the table), replayed
  // while the step arguments match; the run's weight table (device, grown on
  // demand) and its host copy.
  hipGraphExec_t cexec = nullptr;
  StepArgs c_args{};
  float* wtab = nullptr;
  int32_t wtab_cap = 0;
  std::vector<float> wtab_h;
  int32_t sampler = PPS_SAMPLER_UNIFORM;  // sampler of init / reinit
  // Local-search outputs (allocated on first use): x (d x npad), f, argmin.
  float* lx = nullptr;
  float* lf = nullptr;
  uint64_t* lkey = nullptr;
};

#define HIPCHK(expr)                                                   \
  do {                                                                 \
    hipError_t e_ = (expr);                                            \
    if (e_ != hipSuccess) return hip_fail(e_, #expr, __FILE__, __LINE__); \
  } while (0)

// Device scratch of the allocation-free local search (pps_swarm_lbfgs_noalloc).
// Every GPU the library runs on holds its own copy until the process ends.
// In host code `scratch` only names the symbol: kernels and copies use the
// device address looked up once per device (scratch_setup), never the name.
static __device__ __align__(256) unsigned char scratch[PPS_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 256;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxDevices] = {};

// With `device` current: resolve this device's copy of `scratch` (once).
int scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return fail(PPS_E_INVALID, "device %d", device);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_dev[device]) return PPS_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) return hip_fail(e, "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
  g_scratch_dev[device] = p;
  return PPS_OK;
}

char* scratch_of(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return (device >= 0 && device < kMaxDevices) ? (char*)g_scratch_dev[device] : nullptr;
}
}  // namespace

extern "C" {

const char* pps_last_error(void) { return g_err.c_str(); }
int pps_abi_version(void) { return PPS_ABI_VERSION; }

void pps_swarm_params_default(pps_swarm_params* p) {
  if (!p) return;
  memset(p, 0, sizeof *p);
  p->n = 100;
  p->mode = PPS_MODE_SYNC;
  p->gbest_from = PPS_GBEST_PBEST;
  p->clamp = PPS_CLAMP_COMPONENT;
  p->launch = PPS_LAUNCH_AUTO;
  p->seed = 0;
  p->c1 = 1.4962f;
  p->c2 = 1.4962f;
  p->qlock_min_improvers = 2;
  p->grid_per_cu = 0;
  p->sampler = PPS_SAMPLER_UNIFORM;
  p->particles_per_workgroup = 0;
}

int pps_ctx_create(int device, pps_ctx** out) {
  if (!out) return fail(PPS_E_INVALID, "out is NULL");
  *out = nullptr;
  int count = 0;
  HIPCHK(hipGetDeviceCount(&count));
  if (device < 0 || device >= count)
    return fail(PPS_E_INVALID, "device %d not present (%d devices)", device, count);
  HIPCHK(hipSetDevice(device));
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, device));
  const int src = scratch_setup(device);
  if (src != PPS_OK) return src;
  pps_ctx* c = new pps_ctx();
  c->device = device;
  c->num_cus = prop.multiProcessorCount;
  c->clock_khz = prop.clockRate;
  c->total_mem = prop.totalGlobalMem;
  c->arch = prop.gcnArchName;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return hip_fail(e, "hipStreamCreateWithFlags", __FILE__, __LINE__);
  }
  *out = c;
  return PPS_OK;
}

int pps_ctx_destroy(pps_ctx* c) {
  if (!c) return PPS_OK;
  hipSetDevice(c->device);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return PPS_OK;
}

int pps_ctx_info(pps_ctx* c, int32_t* num_cus, int32_t* clock_khz, int64_t* total_mem,
                 int64_t* free_mem) {
  if (!c) return fail(PPS_E_INVALID, "ctx is NULL");
  HIPCHK(hipSetDevice(c->device));
  if (num_cus) *num_cus = c->num_cus;
  if (clock_khz) *clock_khz = c->clock_khz;
  if (total_mem) *total_mem = (int64_t)c->total_mem;
  if (free_mem) {
    size_t fr = 0, tot = 0;
    HIPCHK(hipMemGetInfo(&fr, &tot));
    *free_mem = (int64_t)fr;
  }
  return PPS_OK;
}

// ---------------------------------------------------------------------------
static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Default workgroups per CU of the multi-workgroup sync step (each reduces all
// G block keys at its start, so G stays a small multiple of the CU count).
#ifndef PPS_SYNC_GRID_PER_CU
#define PPS_SYNC_GRID_PER_CU 2
#endif

}  // extern "C"

namespace {

// Autotuned (grid_per_cu, particles_per_workgroup) keyed by
// (objective, d, ceil(log2 n)).
struct TuneTable {
  std::mutex mu;
  std::map<std::tuple<int, int, int>, std::pair<int, int>> best;
  bool loaded = false;
};
TuneTable& tune_table() {
  static TuneTable t;
  return t;
}
int log2_bucket(int64_t n) {
  int b = 0;
  while (((int64_t)1 << b) < n) ++b;
  return b;
}
void tune_load_locked(TuneTable& t) {
  if (t.loaded) return;
  t.loaded = true;
  const char* path = getenv("PPS_TUNE_CACHE");
  if (!path) return;
  FILE* f = fopen(path, "r");
  if (!f) return;
  char line[256];
  while (fgets(line, sizeof line, f)) {
    int obj, d, b, g, ppw = 256;
    const int got = sscanf(line, "%d %d %d %d %d", &obj, &d, &b, &g, &ppw);
    if (got < 4) continue;
    if (g > 0 && g <= 64 && (ppw == 64 || ppw == 256 || ppw == 512 || ppw == 1024))
      t.best[std::make_tuple(obj, d, b)] = std::make_pair(g, ppw);
  }
  fclose(f);
}

}  // namespace

extern "C" {

extern "C" int pps_tuned_launch(int32_t obj, int32_t d, int64_t n, int32_t* grid_per_cu,
                                int32_t* particles_per_workgroup) {
  TuneTable& t = tune_table();
  std::lock_guard<std::mutex> lk(t.mu);
  tune_load_locked(t);
  auto it = t.best.find(std::make_tuple((int)obj, (int)d, log2_bucket(n)));
  if (it == t.best.end()) return PPS_E_NOTFOUND;
  if (grid_per_cu) *grid_per_cu = it->second.first;
  if (particles_per_workgroup) *particles_per_workgroup = it->second.second;
  return PPS_OK;
}

static int swarm_prepare_launch(pps_swarm* s) {
  const int64_t n = s->prm.n;
  const int mode = s->prm.mode;
  int launch = s->prm.launch;
  if (mode == PPS_MODE_SYNC) {
    // Up to 1024 particles one resident workgroup runs the whole loop.  Above,
    // replayed chunk graphs save the host's ~1 us of issue per launch while a
    // launch is short (500 iterations, us per iteration one by one against
    // replayed: 10-D 2^14 5.75 / 5.27, 2^16 6.33 / 5.96; 3-D 2^16 4.80 /
    // 4.25, 2^18 6.10 / 5.75), and from 2^17 particles at d > 4 launches
    // issued one by one run the kernels as close together or closer (10-D
    // 2^17 7.27 / 7.47, 196608 8.54 / 8.75; 32-D 2^17 28.1 / 28.4; 2^20
    // 10-D 37.0 / 37.2), the host issuing far ahead of the device.
    if (launch == PPS_LAUNCH_AUTO) {
      const bool short_launch =
          n < ((int64_t)1 << 17) || (n < ((int64_t)1 << 19) && s->P.d <= 4);
      launch = n <= 1024 ? PPS_LAUNCH_WORKGROUP
                         : (short_launch ? PPS_LAUNCH_GRAPH : PPS_LAUNCH_EAGER);
    }
    if (launch == PPS_LAUNCH_WORKGROUP && n > 1024)
      return fail(PPS_E_INVALID, "workgroup launch holds at most 1024 particles (n=%lld)",
                  (long long)n);
  } else {
    launch = PPS_LAUNCH_EAGER;
  }
  s->launch = launch;
  if (launch == PPS_LAUNCH_WORKGROUP) {
    s->S.G = 1;
    s->S.block = 1024;
  } else {
    int blk = block_of(s->P.obj);
    int tuned_g = 0, tuned_b = 0;
    if (mode == PPS_MODE_SYNC && (s->prm.grid_per_cu == 0 || s->prm.particles_per_workgroup == 0))
      pps_tuned_launch(s->P.obj, s->P.d, n, &tuned_g, &tuned_b);
    if (mode == PPS_MODE_SYNC) {
      // Workgroups of 512 or 1024 particles where the kernels are compiled
      // for them and every state array is below 4 GiB (32-bit row offsets).
      const bool small_arrays = (uint64_t)((n + 63) / 64 * 64) * (uint64_t)s->P.d * 4u < (1ull << 32);
      const int want = s->prm.particles_per_workgroup > 0 ? s->prm.particles_per_workgroup : tuned_b;
      if (want > 0 && want != blk) {
        if (want <= s->ks->max_block && small_arrays)
          blk = want;
        else if (s->prm.particles_per_workgroup > 0)
          return fail(PPS_E_INVALID,
                      "particles_per_workgroup %d is compiled for Rosenbrock, Rastrigin and "
                      "Sphere at d in {1,2,3,4,8,10,16} (512 also at d = 32) without "
                      "constraints, with state arrays below 4 GiB", want);
      }
    }
    s->S.block = blk;
    const int64_t tiles = (n + blk - 1) / blk;
    int per_cu = PPS_SYNC_GRID_PER_CU;
    if (mode == PPS_MODE_ASYNC) {
      per_cu = 8;
    } else if (s->prm.grid_per_cu > 0) {
      per_cu = s->prm.grid_per_cu;
    } else if (tuned_g > 0) {
      per_cu = tuned_g;
    }
    const int64_t cap = (int64_t)s->ctx->num_cus * per_cu;
    s->S.G = (int32_t)std::max<int64_t>(1, std::min<int64_t>(tiles, cap));
  }
  return PPS_OK;
}

static int swarm_alloc(pps_swarm* s, const pps_problem* prob) {
  const int d = s->P.d;
  const int64_t n = s->prm.n;
  const int64_t npad = (n + 63) / 64 * 64;
  const int G = s->S.G;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off = align256(off + bytes);
    return o;
  };
  const size_t o_x = take(sizeof(float) * d * npad);
  const size_t o_v = take(sizeof(float) * d * npad);
  const size_t o_pb = take(sizeof(float) * d * npad);
  const size_t o_pc = take(sizeof(float) * npad);
  const size_t o_bkey = take(sizeof(uint64_t) * 2 * G);
  const size_t o_cand = take(sizeof(float) * 2 * G * d);
  const size_t o_rec = take(sizeof(GbRec));
  const size_t o_t = take(sizeof(uint32_t) * 64);
  const size_t o_obs = take(sizeof(float) * 2 * (size_t)std::max(0, s->P.n_obs));
  const size_t o_cons = take(sizeof(float) * (size_t)s->P.n_cons * (2 * d + 1));
  const size_t o_u0 = take(sizeof(float) * d);
  s->mem_bytes = off;
  HIPCHK(hipSetDevice(s->ctx->device));
  hipError_t e = hipMalloc(&s->mem, off);
  if (e != hipSuccess) {
    s->mem = nullptr;
    return fail(PPS_E_NOMEM, "hipMalloc of %zu bytes of swarm state failed: %s", off,
                hipGetErrorString(e));
  }
  char* base = (char*)s->mem;
  s->S.x = (float*)(base + o_x);
  s->S.v = (float*)(base + o_v);
  s->S.pb = (float*)(base + o_pb);
  s->S.pc = (float*)(base + o_pc);
  s->S.bkey = (uint64_t*)(base + o_bkey);
  s->S.cand = (float*)(base + o_cand);
  s->S.gbrec = (GbRec*)(base + o_rec);
  s->S.tbase = (const uint32_t*)(base + o_t);
  s->S.n = n;
  s->S.npad = npad;
  s->S.key0 = (uint32_t)s->prm.seed;
  s->S.key1 = (uint32_t)(s->prm.seed >> 32);
  HIPCHK(hipMemsetAsync(s->mem, 0, off, s->ctx->stream));
  if (s->P.n_obs > 0) {
    float* obs = (float*)(base + o_obs);
    HIPCHK(hipMemcpyAsync(obs, prob->obs, sizeof(float) * 2 * s->P.n_obs, hipMemcpyHostToDevice,
                          s->ctx->stream));
    s->P.obs = obs;
  }
  s->S.u0 = (const float*)(base + o_u0);
  if (prob->u0)
    HIPCHK(hipMemcpyAsync(base + o_u0, prob->u0, sizeof(float) * d, hipMemcpyHostToDevice,
                          s->ctx->stream));
  if (s->P.n_cons > 0) {
    float* cons = (float*)(base + o_cons);
    HIPCHK(hipMemcpyAsync(cons, prob->cons, sizeof(float) * s->P.n_cons * (2 * d + 1),
                          hipMemcpyHostToDevice, s->ctx->stream));
    s->P.cons = cons;
  }
  return PPS_OK;
}

// Uniform init of the current epoch + global best of the initial swarm.
// Initialise the swarm of the current epoch: sample (or take `positions`,
// n x d particle-major), evaluate, global best of the initial swarm.
static int swarm_do_init(pps_swarm* s, const float* positions) {
  s->iter = 0;
  SwarmDev S = s->S;
  S.sampler = s->sampler;
  if (positions) {
    const int d = s->P.d;
    const int64_t n = s->prm.n, npad = s->S.npad;
    for (int64_t k = 0; k < n * d; ++k)
      if (!std::isfinite(positions[k]))
        return fail(PPS_E_INVALID, "initial position %lld is not finite", (long long)k);
    std::vector<float> rows((size_t)d * npad, 0.0f);
    for (int64_t i = 0; i < n; ++i)
      for (int j = 0; j < d; ++j) rows[(size_t)j * npad + i] = positions[i * d + j];
    HIPCHK(hipMemcpyAsync(S.x, rows.data(), sizeof(float) * rows.size(), hipMemcpyHostToDevice,
                          s->ctx->stream));
    S.sampler = PPS_SAMPLER_POSITIONS;
    HIPCHK(s->ks->init(S, s->P, 0u, s->ctx->stream));
    HIPCHK(launch_finalize(S, s->P.d, 0u, s->ctx->stream));
    HIPCHK(hipStreamSynchronize(s->ctx->stream));  // `rows` is read by the copy
    return PPS_OK;
  }
  HIPCHK(s->ks->init(S, s->P, 0u, s->ctx->stream));
  HIPCHK(launch_finalize(S, s->P.d, 0u, s->ctx->stream));
  HIPCHK(hipStreamSynchronize(s->ctx->stream));
  return PPS_OK;
}

static int swarm_create(pps_ctx* ctx, const pps_problem* prob, const pps_swarm_params* prm,
                        const float* positions, pps_swarm** out) {
  if (!ctx || !prob || !prm || !out) return fail(PPS_E_INVALID, "NULL argument");
  *out = nullptr;
  if (prm->sampler < PPS_SAMPLER_UNIFORM || prm->sampler > PPS_SAMPLER_LATIN_HYPERCUBE)
    return fail(PPS_E_INVALID, "unknown sampler %d", prm->sampler);
  if (prm->n < 1 || prm->n > (int64_t)0x7FFFFFFF)
    return fail(PPS_E_INVALID, "num_particles %lld outside 1..2^31-1", (long long)prm->n);
  if (prm->mode < PPS_MODE_SYNC || prm->mode > PPS_MODE_QLOCK)
    return fail(PPS_E_INVALID, "unknown mode %d", prm->mode);
  if (check_step_coefs(prm->c1, prm->c2) != PPS_OK) return PPS_E_INVALID;
  if (prm->gbest_from != PPS_GBEST_PBEST && prm->gbest_from != PPS_GBEST_CURRENT)
    return fail(PPS_E_INVALID, "unknown gbest_from %d", prm->gbest_from);
  if (prm->clamp != PPS_CLAMP_COMPONENT && prm->clamp != PPS_CLAMP_LEXICOGRAPHIC)
    return fail(PPS_E_INVALID, "unknown clamp %d", prm->clamp);
  if (prm->clamp == PPS_CLAMP_LEXICOGRAPHIC && prob && !is_ode_objective(prob->objective))
    return fail(PPS_E_INVALID,
                "lexicographic clamping is the ODE estimation path's; it needs "
                "an ODE objective (objective %d given)", prob->objective);
  if (prm->grid_per_cu < 0 || prm->grid_per_cu > 64)
    return fail(PPS_E_INVALID, "grid_per_cu %d outside 0..64", prm->grid_per_cu);
  // 0, the objective's own workgroup size (64 particles for the ODE
  // objectives, 256 otherwise) or one of the wider ones
  const int own_block = prob ? block_of(prob->objective) : 256;
  if (prm->particles_per_workgroup != 0 && prm->particles_per_workgroup != own_block &&
      prm->particles_per_workgroup != 256 && prm->particles_per_workgroup != 512 &&
      prm->particles_per_workgroup != 1024)
    return fail(PPS_E_INVALID,
                "particles_per_workgroup %d is not 0, %d (this objective's own), 256, 512 or 1024",
                prm->particles_per_workgroup, own_block);
  pps_swarm* s = new pps_swarm();
  s->ctx = ctx;
  s->prm = *prm;
  int rc = fill_prob_args(prob, &s->P);
  if (rc == PPS_OK) {
    s->P.clamp = prm->clamp;
    s->sampler = prob_unbounded(s->P) ? PPS_SAMPLER_UNBOUNDED : prm->sampler;
    s->ks = find_kernels(s->P.obj, s->P.d, s->P.n_cons > 0);
    if (!s->ks) rc = fail(PPS_E_INVALID, "no kernels for objective %d at d=%d", s->P.obj, s->P.d);
  }
  if (rc == PPS_OK) rc = swarm_prepare_launch(s);
  if (rc == PPS_OK) rc = swarm_alloc(s, prob);
  if (rc == PPS_OK) {
    hipError_t e = hipEventCreate(&s->ev0);
    if (e == hipSuccess) e = hipEventCreate(&s->ev1);
    if (e != hipSuccess) rc = hip_fail(e, "hipEventCreate", __FILE__, __LINE__);
  }
  if (rc == PPS_OK) {
    s->S.epoch = 0;
    rc = swarm_do_init(s, positions);
  }
  if (rc != PPS_OK) {
    std::string keep = g_err;
    pps_swarm_free(s);
    g_err = keep;
    return rc;
  }
  *out = s;
  return PPS_OK;
}

int pps_swarm_init(pps_ctx* ctx, const pps_problem* prob, const pps_swarm_params* prm,
                   pps_swarm** out) {
  return swarm_create(ctx, prob, prm, nullptr, out);
}

int pps_swarm_init_positions(pps_ctx* ctx, const pps_problem* prob, const pps_swarm_params* prm,
                             const float* positions, pps_swarm** out) {
  if (!positions) return fail(PPS_E_INVALID, "positions is NULL");
  return swarm_create(ctx, prob, prm, positions, out);
}

int pps_swarm_reinit(pps_swarm* s) {
  if (!s) return fail(PPS_E_INVALID, "swarm is NULL");
  HIPCHK(hipSetDevice(s->ctx->device));
  s->S.epoch += 1;
  return swarm_do_init(s, nullptr);
}

int pps_swarm_reinit_positions(pps_swarm* s, const float* positions) {
  if (!s || !positions) return fail(PPS_E_INVALID, "NULL argument");
  HIPCHK(hipSetDevice(s->ctx->device));
  s->S.epoch += 1;
  return swarm_do_init(s, positions);
}

// Capture (or reuse) the graph of T sync launches; launched by pps_swarm_run
// after its start event, so the run's device time excludes the capture.  The
// kernels read their stream position (iteration base and epoch) from device
// memory, so a graph stays valid across runs and reinitialisations.
// Sync launches per captured chunk.  A chunk graph holds kChunk launches of
// the step kernel, whose weights come from the run's table, and the advance of
// the stream position; any run replays it T / kChunk times and launches the
// remaining iterations one by one.  Capturing costs ~4 us per launch once per
// swarm (and step arguments); a replayed launch is ~1 us shorter than one
// issued by the host, and each replay costs ~7.7 us at its boundary plus
// ~0.07 us per launch for the table read (10-D, 2048-65536 particles, us per
// iteration: all 512 launches in one graph 4.47-5.09; chunks of 16 4.96-5.59,
// 64 4.66-5.30, 128 4.60-5.25; one by one 4.7-6.0).  32 keeps the capture of
// a one-shot solve cheap.
constexpr int32_t kChunk = 32;

static int prepare_chunk_graph(pps_swarm* s, const StepArgs& A) {
  hipStream_t st = s->ctx->stream;
  if (s->cexec && s->c_args.c1 == A.c1 && s->c_args.c2 == A.c2 &&
      s->c_args.gbest_from == A.gbest_from)
    return PPS_OK;
  if (s->cexec) {
    hipGraphExecDestroy(s->cexec);
    s->cexec = nullptr;
  }
  StepArgs At = A;
  At.w_table = 1;
  hipGraph_t g = nullptr;
  HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  hipError_t le = hipSuccess;
  for (int32_t k = 1; k <= kChunk && le == hipSuccess; ++k)
    le = s->ks->sync_step(s->S, s->P, At, 0.0f, (uint32_t)k, st);
  if (le == hipSuccess) le = launch_advance_tbase((uint32_t*)s->S.tbase, (uint32_t)kChunk, st);
  hipError_t ce = hipStreamEndCapture(st, &g);
  if (le != hipSuccess) {
    if (g) hipGraphDestroy(g);
    return hip_fail(le, "sync_step (capture)", __FILE__, __LINE__);
  }
  if (ce != hipSuccess) return hip_fail(ce, "hipStreamEndCapture", __FILE__, __LINE__);
  hipError_t ie = hipGraphInstantiate(&s->cexec, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (ie != hipSuccess) {
    s->cexec = nullptr;
    return hip_fail(ie, "hipGraphInstantiate", __FILE__, __LINE__);
  }
  s->c_args = A;
  return PPS_OK;
}

// The run's weights w_1 = w, w_{k+1} = w_k * wdamp (FP32, in order, as the
// launches issued one by one get them) into the device table.  A larger
// table moves, and the captured chunk (which holds its address) is dropped.
static int upload_weights(pps_swarm* s, int32_t T, float w, float wdamp) {
  s->wtab_h.resize((size_t)T);
  float wk = w;
  for (int32_t k = 0; k < T; ++k) {
    s->wtab_h[(size_t)k] = wk;
    wk = wk * wdamp;
  }
  if (T > s->wtab_cap) {
    const int32_t cap = std::max(T, std::max(1024, 2 * s->wtab_cap));
    // the captured chunk holds the old table's address: drop both first
    if (s->cexec) {
      hipGraphExecDestroy(s->cexec);
      s->cexec = nullptr;
    }
    if (s->wtab) {
      HIPCHK(hipStreamSynchronize(s->ctx->stream));
      HIPCHK(hipFree(s->wtab));
      s->wtab = nullptr;
      s->wtab_cap = 0;
      s->S.wtab = nullptr;
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, sizeof(float) * (size_t)cap);
    if (e != hipSuccess)
      return fail(PPS_E_NOMEM, "hipMalloc of the %d-entry weight table failed: %s", cap,
                  hipGetErrorString(e));
    s->wtab = (float*)p;
    s->wtab_cap = cap;
    s->S.wtab = s->wtab;
  }
  HIPCHK(hipMemcpyAsync(s->wtab, s->wtab_h.data(), sizeof(float) * (size_t)T,
                        hipMemcpyHostToDevice, s->ctx->stream));
  return PPS_OK;
}

int pps_swarm_run(pps_swarm* s, int32_t maxiters, float w, float wdamp) {
  if (!s) return fail(PPS_E_INVALID, "swarm is NULL");
  if (maxiters < 0) return fail(PPS_E_INVALID, "maxiters < 0");
  HIPCHK(hipSetDevice(s->ctx->device));
  hipStream_t st = s->ctx->stream;
  StepArgs A{s->prm.c1, s->prm.c2, s->prm.gbest_from, s->prm.qlock_min_improvers};
  const uint32_t t0 = s->iter;
  const int32_t T = maxiters;
  // Graph launch: chunks of kChunk launches replayed from a captured graph.
  // The automatic mode captures only for a run long enough to pay for it
  // (or once captured).
  bool graph = T >= kChunk && s->prm.mode == PPS_MODE_SYNC && s->launch == PPS_LAUNCH_GRAPH;
  if (graph && s->prm.launch == PPS_LAUNCH_AUTO && !s->cexec && T < 4 * kChunk) graph = false;
  if (graph) {
    int rc = upload_weights(s, T, w, wdamp);
    if (rc != PPS_OK) return rc;
    rc = prepare_chunk_graph(s, A);
    if (rc != PPS_OK) return rc;
  }
  // Stream position of this run (read by the kernels, so a captured chunk
  // replays after a reinit too); a queue-lock run starts with a clear
  // give-up flag.
  HIPCHK(launch_set_tbase((uint32_t*)s->S.tbase, t0, s->S.epoch, st));
  if (s->prm.mode == PPS_MODE_QLOCK) HIPCHK(hipMemsetAsync(&s->S.gbrec->err, 0, sizeof(uint32_t), st));
  HIPCHK(hipEventRecord(s->ev0, st));
  if (T > 0) {
    switch (s->prm.mode) {
      case PPS_MODE_SYNC:
        if (s->launch == PPS_LAUNCH_WORKGROUP) {
          HIPCHK(s->ks->sync_wg(s->S, s->P, A, w, wdamp, t0, T, st));
        } else if (graph) {
          const int32_t chunks = T / kChunk;
          for (int32_t c = 0; c < chunks; ++c) HIPCHK(hipGraphLaunch(s->cexec, st));
          // the rest one by one: the chunks advanced the device's iteration base
          for (int32_t k = 1; k <= T - chunks * kChunk; ++k)
            HIPCHK(s->ks->sync_step(s->S, s->P, A, s->wtab_h[(size_t)(chunks * kChunk + k - 1)],
                                    (uint32_t)k, st));
        } else {
          float wk = w;
          for (int32_t k = 1; k <= T; ++k) {
            HIPCHK(s->ks->sync_step(s->S, s->P, A, wk, (uint32_t)k, st));
            wk = wk * wdamp;
          }
        }
        break;
      case PPS_MODE_ASYNC:
        HIPCHK(s->ks->async_run(s->S, s->P, A, w, wdamp, t0, T, st));
        break;
      case PPS_MODE_QLOCK: {
        // One workgroup's worth of particles (one queue-lock block): the T
        // iterations in one launch, the live global best in LDS.
        if (s->prm.n <= 1024) {
          HIPCHK(s->ks->qlock_wg(s->S, s->P, A, w, wdamp, T, st));
          break;
        }
        float wk = w;
        for (int32_t k = 1; k <= T; ++k) {
          HIPCHK(s->ks->qlock_step(s->S, s->P, A, wk, (uint32_t)k, st));
          wk = wk * wdamp;
        }
        break;
      }
    }
  }
  HIPCHK(hipEventRecord(s->ev1, st));
  s->iter = t0 + (uint32_t)T;
  if (s->prm.mode != PPS_MODE_QLOCK)
    HIPCHK(launch_finalize(s->S, s->P.d, s->iter & 1u, st));
  HIPCHK(hipStreamSynchronize(st));
  HIPCHK(hipEventElapsedTime(&s->last_ms, s->ev0, s->ev1));
  if (s->prm.mode == PPS_MODE_QLOCK) {
    uint32_t err = 0;
    HIPCHK(hipMemcpy(&err, &s->S.gbrec->err, sizeof err, hipMemcpyDeviceToHost));
    if (err) return fail(PPS_E_STATE, "queue-lock publication gave up waiting for the lock");
  }
  return PPS_OK;
}

int pps_swarm_get_gbest(pps_swarm* s, float* pos, float* cost, int64_t* idx) {
  if (!s) return fail(PPS_E_INVALID, "swarm is NULL");
  HIPCHK(hipSetDevice(s->ctx->device));
  GbRec r;
  HIPCHK(hipStreamSynchronize(s->ctx->stream));
  HIPCHK(hipMemcpy(&r, s->S.gbrec, sizeof r, hipMemcpyDeviceToHost));
  if (pos) memcpy(pos, r.pos, sizeof(float) * s->P.d);
  if (cost) *cost = r.cost;
  if (idx) *idx = (int64_t)r.idx;
  return PPS_OK;
}

int pps_swarm_set_gbest(pps_swarm* s, const float* pos, float cost) {
  if (!s || !pos) return fail(PPS_E_INVALID, "NULL argument");
  HIPCHK(hipSetDevice(s->ctx->device));
  HIPCHK(hipStreamSynchronize(s->ctx->stream));
  // Stage the position in the global-best record, then fan it out to the
  // hand-over slots of the next iteration.
  HIPCHK(hipMemcpy(s->S.gbrec->pos, pos, sizeof(float) * s->P.d, hipMemcpyHostToDevice));
  HIPCHK(launch_set_gbest(s->S, s->P.d, s->iter & 1u, s->S.gbrec->pos, cost, s->ctx->stream));
  HIPCHK(hipStreamSynchronize(s->ctx->stream));
  return PPS_OK;
}

int pps_swarm_get_positions(pps_swarm* s, float* out) {
  if (!s || !out) return fail(PPS_E_INVALID, "NULL argument");
  const int d = s->P.d;
  const int64_t n = s->prm.n;
  std::vector<float> tmp((size_t)d * n);
  HIPCHK(hipSetDevice(s->ctx->device));
  HIPCHK(hipStreamSynchronize(s->ctx->stream));
  HIPCHK(hipMemcpy2D(tmp.data(), sizeof(float) * n, s->S.x, sizeof(float) * s->S.npad,
                     sizeof(float) * n, d, hipMemcpyDeviceToHost));
  for (int64_t i = 0; i < n; ++i)
    for (int j = 0; j < d; ++j) out[i * d + j] = tmp[(size_t)j * n + i];
  return PPS_OK;
}

int pps_swarm_get_state(pps_swarm* s, float* x, float* v, float* best, float* best_cost,
                        uint32_t* iter) {
  if (!s) return fail(PPS_E_INVALID, "swarm is NULL");
  const int d = s->P.d;
  const int64_t n = s->prm.n;
  HIPCHK(hipSetDevice(s->ctx->device));
  HIPCHK(hipStreamSynchronize(s->ctx->stream));
  const size_t w = sizeof(float) * n, sp = sizeof(float) * s->S.npad;
  if (x) HIPCHK(hipMemcpy2D(x, w, s->S.x, sp, w, d, hipMemcpyDeviceToHost));
  if (v) HIPCHK(hipMemcpy2D(v, w, s->S.v, sp, w, d, hipMemcpyDeviceToHost));
  if (best) HIPCHK(hipMemcpy2D(best, w, s->S.pb, sp, w, d, hipMemcpyDeviceToHost));
  if (best_cost) HIPCHK(hipMemcpy(best_cost, s->S.pc, w, hipMemcpyDeviceToHost));
  if (iter) *iter = s->iter;
  return PPS_OK;
}

int pps_swarm_set_state(pps_swarm* s, const float* x, const float* v, const float* best,
                        const float* best_cost, uint32_t iter) {
  if (!s) return fail(PPS_E_INVALID, "swarm is NULL");
  const int d = s->P.d;
  const int64_t n = s->prm.n;
  HIPCHK(hipSetDevice(s->ctx->device));
  HIPCHK(hipStreamSynchronize(s->ctx->stream));
  const size_t w = sizeof(float) * n, sp = sizeof(float) * s->S.npad;
  if (x) HIPCHK(hipMemcpy2D(s->S.x, sp, x, w, w, d, hipMemcpyHostToDevice));
  if (v) HIPCHK(hipMemcpy2D(s->S.v, sp, v, w, w, d, hipMemcpyHostToDevice));
  if (best) HIPCHK(hipMemcpy2D(s->S.pb, sp, best, w, w, d, hipMemcpyHostToDevice));
  if (best_cost) HIPCHK(hipMemcpy(s->S.pc, best_cost, w, hipMemcpyHostToDevice));
  s->iter = iter;
  // Re-derive the hand-over state (block keys, candidates, global best).
  HIPCHK(s->ks->rekey(s->S, s->P, iter & 1u, s->prm.gbest_from, s->ctx->stream));
  HIPCHK(launch_finalize(s->S, d, iter & 1u, s->ctx->stream));
  HIPCHK(hipStreamSynchronize(s->ctx->stream));
  return PPS_OK;
}

int pps_swarm_elapsed_ms(pps_swarm* s, float* ms) {
  if (!s || !ms) return fail(PPS_E_INVALID, "NULL argument");
  *ms = s->last_ms;
  return PPS_OK;
}

int pps_swarm_free(pps_swarm* s) {
  if (!s) return PPS_OK;
  if (s->ctx) hipSetDevice(s->ctx->device);
  if (s->ctx && s->ctx->stream) hipStreamSynchronize(s->ctx->stream);
  if (s->cexec) hipGraphExecDestroy(s->cexec);
  if (s->wtab) hipFree(s->wtab);
  if (s->ev0) hipEventDestroy(s->ev0);
  if (s->ev1) hipEventDestroy(s->ev1);
  if (s->mem) hipFree(s->mem);
  if (s->lx) hipFree(s->lx);
  delete s;
  return PPS_OK;
}

int pps_autotune_sync(pps_ctx* ctx, const pps_problem* prob, int64_t n, int32_t iters,
                      int32_t* best_grid_per_cu, int32_t* best_particles_per_workgroup,
                      float* best_us_per_iter) {
  if (!ctx || !prob) return fail(PPS_E_INVALID, "NULL argument");
  if (n <= 1024) return fail(PPS_E_INVALID, "swarms of at most 1024 particles run in one workgroup");
  if (iters < 1) return fail(PPS_E_INVALID, "iters < 1");
  const KernelSet* ks = find_kernels(prob->objective, prob->d, prob->n_cons > 0);
  if (!ks) return fail(PPS_E_INVALID, "no kernels for objective %d at d=%d", prob->objective, prob->d);
  static const int grids[] = {1, 2, 3, 4, 6, 8};
  // the objective's own workgroup size first (64 particles for the ODE
  // objectives, 256 otherwise), then the wider ones its kernels are compiled for
  const int default_block = block_of(prob->objective);
  const int blocks[] = {default_block, 512, 1024};
  const bool small_arrays = (uint64_t)((n + 63) / 64 * 64) * (uint64_t)prob->d * 4u < (1ull << 32);
  int best_g = 0, best_b = 0;
  float best_us = 0.0f;
  for (int bi = 0; bi < 3; ++bi) {
    const int b = blocks[bi];
    if (bi > 0 && (b <= default_block || b > ks->max_block || !small_arrays)) continue;
    for (int g : grids) {
      pps_swarm_params prm;
      pps_swarm_params_default(&prm);
      prm.n = n;
      prm.launch = PPS_LAUNCH_AUTO;  // as the swarms it is tuned for (n > 1024)
      prm.grid_per_cu = g;
      prm.particles_per_workgroup = b;
      // Two scratch swarms of the size, run in turn, as a workload holding
      // more than one swarm: their states compete for the Infinity Cache,
      // where workgroups of fewer waves hide memory latency worse (2^20 x
      // 10-D in 1000-iteration runs of two swarms in turn, us per launch: 256
      // particles at 1, 2, 3 per CU 44.2, 43.0, 41.9; 512 at 2 37.3), while a
      // lone scratch swarm had timed 256 at 1 and 3 per CU at 35.5-35.7.  One
      // swarm only if a second does not fit.
      pps_swarm* sw[2] = {nullptr, nullptr};
      int rc = PPS_OK;
      for (int q = 0; q < 2 && rc == PPS_OK; ++q) {
        prm.seed = 0x7E57u + (uint64_t)q;
        rc = pps_swarm_init(ctx, prob, &prm, &sw[q]);
        if (rc == PPS_E_NOMEM && q == 1) {
          sw[1] = nullptr;
          rc = PPS_OK;
        }
      }
      if (rc != PPS_OK) {
        pps_swarm_free(sw[0]);
        return rc;
      }
      const int ns = sw[1] ? 2 : 1;
      float ms = 0.0f, best_ms = 1e30f;
      // Timed past the first thousand iterations, where a solve spends most of
      // its time: early on most particles improve and write their best
      // positions, and the fastest shape there is not always the fastest
      // later (2^20 x 10-D: 4 workgroups per CU lead 2 per CU by 3 us at
      // iteration 100 and trail or tie them from iteration 1000).  These
      // runs launch as the automatic mode does at this size (replayed chunk
      // graphs or one by one); the best of three is kept.
      for (int q = 0; q < ns && rc == PPS_OK; ++q)
        rc = pps_swarm_run(sw[q], std::max(iters, 1000), 0.7298f, 1.0f);
      for (int rep = 0; rep < 3 && rc == PPS_OK; ++rep) {
        float sum = 0.0f;
        for (int q = 0; q < ns && rc == PPS_OK; ++q) {
          rc = pps_swarm_run(sw[q], iters, 0.7298f, 1.0f);
          if (rc == PPS_OK) rc = pps_swarm_elapsed_ms(sw[q], &ms);
          sum += ms;
        }
        best_ms = std::min(best_ms, sum / (float)ns);
      }
      pps_swarm_free(sw[0]);
      pps_swarm_free(sw[1]);
      if (rc != PPS_OK) return rc;
      const float us = best_ms * 1e3f / (float)iters;
      if (best_g == 0 || us < best_us) {
        best_g = g;
        best_b = b;
        best_us = us;
      }
    }
  }
  if (best_g == 0) return fail(PPS_E_STATE, "no launch shape was timed");
  {
    TuneTable& t = tune_table();
    std::lock_guard<std::mutex> lk(t.mu);
    tune_load_locked(t);
    t.best[std::make_tuple((int)prob->objective, (int)prob->d, log2_bucket(n))] =
        std::make_pair(best_g, best_b);
    if (const char* path = getenv("PPS_TUNE_CACHE")) {
      if (FILE* f = fopen(path, "a")) {
        fprintf(f, "%d %d %d %d %d\n", prob->objective, prob->d, log2_bucket(n), best_g, best_b);
        fclose(f);
      }
    }
  }
  if (best_grid_per_cu) *best_grid_per_cu = best_g;
  if (best_particles_per_workgroup) *best_particles_per_workgroup = best_b;
  if (best_us_per_iter) *best_us_per_iter = best_us;
  return PPS_OK;
}

void pps_lbfgs_params_default(pps_lbfgs_params* p) {
  if (!p) return;
  p->maxiters = 100;
  p->history = 10;
  p->abstol = 0.0f;
  p->armijo_c = 1e-4f;
  p->armijo_rho = 0.5f;
  p->max_backtracks = 20;
  p->linesearch = PPS_LINESEARCH_STRONG_WOLFE;
  p->wolfe_c2 = 0.9f;
  p->alpha_init = 1.0f;
  p->alpha_max = 65536.0f;
  p->ls_maxiters = 20;
  p->zoom_maxiters = 30;
  p->reltol = 0.0f;
}

}  // extern "C"

static int lbfgs_check(const pps_swarm* s, const pps_lbfgs_params* lp) {
  if (!s || !lp) return fail(PPS_E_INVALID, "NULL argument");
  if (!lbfgs_supported(s->P.obj, s->P.d))
    return fail(PPS_E_INVALID, "no analytic gradient for objective %d at d=%d", s->P.obj, s->P.d);
  if (s->P.n_cons > 0)
    return fail(PPS_E_INVALID,
                "the local L-BFGS phase minimises the objective alone; it has no constraint "
                "handling (the swarm has %d constraints)", s->P.n_cons);
  if (lp->history < 1 || lp->history > 16) return fail(PPS_E_INVALID, "history must be 1..16");
  if (lp->maxiters < 0 || lp->max_backtracks < 0) return fail(PPS_E_INVALID, "negative iteration count");
  if (!(lp->reltol >= 0.0f && std::isfinite(lp->reltol)))
    return fail(PPS_E_INVALID, "reltol must be finite and >= 0");
  if (lp->linesearch == PPS_LINESEARCH_ARMIJO) {
    if (!(lp->armijo_rho > 0.0f && lp->armijo_rho < 1.0f))
      return fail(PPS_E_INVALID, "armijo_rho must be in (0,1)");
  } else if (lp->linesearch == PPS_LINESEARCH_STRONG_WOLFE) {
    if (!(lp->armijo_c > 0.0f && lp->armijo_c < lp->wolfe_c2 && lp->wolfe_c2 < 1.0f))
      return fail(PPS_E_INVALID, "Strong-Wolfe constants need 0 < c1 < c2 < 1 (c1 %g, c2 %g)",
                  (double)lp->armijo_c, (double)lp->wolfe_c2);
    if (!(lp->alpha_init > 0.0f && lp->alpha_max >= lp->alpha_init && std::isfinite(lp->alpha_max)))
      return fail(PPS_E_INVALID, "Strong-Wolfe steps need 0 < alpha_init <= alpha_max < Inf");
    if (lp->ls_maxiters < 1 || lp->zoom_maxiters < 0)
      return fail(PPS_E_INVALID, "Strong-Wolfe iteration limits: ls_maxiters >= 1, zoom_maxiters >= 0");
  } else {
    return fail(PPS_E_INVALID, "unknown line search %d", lp->linesearch);
  }
  return PPS_OK;
}

static LbfgsArgs lbfgs_args(const pps_lbfgs_params* lp) {
  return LbfgsArgs{lp->maxiters,   lp->history,    lp->abstol,      lp->armijo_c,
                   lp->armijo_rho, lp->max_backtracks, lp->linesearch, lp->wolfe_c2,
                   lp->alpha_init, lp->alpha_max,  lp->ls_maxiters, lp->zoom_maxiters,
                   lp->reltol};
}

extern "C" {

int pps_swarm_lbfgs(pps_swarm* s, const pps_lbfgs_params* lp, float* out_x, float* out_f,
                    float* best_x, float* best_f, int64_t* best_idx) {
  const int chk = lbfgs_check(s, lp);
  if (chk != PPS_OK) return chk;
  HIPCHK(hipSetDevice(s->ctx->device));
  const int d = s->P.d;
  const int64_t n = s->prm.n, npad = s->S.npad;
  hipStream_t st = s->ctx->stream;
  if (!s->lx) {
    const size_t bx = align256(sizeof(float) * d * npad), bf = align256(sizeof(float) * npad);
    void* m = nullptr;
    hipError_t e = hipMalloc(&m, bx + bf + 256);
    if (e != hipSuccess) return fail(PPS_E_NOMEM, "hipMalloc of local-search buffers failed");
    s->lx = (float*)m;
    s->lf = (float*)((char*)m + bx);
    s->lkey = (uint64_t*)((char*)m + bx + bf);
  }
  const LbfgsArgs A = lbfgs_args(lp);
  HIPCHK(hipEventRecord(s->ev0, st));
  HIPCHK(launch_lbfgs(s->S, s->P, A, s->lx, s->lf, st));
  HIPCHK(launch_argmin_f32(s->lf, n, s->lkey, st));
  HIPCHK(hipEventRecord(s->ev1, st));
  HIPCHK(hipStreamSynchronize(st));
  HIPCHK(hipEventElapsedTime(&s->last_ms, s->ev0, s->ev1));
  uint64_t key = 0;
  HIPCHK(hipMemcpy(&key, s->lkey, sizeof key, hipMemcpyDeviceToHost));
  const int64_t bi = (int64_t)(uint32_t)key;
  if (best_idx) *best_idx = bi;
  if (best_f) HIPCHK(hipMemcpy(best_f, s->lf + bi, sizeof(float), hipMemcpyDeviceToHost));
  if (best_x)
    HIPCHK(hipMemcpy2D(best_x, sizeof(float), s->lx + bi, sizeof(float) * npad, sizeof(float), d,
                       hipMemcpyDeviceToHost));
  if (out_f) HIPCHK(hipMemcpy(out_f, s->lf, sizeof(float) * n, hipMemcpyDeviceToHost));
  if (out_x) {
    std::vector<float> tmp((size_t)d * n);
    HIPCHK(hipMemcpy2D(tmp.data(), sizeof(float) * n, s->lx, sizeof(float) * npad,
                       sizeof(float) * n, d, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < n; ++i)
      for (int j = 0; j < d; ++j) out_x[i * d + j] = tmp[(size_t)j * n + i];
  }
  return PPS_OK;
}

// The kernels of pps_swarm_lbfgs with their outputs in this device's
// `scratch` instead of library-allocated buffers: SoA results (d x npad),
// objective values (npad), the argmin key, and the AoS staging of out_x.
// Nothing there needs clearing: every word read is written by this call.
int pps_swarm_lbfgs_noalloc(pps_swarm* s, const pps_lbfgs_params* lp, float* out_x,
                            float* out_f, float* best_x, float* best_f, int64_t* best_idx) {
  const int chk = lbfgs_check(s, lp);
  if (chk != PPS_OK) return chk;
  const int d = s->P.d;
  const int64_t n = s->prm.n, npad = s->S.npad;
  const size_t o_f = align256(sizeof(float) * d * npad);
  const size_t o_key = o_f + align256(sizeof(float) * npad);
  const size_t o_aos = o_key + 256;
  const size_t need = o_aos + (out_x ? align256(sizeof(float) * d * n) : 0);
  if (need > (size_t)PPS_NOALLOC_SCRATCH_BYTES)
    return fail(PPS_E_INVALID,
                "the allocation-free local search needs %zu bytes of scratch for n=%lld, d=%d; "
                "it has %zu (about 4 n (2d + 1) bytes fit)",
                need, (long long)n, d, (size_t)PPS_NOALLOC_SCRATCH_BYTES);
  char* base = scratch_of(s->ctx->device);
  if (!base) return fail(PPS_E_STATE, "no scratch for device %d", s->ctx->device);
  float* lx = (float*)base;
  float* lf = (float*)(base + o_f);
  uint64_t* lkey = (uint64_t*)(base + o_key);
  float* aos = (float*)(base + o_aos);
  HIPCHK(hipSetDevice(s->ctx->device));
  hipStream_t st = s->ctx->stream;
  const LbfgsArgs A = lbfgs_args(lp);
  HIPCHK(hipEventRecord(s->ev0, st));
  HIPCHK(launch_lbfgs(s->S, s->P, A, lx, lf, st));
  HIPCHK(launch_argmin_f32(lf, n, lkey, st));
  HIPCHK(hipEventRecord(s->ev1, st));
  if (out_x) HIPCHK(launch_soa_to_aos(lx, npad, n, d, aos, st));
  HIPCHK(hipStreamSynchronize(st));
  HIPCHK(hipEventElapsedTime(&s->last_ms, s->ev0, s->ev1));
  uint64_t key = 0;
  HIPCHK(hipMemcpy(&key, lkey, sizeof key, hipMemcpyDeviceToHost));
  const int64_t bi = (int64_t)(uint32_t)key;
  if (best_idx) *best_idx = bi;
  if (best_f) HIPCHK(hipMemcpy(best_f, lf + bi, sizeof(float), hipMemcpyDeviceToHost));
  if (best_x)
    HIPCHK(hipMemcpy2D(best_x, sizeof(float), lx + bi, sizeof(float) * npad, sizeof(float), d,
                       hipMemcpyDeviceToHost));
  if (out_f) HIPCHK(hipMemcpy(out_f, lf, sizeof(float) * n, hipMemcpyDeviceToHost));
  if (out_x) HIPCHK(hipMemcpy(out_x, aos, sizeof(float) * d * n, hipMemcpyDeviceToHost));
  return PPS_OK;
}

}  // extern "C"
