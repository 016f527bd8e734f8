// pps_serve.hip — continuous-batching request server for independent sync-PSO
// problems (no reference counterpart: the reference solves one problem per
// call, src/solve.jl:61-67).
//
// Requests are queued by pps_server_submit from any thread.  A server thread
// repeats:
//   1. retire problems whose workgroup has published done = 1 into pinned
//      host memory (their result is read there and their HBM freed);
//   2. admit queued requests in arrival order while a descriptor slot is free
//      and the problem's SoA state fits in the HBM arena (first fit);
//   3. launch one time slice over every resident problem (work items, one
//      workgroup each): a small problem's workgroup runs its iteration loop
//      until the problem ends or the slice's device-clock budget is spent; a
//      large problem runs one iteration per slice split over several
//      workgroups, then a final item publishes its result.  Requests that
//      arrive during a slice start in the next one while the others keep
//      running; two slices are in flight at a time.
// A problem's random stream is keyed by its own seed and particle indices, so
// its result does not depend on what it was packed with.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "pps_batch.h"
#include "pps_kernels.h"
#include "pps_runtime.h"

using namespace pps;

namespace {

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Swarms with n * d at or above this run as `batch_parts(n, .)` workgroups, one
// iteration per slice; smaller ones run their own loop in one workgroup.  A
// slice lasts at least the longest single iteration of a SMALL item (about
// n * d / 2048 dependent rounds of loads in one workgroup), and one slice is
// one iteration of every large problem: at 32768 that was ~16 rounds and an
// 80 us slice whatever the budget below; at 8192, with PART items of 256
// particles, a slice follows the budget.
constexpr int64_t kLargeWork = 8192;
// Device-clock budget of SMALL items while large problems are resident, so a
// slice (and with it one iteration of every large problem) stays short.  The
// slice lasts as long as its longest item: a SMALL item's budget, or a large
// problem's 32-coordinate PART tile (~14 us).  On the 4096-request mix at half
// the saturating rate, served p50/p99: 10.2/23.2 ms at 20 us, 9.6/21.5 at
// 15 us, 10.7/21.7 at 10 us, 11.3/21.7 at 5 us, with the same throughput.
constexpr int kSliceUsWithLarge = 15;
// Tiles per PART item of a large problem.  A PART item spends a few dependent
// round trips (its descriptor, the part keys, the winner's position, and at
// the end its own winner's row) around the tiles it streams; two tiles per
// item amortise them (batched mix 5488 -> 6003 solves/s) but lengthen the
// slice that paces every large problem (served p99 28.8 -> 44.8 ms; four
// tiles: 5883 solves/s, p99 72.9 ms).  So a problem admitted while at least
// kBacklog requests wait behind it (throughput) gets two-tile parts, and one
// admitted without such a backlog (latency) one-tile parts.
constexpr size_t kBacklog = 64;

enum { R_PENDING = 0, R_RUNNING = 1, R_DONE = 2, R_CANCELLED = 3, R_FAILED = 4 };

struct Req {
  uint64_t ticket = 0;
  ProbArgs P{};
  std::vector<float> lb, ub, par, shift;
  int64_t n = 0;
  uint64_t seed = 0;
  float c1 = 1.4962f, c2 = 1.4962f;
  int32_t maxiters = 0;
  float w = 0.7298f, wdamp = 1.0f;
  int state = R_PENDING;
  int slot = -1;
  size_t off = 0, bytes = 0;
  // large problems: parts per iteration, next iteration to issue (0 = init),
  // inertia of that iteration (iterated FP32 product, as pps_swarm_run)
  bool large = false;
  int parts = 0;
  uint32_t t_next = 0;
  float w_next = 0.0f;
  pps_result res{};
};

// First-fit allocator over [0, size) in 256-byte units.
struct Arena {
  std::map<size_t, size_t> free_;  // offset -> length
  size_t size = 0, used = 0, peak = 0;
  void reset(size_t s) {
    size = s;
    used = peak = 0;
    free_.clear();
    if (s) free_[0] = s;
  }
  bool alloc(size_t n, size_t* off) {
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= n) {
        *off = it->first;
        const size_t rest = it->second - n;
        const size_t o = it->first;
        free_.erase(it);
        if (rest) free_[o + n] = rest;
        used += n;
        peak = std::max(peak, used);
        return true;
      }
    }
    return false;
  }
  void release(size_t off, size_t n) {
    used -= n;
    auto it = free_.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
};

}  // namespace

struct pps_server {
  pps_ctx* ctx = nullptr;
  pps_server_config cfg{};
  char* arena_dev = nullptr;
  Arena arena;
  BatchProb* d_desc = nullptr;  // [max_active] device
  BatchProb* h_desc = nullptr;  // [max_active] pinned staging
  BatchOut* h_out = nullptr;    // [max_active] pinned results
  BatchItem* h_items[2] = {};    // [max_items] pinned work lists (two slices in flight)
  int64_t max_items = 0;
  hipEvent_t slice_done[2] = {};
  uint64_t slices_enqueued = 0, slices_completed = 0;
  // retired slots/arena ranges wait until every slice that may still list
  // them has completed
  struct Quarantined {
    uint64_t after;  // free once slices_completed >= after
    int slot;
    size_t off, bytes;
  };
  std::deque<Quarantined> quarantine;
  BatchXchg* d_xchg = nullptr;   // [max_active] device exchange areas
  std::vector<int> free_slots;
  std::vector<Req*> slot_req;
  std::vector<char> slot_fresh;  // SMALL slot whose descriptor the next slice uploads
  hipStream_t stream = nullptr;
  uint64_t budget_ticks = 0, budget_ticks_large = 0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Req*> pending;
  std::unordered_map<uint64_t, Req*> reqs;
  uint64_t next_ticket = 1;
  int running = 0;
  int64_t submitted = 0, completed = 0, launches = 0;
  bool stop = false;
  std::string fatal;
  std::thread th;

  size_t bytes_for(int d, int64_t n) const {
    const int64_t npad = (n + 63) / 64 * 64;
    return 3 * align256(sizeof(float) * d * npad) + align256(sizeof(float) * npad);
  }

  bool retire_locked();
  void loop();
};

// Retire every running problem whose done flag is set (caller holds mu).
bool pps_server::retire_locked() {
  bool any = false;
  for (size_t s = 0; s < slot_req.size(); ++s) {
    Req* r = slot_req[s];
    if (!r) continue;
    const uint32_t done = __atomic_load_n(&h_out[s].done, __ATOMIC_ACQUIRE);
    if (!done) continue;
    const BatchOut& o = h_out[s];
    r->res.gbest_cost = o.gb_cost;
    r->res.gbest_idx = o.gb_idx;
    r->res.d = r->P.d;
    for (int j = 0; j < r->P.d; ++j) r->res.gbest_pos[j] = o.gb[j];
    r->res.t_done_s = now_s();
    // FP64 re-evaluation of the final minimum
    pps_problem pp{};
    pp.objective = r->P.obj;
    pp.d = r->P.d;
    pp.lb = r->lb.data();
    pp.ub = r->ub.data();
    pp.params = r->par.data();
    pp.n_params = (int32_t)r->par.size();
    pp.shift = r->shift.empty() ? nullptr : r->shift.data();
    double xd[PPS_BMAXD];
    for (int j = 0; j < r->P.d; ++j) xd[j] = (double)o.gb[j];
    double f64 = 0.0;
    pps_eval_f64(&pp, 1, xd, &f64);
    r->res.gbest_cost_f64 = f64;
    r->state = R_DONE;
    quarantine.push_back({slices_enqueued, (int)s, r->off, r->bytes});
    slot_req[s] = nullptr;
    --running;
    ++completed;
    any = true;
  }
  return any;
}

void pps_server::loop() {
  (void)hipSetDevice(ctx->device);
  std::vector<BatchItem> items;
  std::vector<int32_t> slots_in_slice;
  while (true) {
    items.clear();
    slots_in_slice.clear();
    bool any_large = false;
    {
      std::unique_lock<std::mutex> lk(mu);
      if (retire_locked()) cv_done.notify_all();
      while (!quarantine.empty() && quarantine.front().after <= slices_completed) {
        const Quarantined& qd = quarantine.front();
        arena.release(qd.off, qd.bytes);
        free_slots.push_back(qd.slot);
        quarantine.pop_front();
      }
      if (stop && pending.empty() && running == 0) break;
      while (!pending.empty() && !free_slots.empty()) {
        Req* r = pending.front();
        size_t off = 0;
        if (!arena.alloc(r->bytes, &off)) break;
        pending.pop_front();
        const int s = free_slots.back();
        free_slots.pop_back();
        r->slot = s;
        r->off = off;
        r->state = R_RUNNING;
        r->res.t_admit_s = now_s();
        // clear the previous occupant's flag before the slot can be polled
        memset(&h_out[s], 0, sizeof(BatchOut));
        BatchProb& B = h_desc[s];
        memset(&B, 0, sizeof B);
        B.P = r->P;
        B.n = r->n;
        B.npad = (r->n + 63) / 64 * 64;
        B.key0 = (uint32_t)r->seed;
        B.key1 = (uint32_t)(r->seed >> 32);
        B.remaining = r->maxiters;
        B.w = r->w;
        B.wdamp = r->wdamp;
        B.c1 = r->c1;
        B.c2 = r->c2;
        const size_t row = align256(sizeof(float) * r->P.d * B.npad);
        char* base = arena_dev + r->off;
        B.x = (float*)base;
        B.v = (float*)(base + row);
        B.pb = (float*)(base + 2 * row);
        B.pc = (float*)(base + 3 * row);
        r->large = r->n * r->P.d >= kLargeWork;
        if (r->large) {
          // parts read the device copy; upload it ahead of the slice
          r->parts = batch_parts(r->n, pending.size() >= kBacklog ? 2 : 1);
          r->t_next = 0;
          r->w_next = r->w;
          (void)hipMemcpyAsync(d_desc + s, &B, sizeof(BatchProb), hipMemcpyHostToDevice, stream);
        } else {
          slot_fresh[s] = 1;
        }
        slot_req[s] = r;
        ++running;
      }
      // A slice's work list holds at most max_items; a problem that does not
      // fit whole (all parts of a large problem's iteration) waits for the
      // next slice with its state untouched.
      const size_t cap = (size_t)max_items;
      for (size_t s = 0; s < slot_req.size(); ++s) {
        Req* r = slot_req[s];
        if (!r) continue;
        if (!r->large) {
          if (!slot_fresh[s] && __atomic_load_n(&h_out[s].done, __ATOMIC_ACQUIRE)) continue;
          if (items.size() + 1 > cap) continue;
          BatchItem it{};
          it.slot = (int32_t)s;
          it.kind = BATCH_SMALL;
          it.fresh = slot_fresh[s];
          slot_fresh[s] = 0;
          items.push_back(it);
          continue;
        }
        any_large = true;
        if (r->t_next > (uint32_t)r->maxiters) {
          if (r->t_next == (uint32_t)r->maxiters + 1 && items.size() + 1 <= cap) {
            BatchItem it{};
            it.slot = (int32_t)s;
            it.kind = BATCH_FINAL;
            it.nparts = r->parts;
            it.t = (uint32_t)r->maxiters;
            items.push_back(it);
            r->t_next++;  // FINAL issued; retired when its done flag appears
          }
          continue;
        }
        if (items.size() + (size_t)r->parts > cap) continue;
        for (int q = 0; q < r->parts; ++q) {
          BatchItem it{};
          it.slot = (int32_t)s;
          it.kind = BATCH_PART;
          it.part = q;
          it.nparts = r->parts;
          it.t = r->t_next;
          it.w = r->w_next;
          items.push_back(it);
        }
        if (r->t_next > 0) r->w_next = r->w_next * r->wdamp;
        r->t_next++;
      }
      if (items.empty()) {
        if (slices_completed < slices_enqueued) {
          // nothing new to issue: let the last slice finish (frees quarantine)
          lk.unlock();
          (void)hipStreamSynchronize(stream);
          lk.lock();
          slices_completed = slices_enqueued;
          continue;
        }
        cv_work.wait_for(lk, std::chrono::microseconds(running > 0 ? 30 : (stop ? 1000 : 5000)));
        continue;
      }
    }
    // Two slices in flight: work list i of slice n is list n & 1, rewritten
    // only after slice n - 2 (its previous user) has completed.
    hipError_t e = hipSuccess;
    const int buf = (int)(slices_enqueued & 1u);
    if (slices_enqueued >= 2 && slices_completed + 2 <= slices_enqueued) {
      e = hipEventSynchronize(slice_done[buf]);
      if (e == hipSuccess) slices_completed = slices_enqueued - 1;
    }
    if (e == hipSuccess) {
      memcpy(h_items[buf], items.data(), sizeof(BatchItem) * items.size());
      e = launch_slice(d_desc, h_desc, h_items[buf], (int32_t)items.size(),
                       any_large ? budget_ticks_large : budget_ticks, h_out, d_xchg, stream);
    }
    if (e == hipSuccess) e = hipEventRecord(slice_done[buf], stream);
    std::unique_lock<std::mutex> lk(mu);
    if (e == hipSuccess) {
      ++slices_enqueued;
      ++launches;
      // wait for the older slice so the next work list (and the flags of the
      // problems it finished) are available
      if (slices_enqueued >= 2) {
        const int older = (int)((slices_enqueued - 2) & 1u);
        lk.unlock();
        e = hipEventSynchronize(slice_done[older]);
        lk.lock();
        if (e == hipSuccess) slices_completed = slices_enqueued - 1;
      }
    }
    if (e != hipSuccess) {
      fatal = hipGetErrorString(e);
      (void)hipStreamSynchronize(stream);
      slices_completed = slices_enqueued;
      for (size_t s = 0; s < slot_req.size(); ++s) {
        Req* r = slot_req[s];
        if (!r) continue;
        r->state = R_FAILED;
        arena.release(r->off, r->bytes);
        slot_req[s] = nullptr;
        free_slots.push_back((int)s);
        --running;
      }
      cv_done.notify_all();
    }
  }
  (void)hipStreamSynchronize(stream);
}

#define SVCHK(expr)                                                        \
  do {                                                                     \
    hipError_t e_ = (expr);                                                \
    if (e_ != hipSuccess) {                                                \
      int rc_ = hip_fail(e_, #expr, __FILE__, __LINE__);                   \
      destroy_partial(sv);                                                 \
      return rc_;                                                          \
    }                                                                      \
  } while (0)

static void destroy_partial(pps_server* sv) {
  if (!sv) return;
  if (sv->stream) {
    (void)hipStreamSynchronize(sv->stream);
    (void)hipStreamDestroy(sv->stream);
  }
  if (sv->arena_dev) (void)hipFree(sv->arena_dev);
  if (sv->d_desc) (void)hipFree(sv->d_desc);
  if (sv->h_desc) (void)hipHostFree(sv->h_desc);
  if (sv->h_out) (void)hipHostFree(sv->h_out);
  for (int b = 0; b < 2; ++b) {
    if (sv->h_items[b]) (void)hipHostFree(sv->h_items[b]);
    if (sv->slice_done[b]) (void)hipEventDestroy(sv->slice_done[b]);
  }
  if (sv->d_xchg) (void)hipFree(sv->d_xchg);
  for (auto& kv : sv->reqs) delete kv.second;
  delete sv;
}

extern "C" {

int pps_server_create(pps_ctx* ctx, const pps_server_config* cfg, pps_server** out) {
  if (!ctx || !out) return fail(PPS_E_INVALID, "NULL argument");
  *out = nullptr;
  pps_server* sv = new pps_server();
  sv->ctx = ctx;
  if (cfg) sv->cfg = *cfg;
  SVCHK(hipSetDevice(ctx->device));
  int32_t max_active = sv->cfg.max_active > 0 ? sv->cfg.max_active : 4096;
  if (max_active > (1 << 16)) max_active = 1 << 16;
  sv->cfg.max_active = max_active;
  size_t budget = (size_t)std::max<int64_t>(0, sv->cfg.hbm_budget);
  if (budget == 0) {
    size_t fr = 0, tot = 0;
    SVCHK(hipMemGetInfo(&fr, &tot));
    budget = std::min<size_t>(fr / 10 * 8, (size_t)16 << 30);
  }
  budget = budget / 256 * 256;
  sv->cfg.hbm_budget = (int64_t)budget;
  SVCHK(hipMalloc(&sv->arena_dev, budget));
  sv->arena.reset(budget);
  SVCHK(hipMalloc(&sv->d_desc, sizeof(BatchProb) * max_active));
  SVCHK(hipHostMalloc(&sv->h_desc, sizeof(BatchProb) * max_active, hipHostMallocCoherent | hipHostMallocMapped));
  SVCHK(hipHostMalloc(&sv->h_out, sizeof(BatchOut) * max_active, hipHostMallocCoherent | hipHostMallocMapped));
  sv->max_items = 1 << 20;  // workgroups per launch
  for (int b = 0; b < 2; ++b) {
    SVCHK(hipHostMalloc(&sv->h_items[b], sizeof(BatchItem) * sv->max_items,
                        hipHostMallocCoherent | hipHostMallocMapped));
    // Without the event's system-scope fence (a cache writeback and
    // invalidation after every slice): the host waits on these events only
    // for completion order, to rewrite a slice's work list, and every result
    // and done flag it reads is published by the kernel itself with a
    // system-scope release.
    SVCHK(hipEventCreateWithFlags(&sv->slice_done[b],
                                  hipEventDisableTiming | hipEventDisableSystemFence));
  }
  SVCHK(hipMalloc(&sv->d_xchg, sizeof(BatchXchg) * max_active));
  memset(sv->h_out, 0, sizeof(BatchOut) * max_active);
  sv->slot_req.assign(max_active, nullptr);
  sv->slot_fresh.assign(max_active, 0);
  for (int s = max_active - 1; s >= 0; --s) sv->free_slots.push_back(s);
  SVCHK(hipStreamCreateWithFlags(&sv->stream, hipStreamNonBlocking));
  int wclk_khz = 0;
  SVCHK(hipDeviceGetAttribute(&wclk_khz, hipDeviceAttributeWallClockRate, ctx->device));
  const int slice_us = sv->cfg.slice_us > 0 ? sv->cfg.slice_us : 1000;
  sv->cfg.slice_us = slice_us;
  sv->budget_ticks = (uint64_t)wclk_khz * (uint64_t)slice_us / 1000u;
  sv->budget_ticks_large =
      (uint64_t)wclk_khz * (uint64_t)std::min(slice_us, kSliceUsWithLarge) / 1000u;
  sv->th = std::thread([sv] { sv->loop(); });
  *out = sv;
  return PPS_OK;
}

int pps_server_submit(pps_server* sv, const pps_problem* prob, const pps_swarm_params* prm,
                      int32_t maxiters, float w, float wdamp, uint64_t* ticket) {
  if (!sv || !prob || !prm || !ticket) return fail(PPS_E_INVALID, "NULL argument");
  if (!batch_supported(prob->objective, prob->d))
    return fail(PPS_E_INVALID,
                "served problems are Rosenbrock or Rastrigin with d in 1..32 (got objective %d, "
                "d=%d)", prob->objective, prob->d);
  if (prob->n_cons != 0)
    return fail(PPS_E_INVALID, "served problems are unconstrained (n_cons = %d)", prob->n_cons);
  if (prm->sampler != PPS_SAMPLER_UNIFORM)
    return fail(PPS_E_INVALID, "served problems use the uniform initialiser");
  if (prm->mode != PPS_MODE_SYNC || prm->gbest_from != PPS_GBEST_PBEST ||
      prm->clamp != PPS_CLAMP_COMPONENT)
    return fail(PPS_E_INVALID, "served problems use sync PSO with best-position global best");
  if (prm->n < 1 || prm->n > (1 << 24)) return fail(PPS_E_INVALID, "num_particles outside 1..2^24");
  if (maxiters < 0) return fail(PPS_E_INVALID, "maxiters < 0");
  if (check_step_coefs(prm->c1, prm->c2) != PPS_OK) return PPS_E_INVALID;
  Req* r = new Req();
  int rc = fill_prob_args(prob, &r->P);
  if (rc != PPS_OK) {
    delete r;
    return rc;
  }
  r->P.clamp = PPS_CLAMP_COMPONENT;
  if (std::isinf(r->P.lb[0])) {
    delete r;
    return fail(PPS_E_INVALID, "served problems need finite bounds");
  }
  r->lb.assign(prob->lb, prob->lb + prob->d);
  r->ub.assign(prob->ub, prob->ub + prob->d);
  if (prob->n_params > 0) r->par.assign(prob->params, prob->params + prob->n_params);
  if (prob->shift) r->shift.assign(prob->shift, prob->shift + prob->d);
  r->n = prm->n;
  r->seed = prm->seed;
  r->c1 = prm->c1;
  r->c2 = prm->c2;
  r->maxiters = maxiters;
  r->w = w;
  r->wdamp = wdamp;
  r->bytes = sv->bytes_for(prob->d, prm->n);
  if (r->bytes > (size_t)sv->cfg.hbm_budget) {
    delete r;
    return fail(PPS_E_NOMEM, "problem state (%zu bytes) exceeds the server's HBM budget",
                sv->bytes_for(prob->d, prm->n));
  }
  r->res.t_submit_s = now_s();
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    if (sv->stop) {
      delete r;
      return fail(PPS_E_STATE, "server is shutting down");
    }
    r->ticket = sv->next_ticket++;
    sv->reqs[r->ticket] = r;
    sv->pending.push_back(r);
    ++sv->submitted;
    *ticket = r->ticket;
  }
  sv->cv_work.notify_one();
  return PPS_OK;
}

// Final status of a request (caller holds mu).  A finished request (done,
// cancelled or failed) is returned once and then forgotten: the server
// thread no longer refers to it, so a long-running server holds only the
// requests that are unfinished or not yet collected.
static int collect_locked(pps_server* sv, std::unordered_map<uint64_t, Req*>::iterator it,
                          pps_result* out) {
  Req* r = it->second;
  int rc;
  switch (r->state) {
    case R_DONE:
      if (out) *out = r->res;
      rc = PPS_OK;
      break;
    case R_CANCELLED: rc = fail(PPS_E_CANCELLED, "request cancelled"); break;
    case R_FAILED: rc = fail(PPS_E_HIP, "request failed: %s", sv->fatal.c_str()); break;
    default: return PPS_E_PENDING;
  }
  sv->reqs.erase(it);
  delete r;
  return rc;
}

int pps_server_poll(pps_server* sv, uint64_t ticket, pps_result* out) {
  if (!sv) return fail(PPS_E_INVALID, "server is NULL");
  std::lock_guard<std::mutex> lk(sv->mu);
  auto it = sv->reqs.find(ticket);
  if (it == sv->reqs.end()) return fail(PPS_E_NOTFOUND, "unknown ticket %llu", (unsigned long long)ticket);
  return collect_locked(sv, it, out);
}

int pps_server_wait(pps_server* sv, uint64_t ticket, double timeout_s, pps_result* out) {
  if (!sv) return fail(PPS_E_INVALID, "server is NULL");
  std::unique_lock<std::mutex> lk(sv->mu);
  const auto deadline = std::chrono::steady_clock::now() +
                        std::chrono::duration<double>(timeout_s > 0 ? timeout_s : 1e9);
  while (true) {
    // looked up again after every wake-up: another thread may have collected it
    auto it = sv->reqs.find(ticket);
    if (it == sv->reqs.end())
      return fail(PPS_E_NOTFOUND, "unknown ticket %llu", (unsigned long long)ticket);
    const int rc = collect_locked(sv, it, out);
    if (rc != PPS_E_PENDING) return rc;
    if (sv->cv_done.wait_until(lk, deadline) == std::cv_status::timeout) {
      it = sv->reqs.find(ticket);
      if (it == sv->reqs.end())
        return fail(PPS_E_NOTFOUND, "unknown ticket %llu", (unsigned long long)ticket);
      const int rc2 = collect_locked(sv, it, out);
      return rc2 == PPS_E_PENDING ? fail(PPS_E_TIMEOUT, "wait timed out") : rc2;
    }
  }
}

int pps_server_cancel(pps_server* sv, uint64_t ticket) {
  if (!sv) return fail(PPS_E_INVALID, "server is NULL");
  std::lock_guard<std::mutex> lk(sv->mu);
  auto it = sv->reqs.find(ticket);
  if (it == sv->reqs.end()) return fail(PPS_E_NOTFOUND, "unknown ticket");
  Req* r = it->second;
  if (r->state != R_PENDING) return fail(PPS_E_STATE, "only queued requests can be cancelled");
  sv->pending.erase(std::find(sv->pending.begin(), sv->pending.end(), r));
  r->state = R_CANCELLED;
  sv->cv_done.notify_all();
  return PPS_OK;
}

int pps_server_stats(pps_server* sv, int64_t* submitted, int64_t* completed, int64_t* launches,
                     int64_t* peak_bytes) {
  if (!sv) return fail(PPS_E_INVALID, "server is NULL");
  std::lock_guard<std::mutex> lk(sv->mu);
  if (submitted) *submitted = sv->submitted;
  if (completed) *completed = sv->completed;
  if (launches) *launches = sv->launches;
  if (peak_bytes) *peak_bytes = (int64_t)sv->arena.peak;
  return PPS_OK;
}

int pps_server_destroy(pps_server* sv) {
  if (!sv) return PPS_OK;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->stop = true;
    for (Req* r : sv->pending) r->state = R_CANCELLED;
    sv->pending.clear();
  }
  sv->cv_work.notify_all();
  if (sv->th.joinable()) sv->th.join();
  (void)hipSetDevice(sv->ctx->device);
  destroy_partial(sv);
  return PPS_OK;
}

int pps_batch_solve(pps_ctx* ctx, int32_t count, const pps_problem* probs,
                    const pps_swarm_params* params, const int32_t* maxiters, const float* w,
                    const float* wdamp, pps_result* out, float* elapsed_ms) {
  if (!ctx || count < 0 || (count > 0 && (!probs || !params || !maxiters || !w || !wdamp || !out)))
    return fail(PPS_E_INVALID, "NULL argument");
  pps_server_config cfg{};
  cfg.max_active = std::max(1, std::min(count, 4096));
  pps_server* sv = nullptr;
  int rc = pps_server_create(ctx, &cfg, &sv);
  if (rc != PPS_OK) return rc;
  std::vector<uint64_t> tickets(count);
  const double t0 = now_s();
  for (int32_t i = 0; i < count && rc == PPS_OK; ++i)
    rc = pps_server_submit(sv, &probs[i], &params[i], maxiters[i], w[i], wdamp[i], &tickets[i]);
  for (int32_t i = 0; i < count && rc == PPS_OK; ++i) rc = pps_server_wait(sv, tickets[i], 0, &out[i]);
  const double t1 = now_s();
  std::string keep = g_err;
  pps_server_destroy(sv);
  g_err = keep;
  if (elapsed_ms) *elapsed_ms = (float)((t1 - t0) * 1e3);
  return rc;
}

}  // extern "C"
