// pps_batch.hip — packed launches of many independent sync-PSO problems.
//
// A launch ("slice") is a list of work items, one workgroup each:
//   SMALL  one problem in one workgroup: its iteration loop runs until the
//          problem ends or one more iteration as long as the last would end
//          past the slice's device-clock budget (checked after every
//          iteration by thread 0, broadcast through LDS); the winner's
//          position is handed over through LDS.
//   PART   part q of k of a large problem: one iteration t over the particle
//          tiles q, q+k, q+2k, ... (256 particles per tile, one per thread,
//          sixteen coordinates per round of loads).  The global best
//          of iteration t-1 is the minimum of the k part keys the previous
//          slice wrote to parity (t-1)&1 of the problem's exchange area; the
//          part writes its own key and candidate position to parity t&1, so
//          no workgroup reads what another writes within one launch.
//   FINAL  reduce the k part keys of a large problem's last iteration into
//          its result.
// Results and done flags go to host-visible pinned memory.  A problem's random
// stream is keyed by its own seed, particle indices and iteration numbers, and
// every path below applies the same per-particle arithmetic as the plain sync
// kernel, so a result is the same alone (pps_swarm_init + pps_swarm_run) or
// packed with anything, in any slicing.
#include <hip/hip_runtime.h>

#include "pps_batch.h"
#include "pps_device.h"

namespace pps {

namespace {

constexpr uint64_t KEY_NONE = ~0ull;
constexpr int PPT = 4;  // particles in flight per thread (SMALL items)
constexpr int TILE = BATCH_BLOCK * PPT;
constexpr int PTILE = BATCH_BLOCK;  // particles per tile of a PART item
constexpr int PCOORD = 16;          // coordinates per round of loads (PART)

template <int BLOCK>
__device__ __forceinline__ uint64_t block_min(uint64_t k, uint64_t* s_red) {
  k = wave_min_u64(k);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) s_red[wid] = k;
  __syncthreads();
  uint64_t r = s_red[0];
#pragma unroll
  for (int i = 1; i < BLOCK / 64; ++i) r = umin64(r, s_red[i]);
  __syncthreads();
  return r;
}

// Per-problem constants staged in LDS (runtime-indexed by coordinate).
struct ProbLds {
  float lb[PPS_BMAXD], ub[PPS_BMAXD], shift[PPS_BMAXD], gb[PPS_BMAXD];
};

// Objective accumulated coordinate by coordinate, in the order of
// objective<>() (pps_device.h): Rastrigin adds z_j^2 - 10 cos(2 pi z_j) to
// 10 d for j = 0, 1, ...; Rosenbrock adds term(z_{j-1}, z_j) for j = 1, 2, ...
template <int OBJ>
struct Accum {
  float res, zprev;
  __device__ __forceinline__ void start(int d) {
    res = OBJ == PPS_OBJ_RASTRIGIN ? 10.0f * (float)d : 0.0f;
    zprev = 0.0f;
  }
  __device__ __forceinline__ void add(int j, float z, float p1, float p2) {
    if constexpr (OBJ == PPS_OBJ_RASTRIGIN) {
      res = res + (z * z - 10.0f * cos2pi(z));
    } else {
      if (j > 0) {
        const float dd = z - zprev * zprev;
        const float e = p1 - zprev;
        res = res + (p2 * (dd * dd) + e * e);
      }
      zprev = z;
    }
  }
};

struct Swarm {
  int d;
  int64_t n, np;
  uint32_t k0, k1;
  float p1, p2, c1, c2;
  bool has_shift;
  float* __restrict__ X;
  float* __restrict__ V;
  float* __restrict__ PB;
  float* __restrict__ PC;
};

// Coordinate j of particle i of a [d][np] state array at byte offset
// i * 4 + j * np * 4: every state array of a batched problem is below 2 GiB
// (n <= 2^24, d <= 32), so the offset fits 32 bits and the array base stays
// in scalar registers.
__device__ __forceinline__ float& at32(float* base, uint32_t off) {
  return *reinterpret_cast<float*>(reinterpret_cast<char*>(base) + off);
}

// Copies coordinates j < d of one particle from the rows of src to those of
// dst (byte offsets i4 + j * row4) in chunks of CH coordinates, each chunk's
// loads issued before its stores: ceil(d / CH) memory round trips instead of
// d (a load after a store that may alias it waits for the store).
template <int CH>
__device__ __forceinline__ void copy_particle_rows(float* dst, float* src, uint32_t i4,
                                                   uint32_t row4, int d) {
  for (int j0 = 0; j0 < d; j0 += CH) {
    float t[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c)
      if (j0 + c < d) t[c] = at32(src, i4 + (uint32_t)(j0 + c) * row4);
#pragma unroll
    for (int c = 0; c < CH; ++c)
      if (j0 + c < d) at32(dst, i4 + (uint32_t)(j0 + c) * row4) = t[c];
  }
}

__device__ __forceinline__ Swarm load_swarm(const BatchProb* B, ProbLds* L) {
  Swarm S;
  S.d = B->P.d;
  S.n = B->n;
  S.np = B->npad;
  S.k0 = B->key0;
  S.k1 = B->key1;
  S.p1 = B->P.par[0];
  S.p2 = B->P.par[1];
  S.c1 = B->c1;
  S.c2 = B->c2;
  S.has_shift = B->P.has_shift != 0;
  S.X = B->x;
  S.V = B->v;
  S.PB = B->pb;
  S.PC = B->pc;
  if ((int)threadIdx.x < S.d) {
    L->lb[threadIdx.x] = B->P.lb[threadIdx.x];
    L->ub[threadIdx.x] = B->P.ub[threadIdx.x];
    L->shift[threadIdx.x] = S.has_shift ? B->P.shift[threadIdx.x] : 0.0f;
  }
  return S;
}

// Uniform init (as init_position + objective) of the tiles first,
// first + stride, ... of Q * BATCH_BLOCK particles; returns this thread's
// minimum key.
template <int OBJ, int Q>
__device__ uint64_t init_sweep(const Swarm& S, const ProbLds* L, int64_t first, int64_t stride) {
  uint64_t mykey = KEY_NONE;
  for (int64_t tile = first; tile * (Q * BATCH_BLOCK) < S.n; tile += stride) {
    for (int q = 0; q < Q; ++q) {
      const int64_t i = tile * (Q * BATCH_BLOCK) + (int64_t)q * BATCH_BLOCK + threadIdx.x;
      if (i >= S.n) continue;
      Accum<OBJ> acc;
      acc.start(S.d);
      u4 r{};
      for (int j = 0; j < S.d; ++j) {
        if ((j & 3) == 0)
          r = philox4x32_10(u4{(uint32_t)i, 0u, (uint32_t)(j >> 2) | (PPS_TAG_INIT << 16), 0u},
                            S.k0, S.k1);
        const uint32_t rr = (j & 3) == 0 ? r.x : (j & 3) == 1 ? r.y : (j & 3) == 2 ? r.z : r.w;
        const float x = u01(rr) * (L->ub[j] - L->lb[j]) + L->lb[j];
        const uint32_t o = (uint32_t)i * 4u + (uint32_t)j * ((uint32_t)S.np * 4u);
        at32(S.X, o) = x;
        at32(S.V, o) = 0.0f;
        at32(S.PB, o) = x;
        acc.add(j, S.has_shift ? x - L->shift[j] : x, S.p1, S.p2);
      }
      S.PC[i] = acc.res;
      mykey = umin64(mykey, make_key(acc.res, (uint32_t)i));
    }
  }
  return mykey;
}

// One iteration t (inertia w, attractor L->gb) over the tiles first,
// first + stride, ...; returns this thread's minimum key.  Coordinates are
// streamed pair by pair (one Philox call per pair, as velocity_position()),
// PPT particles per thread in flight; an improving particle re-reads its new
// position to store its best.
template <int OBJ>
__device__ uint64_t step_sweep(const Swarm& S, const ProbLds* L, int64_t first, int64_t stride,
                               uint32_t t, float w) {
  uint64_t mykey = KEY_NONE;
  const int d = S.d;
  const uint32_t row4 = (uint32_t)S.np * 4u;
  const float c1s = scaled_coef(S.c1), c2s = scaled_coef(S.c2);
  for (int64_t tile = first; tile * TILE < S.n; tile += stride) {
    const int64_t i0 = tile * TILE + threadIdx.x;
    Accum<OBJ> acc[PPT];
    float pcq[PPT];
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
      acc[q].start(d);
      // the best costs go out with the first coordinate pair's loads
      const int64_t i = i0 + (int64_t)q * BATCH_BLOCK;
      pcq[q] = i < S.n ? S.PC[i] : 0.0f;
    }
    for (int k = 0; 2 * k < d; ++k) {
      float x[PPT][2], v[PPT][2], pbv[PPT][2];
#pragma unroll
      for (int q = 0; q < PPT; ++q) {
        const int64_t i = i0 + (int64_t)q * BATCH_BLOCK;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int j = 2 * k + h;
          const uint32_t o = (uint32_t)i * 4u + (uint32_t)j * row4;
          const bool ok = i < S.n && j < d;
          x[q][h] = ok ? at32(S.X, o) : 0.0f;
          v[q][h] = ok ? at32(S.V, o) : 0.0f;
          pbv[q][h] = ok ? at32(S.PB, o) : 0.0f;
        }
      }
#pragma unroll
      for (int q = 0; q < PPT; ++q) {
        const int64_t i = i0 + (int64_t)q * BATCH_BLOCK;
        if (i >= S.n) continue;
        const u4 r = philox4x32_10(u4{(uint32_t)i, t, (uint32_t)k | (PPS_TAG_STEP << 16), 0u},
                                   S.k0, S.k1);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int j = 2 * k + h;
          if (j < d) {
            const float m1 = draw_m(h == 0 ? r.x : r.z);
            const float m2 = draw_m(h == 0 ? r.y : r.w);
            const float a = w * v[q][h];
            const float b = (c1s * m1) * (pbv[q][h] - x[q][h]);
            const float c = (c2s * m2) * (L->gb[j] - x[q][h]);
            const float vn = (a + b) + c;
            float xn = x[q][h] + vn;
            xn = jl_min(jl_max(xn, L->lb[j]), L->ub[j]);
            const uint32_t o = (uint32_t)i * 4u + (uint32_t)j * row4;
            at32(S.X, o) = xn;
            at32(S.V, o) = vn;
            acc[q].add(j, S.has_shift ? xn - L->shift[j] : xn, S.p1, S.p2);
          }
        }
      }
    }
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
      const int64_t i = i0 + (int64_t)q * BATCH_BLOCK;
      if (i >= S.n) continue;
      float pc = pcq[q];
      if (acc[q].res < pc) {
        copy_particle_rows<4>(S.PB, S.X, (uint32_t)i * 4u, row4, d);
        pc = acc[q].res;
        S.PC[i] = pc;
      }
      mykey = umin64(mykey, make_key(pc, (uint32_t)i));
    }
  }
  return mykey;
}

// One iteration t of a PART item over the tiles first, first + stride, ... of
// PTILE particles, one particle per thread, PCOORD coordinates of x, v and
// the best position loaded per round (the same arithmetic as step_sweep);
// returns this thread's minimum key.
template <int OBJ>
__device__ uint64_t part_step_sweep(const Swarm& S, const ProbLds* L, int64_t first,
                                    int64_t stride, uint32_t t, float w) {
  uint64_t mykey = KEY_NONE;
  const int d = S.d;
  const float c1s = scaled_coef(S.c1), c2s = scaled_coef(S.c2);
  for (int64_t tile = first; tile * PTILE < S.n; tile += stride) {
    const int64_t i = tile * PTILE + threadIdx.x;
    if (i >= S.n) continue;
    const uint32_t i4 = (uint32_t)i * 4u, row4 = (uint32_t)S.np * 4u;
    // the best cost goes out with the first round of loads rather than after
    // the last (one dependent round trip less per tile)
    float pc = S.PC[i];
    Accum<OBJ> acc;
    acc.start(d);
    for (int g = 0; g < d; g += PCOORD) {
      float x[PCOORD], v[PCOORD], pbv[PCOORD];
#pragma unroll
      for (int h = 0; h < PCOORD; ++h) {
        const int j = g + h;
        const uint32_t o = i4 + (uint32_t)j * row4;
        const bool ok = j < d;
        x[h] = ok ? at32(S.X, o) : 0.0f;
        v[h] = ok ? at32(S.V, o) : 0.0f;
        pbv[h] = ok ? at32(S.PB, o) : 0.0f;
      }
#pragma unroll
      for (int h2 = 0; h2 < PCOORD; h2 += 2) {
        const int j0 = g + h2;
        if (j0 >= d) break;
        const u4 r = philox4x32_10(
            u4{(uint32_t)i, t, (uint32_t)(j0 >> 1) | (PPS_TAG_STEP << 16), 0u}, S.k0, S.k1);
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          const int h = h2 + e, j = j0 + e;
          if (j < d) {
            const float m1 = draw_m(e == 0 ? r.x : r.z);
            const float m2 = draw_m(e == 0 ? r.y : r.w);
            const float a = w * v[h];
            const float b = (c1s * m1) * (pbv[h] - x[h]);
            const float c = (c2s * m2) * (L->gb[j] - x[h]);
            const float vn = (a + b) + c;
            float xn = x[h] + vn;
            xn = jl_min(jl_max(xn, L->lb[j]), L->ub[j]);
            const uint32_t o = i4 + (uint32_t)j * row4;
            at32(S.X, o) = xn;
            at32(S.V, o) = vn;
            acc.add(j, S.has_shift ? xn - L->shift[j] : xn, S.p1, S.p2);
          }
        }
      }
    }
    if (acc.res < pc) {
      copy_particle_rows<8>(S.PB, S.X, i4, row4, d);
      pc = acc.res;
      S.PC[i] = pc;
    }
    mykey = umin64(mykey, make_key(pc, (uint32_t)i));
  }
  return mykey;
}

// The best position of the workgroup minimum `bk` (known to every thread) to
// dst[0..d), one coordinate per thread.  Called after the block reduction,
// whose barriers order this workgroup's best-position writes before the reads.
__device__ __forceinline__ void copy_winner(const Swarm& S, uint64_t bk, float* dst) {
  const int j = (int)threadIdx.x;
  if (bk != KEY_NONE && j < S.d)
    dst[j] = at32(S.PB, (uint32_t)bk * 4u + (uint32_t)j * ((uint32_t)S.np * 4u));
}

// SMALL item: run iterations until the problem ends or the deadline passes.
// Returns true (in every thread) when the problem ended in this item: it ran
// the last iteration, or it has none and was initialised here.  Only that
// item publishes the result, so no kernel writes a result slot after the host
// may have seen its done flag (a finished problem can still be listed in a
// slice the host built while the finishing slice ran).
template <int OBJ>
__device__ bool run_small(BatchProb* B, uint64_t deadline, uint64_t* s_red, ProbLds* L,
                          int* s_stop) {
  const Swarm S = load_swarm(B, L);
  const bool was_initialized = B->initialized != 0;
  __syncthreads();
  if (!was_initialized) {
    const uint64_t mykey = init_sweep<OBJ, PPT>(S, L, 0, 1);
    const uint64_t bk = block_min<BATCH_BLOCK>(mykey, s_red);
    copy_winner(S, bk, L->gb);
    __syncthreads();
    if ((int)threadIdx.x < S.d) B->gb[threadIdx.x] = L->gb[threadIdx.x];
    if (threadIdx.x == 0) {
      B->gb_cost = key_cost(bk);
      B->gb_idx = (uint32_t)bk;
      B->initialized = 1;
    }
    __syncthreads();
  }
  const int32_t steps = B->remaining;
  if (steps <= 0) return !was_initialized;
  if ((int)threadIdx.x < S.d) L->gb[threadIdx.x] = B->gb[threadIdx.x];
  const uint32_t t0 = B->iter;
  float w = B->w;
  const float wdamp = B->wdamp;
  if (threadIdx.x == 0) *s_stop = 0;
  __syncthreads();
  uint64_t bk = KEY_NONE;
  int32_t done = 0;
  uint64_t prev = wall_clock64();
  while (done < steps) {
    const uint32_t t = t0 + (uint32_t)(++done);
    // d > 4: sixteen coordinates of one particle per round of loads; else
    // four particles of one coordinate pair
    const uint64_t mykey = S.d > 4 ? part_step_sweep<OBJ>(S, L, 0, 1, t, w)
                                   : step_sweep<OBJ>(S, L, 0, 1, t, w);
    bk = block_min<BATCH_BLOCK>(mykey, s_red);
    copy_winner(S, bk, L->gb);
    if (threadIdx.x == 0) {
      // stop when another iteration as long as this one would end past the
      // deadline (the first iteration always runs)
      const uint64_t now = wall_clock64();
      *s_stop = now + (now - prev) > deadline;
      prev = now;
    }
    __syncthreads();
    w = w * wdamp;
    if (*s_stop) break;
  }
  if ((int)threadIdx.x < S.d) B->gb[threadIdx.x] = L->gb[threadIdx.x];
  if (threadIdx.x == 0) {
    B->gb_cost = key_cost(bk);
    B->gb_idx = (uint32_t)bk;
    B->iter = t0 + (uint32_t)done;
    B->remaining -= done;
    B->w = w;
  }
  return done >= steps;
}

// Global best of a large problem from this thread's share `key` of its part
// keys of parity `par`: the block minimum, and the winner's position (the
// winning part is the one owning the winner's tile; tiles of PTILE particles
// are dealt round robin to the k parts) copied to dst.
__device__ __forceinline__ uint64_t part_gbest_from(const BatchXchg* XC, uint64_t key, int k,
                                                    uint32_t par, int d, float* dst,
                                                    uint64_t* s_red) {
  key = block_min<BATCH_BLOCK>(key, s_red);
  const int64_t wi = (int64_t)(uint32_t)key;
  const int wpart = (int)((wi / PTILE) % k);
  if ((int)threadIdx.x < d) dst[threadIdx.x] = XC->cand[par][wpart][threadIdx.x];
  __syncthreads();
  return key;
}
// This thread's share of the part keys of parity `par`.
__device__ __forceinline__ uint64_t part_keys(const BatchXchg* XC, int k, uint32_t par) {
  uint64_t key = KEY_NONE;
  for (int q = threadIdx.x; q < k; q += BATCH_BLOCK) key = umin64(key, XC->key[par][q]);
  return key;
}
// Global best of a large problem after iteration `t`: minimum of its part keys
// of parity t & 1.
__device__ __forceinline__ uint64_t part_gbest(const BatchXchg* XC, int k, uint32_t t, int d,
                                               float* dst, uint64_t* s_red) {
  return part_gbest_from(XC, part_keys(XC, k, t & 1u), k, t & 1u, d, dst, s_red);
}

// PART item: iteration t (t = 0: init) of part q of k.
template <int OBJ>
__device__ void run_part(BatchProb* B, const BatchItem& it, BatchXchg* XC, uint64_t* s_red,
                         ProbLds* L) {
  // the part keys of the previous iteration are requested with the
  // descriptor (both need only the work item)
  const uint64_t keys = it.t == 0 ? KEY_NONE : part_keys(XC, it.nparts, (it.t - 1u) & 1u);
  const Swarm S = load_swarm(B, L);
  __syncthreads();
  uint64_t mykey;
  if (it.t == 0) {
    mykey = init_sweep<OBJ, 1>(S, L, it.part, it.nparts);
  } else {
    part_gbest_from(XC, keys, it.nparts, (it.t - 1u) & 1u, S.d, L->gb, s_red);
    mykey = part_step_sweep<OBJ>(S, L, it.part, it.nparts, it.t, it.w);
  }
  const uint64_t bk = block_min<BATCH_BLOCK>(mykey, s_red);
  const uint32_t par = it.t & 1u;
  copy_winner(S, bk, XC->cand[par][it.part]);
  if (threadIdx.x == 0) XC->key[par][it.part] = bk;
}

}  // namespace

// Result of a finished problem into host-visible memory (one thread): every
// coordinate read from LDS before the stores, the done flag last with a
// system-scope release.
__device__ __forceinline__ void publish_result(BatchOut* o, float cost, uint32_t idx,
                                               uint32_t iter, const float* gb, int d) {
  float g[PPS_BMAXD];
#pragma unroll
  for (int j = 0; j < PPS_BMAXD; ++j) g[j] = j < d ? gb[j] : 0.0f;
  o->gb_cost = cost;
  o->gb_idx = idx;
  o->iter = iter;
  o->remaining = 0;
#pragma unroll
  for (int j = 0; j < PPS_BMAXD; ++j)
    if (j < d) o->gb[j] = g[j];
  __threadfence_system();
  __hip_atomic_store(&o->done, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

__global__ __launch_bounds__(BATCH_BLOCK) void k_slice(BatchProb* probs, const BatchProb* fresh,
                                                       const BatchItem* items,
                                                       uint64_t budget_ticks, BatchOut* out,
                                                       BatchXchg* xchg) {
  __shared__ uint64_t s_red[BATCH_BLOCK / 64];
  __shared__ ProbLds s_lds;
  __shared__ int s_stop;
  const uint64_t deadline = wall_clock64() + budget_ticks;
  const BatchItem it = items[blockIdx.x];
  BatchProb* B = probs + it.slot;
  if (it.kind == BATCH_SMALL) {
    if (it.fresh) {
      // newly admitted: pull the descriptor from host-visible memory
      const uint32_t* src = reinterpret_cast<const uint32_t*>(fresh + it.slot);
      uint32_t* dst = reinterpret_cast<uint32_t*>(B);
      constexpr int W = (int)(sizeof(BatchProb) / 4);
      constexpr int R = (W + BATCH_BLOCK - 1) / BATCH_BLOCK;
      uint32_t t[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {  // every load before the first store
        const int k = (int)threadIdx.x + r * BATCH_BLOCK;
        if (k < W) t[r] = src[k];
      }
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int k = (int)threadIdx.x + r * BATCH_BLOCK;
        if (k < W) dst[k] = t[r];
      }
      __syncthreads();
    }
    const bool finished = B->P.obj == PPS_OBJ_RASTRIGIN
                              ? run_small<PPS_OBJ_RASTRIGIN>(B, deadline, s_red, &s_lds, &s_stop)
                              : run_small<PPS_OBJ_ROSENBROCK>(B, deadline, s_red, &s_lds, &s_stop);
    __syncthreads();
    // s_lds.gb holds the problem's final global best position (run_small
    // leaves it there whenever it returns true)
    if (threadIdx.x == 0 && finished)
      publish_result(out + it.slot, B->gb_cost, B->gb_idx, B->iter, s_lds.gb, B->P.d);
  } else if (it.kind == BATCH_PART) {
    if (B->P.obj == PPS_OBJ_RASTRIGIN)
      run_part<PPS_OBJ_RASTRIGIN>(B, it, xchg + it.slot, s_red, &s_lds);
    else
      run_part<PPS_OBJ_ROSENBROCK>(B, it, xchg + it.slot, s_red, &s_lds);
  } else {  // BATCH_FINAL: it.t = last iteration
    const int d = B->P.d;
    const uint64_t key = part_gbest(xchg + it.slot, it.nparts, it.t, d, s_lds.gb, s_red);
    if (threadIdx.x == 0)
      publish_result(out + it.slot, key_cost(key), (uint32_t)key, it.t, s_lds.gb, d);
  }
}

bool batch_supported(int obj, int d) {
  if (obj != PPS_OBJ_ROSENBROCK && obj != PPS_OBJ_RASTRIGIN) return false;
  return d >= 1 && d <= PPS_BMAXD;
}

int batch_parts(int64_t n, int tiles_per_part) {
  const int64_t tiles = (n + PTILE - 1) / PTILE;
  const int64_t parts = (tiles + tiles_per_part - 1) / tiles_per_part;
  return (int)(parts < BATCH_MAXPARTS ? parts : BATCH_MAXPARTS);
}

hipError_t launch_slice(BatchProb* probs, const BatchProb* fresh, const BatchItem* items,
                        int32_t count, uint64_t budget_ticks, BatchOut* out, BatchXchg* xchg,
                        hipStream_t st) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_slice, dim3(count), dim3(BATCH_BLOCK), 0, st, probs, fresh, items,
                     budget_ticks, out, xchg);
  return hipGetLastError();
}

}  // namespace pps
