#pragma once
// pps_kernels_impl.h — particle-swarm kernel templates for gfx950 (included by
// one translation unit per objective family).
//
// Layout: swarm state is structure-of-arrays in HBM (x, v, best position:
// d rows of npad floats; best cost: npad floats), so a wavefront's load of
// coordinate j of 64 consecutive particles is one 256-byte coalesced access.
// One work-item steps one particle per tile; workgroup b owns tiles b, b+G, ...
//
// Kernels (reference counterparts in brackets):
//   k_init      initial positions (uniform / unbounded / Latin hypercube /
//               given) + cost + block argmin   [gpu_init_particles!,
//               src/utils.jl:23-112; minimum, src/init.jl:49-51]
//   k_sync_step fused velocity/position/clamp/objective/pbest + block argmin
//               [update_particle_states! sync, src/kernels.jl:127-194 and the
//               per-iteration minimum of src/lowerlevel_solve.jl:23-33]
//   k_sync_wg   the whole sync loop in one workgroup, swarm in registers
//   k_sync_wave the same for a swarm of one wavefront (no LDS, no barrier)
//   k_async     all iterations per particle in registers
//               [update_particles_states_async!, src/kernels.jl:196-225]
//   k_qlock     one iteration against the live global best, lock-protected
//               publication [src/kernels.jl:38-125]
//   k_finalize  grid pass over the block keys -> global best record
#include <hip/hip_runtime.h>

#include "pps_device.h"
#include "pps_kernels.h"

namespace pps {

constexpr uint64_t KEY_NONE = ~0ull;
#define NO_BOUND __uint_as_float(0x7FC00000u)  // NaN: an ODE loss is summed in full

// Minimum over a workgroup.  Every thread receives the result.
template <int BLOCK>
__device__ __forceinline__ uint64_t block_min_u64(uint64_t k, uint64_t* s_red) {
  k = wave_min_u64(k);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (BLOCK == 64) return k;
  if (lane == 0) s_red[wid] = k;
  __syncthreads();
  uint64_t r = s_red[0];
#pragma unroll
  for (int i = 1; i < BLOCK / 64; ++i) r = umin64(r, s_red[i]);
  __syncthreads();
  return r;
}

// The same without the trailing barrier: for a kernel's last reduction, or
// one followed by a barrier before its slots are written again.
template <int BLOCK>
__device__ __forceinline__ uint64_t block_min_u64_once(uint64_t k, uint64_t* s_red) {
  k = wave_min_u64(k);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (BLOCK == 64) return k;
  if (lane == 0) s_red[wid] = k;
  __syncthreads();
  uint64_t r = s_red[0];
#pragma unroll
  for (int i = 1; i < BLOCK / 64; ++i) r = umin64(r, s_red[i]);
  return r;
}

// Minimum over a workgroup of nw wavefronts (nw = blockDim.x / 64, uniform);
// with one wavefront no barrier is taken.  Every thread receives the result.
__device__ __forceinline__ uint64_t block_min_u64_n(uint64_t k, uint64_t* s_red, int nw) {
  k = wave_min_u64(k);
  if (nw == 1) return k;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) s_red[wid] = k;
  __syncthreads();
  uint64_t r = s_red[0];
  for (int i = 1; i < nw; ++i) r = umin64(r, s_red[i]);
  __syncthreads();
  return r;
}

template <int BLOCK>
__device__ __forceinline__ int block_sum_i32(int c, int* s_cnt) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) c += __shfl_xor(c, off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (BLOCK == 64) return c;
  if (lane == 0) s_cnt[wid] = c;
  __syncthreads();
  int r = 0;
#pragma unroll
  for (int i = 0; i < BLOCK / 64; ++i) r += s_cnt[i];
  __syncthreads();
  return r;
}

template <int MAXD>
__device__ __forceinline__ void load_vec(float (&r)[MAXD], const float* base, int64_t npad,
                                         int64_t i, int d) {
#pragma unroll
  for (int j = 0; j < MAXD; ++j)
    if (j < d) r[j] = base[(int64_t)j * npad + i];
}
template <int MAXD>
__device__ __forceinline__ void store_vec(const float (&r)[MAXD], float* base, int64_t npad,
                                          int64_t i, int d) {
#pragma unroll
  for (int j = 0; j < MAXD; ++j)
    if (j < d) base[(int64_t)j * npad + i] = r[j];
}

// Global best of parity `par`: minimum block key, then the owning block's
// candidate position into s_gb (all threads must call).
template <int BLOCK>
__device__ __forceinline__ uint64_t fetch_gbest(const SwarmDev& S, uint32_t par, int d,
                                                float* s_gb, uint64_t* s_red) {
  const int G = S.G;
  uint64_t k = KEY_NONE;
  for (int j = threadIdx.x; j < G; j += BLOCK) k = umin64(k, S.bkey[(int64_t)par * G + j]);
  k = block_min_u64<BLOCK>(k, s_red);
  const uint32_t gidx = (uint32_t)k;
  const int gblk = (int)((gidx / (uint32_t)S.block) % (uint32_t)G);
  if ((int)threadIdx.x < d) s_gb[threadIdx.x] = S.cand[((int64_t)par * G + gblk) * d + threadIdx.x];
  __syncthreads();
  return k;
}

// Block argmin over `mykey`; the winner's position (its rows of src) is
// copied into cand[par][b], one coordinate per thread: the reduction's
// barriers order this workgroup's writes of src before the reads, and the d
// loads are one round trip instead of d dependent ones.
template <int BLOCK>
__device__ __forceinline__ void publish_block(const SwarmDev& S, uint32_t par, int d,
                                              uint64_t mykey, const float* src,
                                              uint64_t* s_red) {
  const uint64_t bk = block_min_u64_once<BLOCK>(mykey, s_red);
  const int b = blockIdx.x;
  if (bk != KEY_NONE) {
    const int64_t wi = (int64_t)(uint32_t)bk;
    for (int j = threadIdx.x; j < d; j += BLOCK)
      S.cand[((int64_t)par * S.G + b) * d + j] = src[(int64_t)j * S.npad + wi];
  }
  if (threadIdx.x == 0) S.bkey[(int64_t)par * S.G + b] = bk;
}

// Cost of a position: the objective, plus the constraint penalty when the
// problem has constraints (src/kernels.jl:28-36).  `iter1` is h's argument
// as the reference computes it at each call site (init: 1; sync and queue-lock
// kernels: 1-based particle index + 1; async: 1-based iteration of the launch
// + 1).  Constrained problems run on the runtime-dimension instantiations, so
// the branch is compiled only there.
// `best` is the particle's best cost when the caller only compares the result
// with it (an adaptive ODE loss may then stop early, see ode_sse_tsit5), NaN
// otherwise.
template <int MAXD, bool EXACT, int OBJ>
__device__ __forceinline__ float cost_at(const float (&x)[MAXD], const int d, const ProbArgs& P,
                                         float iter1, float best) {
  float c = objective<MAXD, OBJ>(x, d, P, P.ode_stop ? best : NO_BOUND);
  if constexpr (!EXACT && (OBJ == PPS_OBJ_ROSENBROCK || OBJ == PPS_OBJ_RASTRIGIN ||
                           OBJ == PPS_OBJ_SPHERE)) {
    if (P.n_cons > 0) c = c + constraint_penalty<MAXD>(x, d, P, sqrtf(iter1));
  }
  return c;
}

// ---------------------------------------------------------------------------
template <int MAXD, bool EXACT, int OBJ, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_init(SwarmDev S, ProbArgs P, uint32_t par) {
  __shared__ uint64_t s_red[BLOCK / 64];
  const int d = EXACT ? MAXD : P.d;
  uint64_t mykey = KEY_NONE;
  for (int64_t tile = blockIdx.x; tile * BLOCK < S.n; tile += gridDim.x) {
    const int64_t i = tile * BLOCK + threadIdx.x;
    if (i < S.n) {
      float x[MAXD], z[MAXD];
      // sampler: uniform (src/utils.jl:57-84), unbounded (src/utils.jl:23-55),
      // Latin hypercube, or positions already in the x rows (the QMC path,
      // src/utils.jl:86-112)
      if (S.sampler == PPS_SAMPLER_POSITIONS)
        load_vec<MAXD>(x, S.x, S.npad, i, d);
      else if (S.sampler == PPS_SAMPLER_UNBOUNDED)
        unbounded_position<MAXD>(x, (uint32_t)i, S.epoch, S.key0, S.key1, d, S.u0);
      else if (S.sampler == PPS_SAMPLER_LATIN_HYPERCUBE)
        lhs_position<MAXD>(x, (uint32_t)i, (uint32_t)S.n, S.epoch, S.key0, S.key1, d, P);
      else
        init_position<MAXD>(x, (uint32_t)i, S.epoch, S.key0, S.key1, d, P);
#pragma unroll
      for (int j = 0; j < MAXD; ++j) z[j] = 0.0f;
      const float c = cost_at<MAXD, EXACT, OBJ>(x, d, P, 1.0f, NO_BOUND);
      store_vec<MAXD>(x, S.x, S.npad, i, d);
      store_vec<MAXD>(z, S.v, S.npad, i, d);
      store_vec<MAXD>(x, S.pb, S.npad, i, d);
      S.pc[i] = c;
      mykey = umin64(mykey, make_key(c, (uint32_t)i));
    }
  }
  publish_block<BLOCK>(S, par, d, mykey, S.pb, s_red);
}

template <int MAXD, bool EXACT, int OBJ, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_rekey(SwarmDev S, ProbArgs P, uint32_t par,
                                                 int32_t gbest_from) {
  __shared__ uint64_t s_red[BLOCK / 64];
  const int d = EXACT ? MAXD : P.d;
  uint64_t mykey = KEY_NONE;
  for (int64_t tile = blockIdx.x; tile * BLOCK < S.n; tile += gridDim.x) {
    const int64_t i = tile * BLOCK + threadIdx.x;
    if (i < S.n) mykey = umin64(mykey, make_key(S.pc[i], (uint32_t)i));
  }
  publish_block<BLOCK>(S, par, d, mykey, gbest_from == PPS_GBEST_CURRENT ? S.x : S.pb, s_red);
}

// Occupancy target of the HBM-streaming sync step (waves per SIMD); 0 leaves
// register allocation to the compiler.
#ifndef PPS_SYNC_WAVES_PER_EU
#define PPS_SYNC_WAVES_PER_EU 0
#endif
#if PPS_SYNC_WAVES_PER_EU > 0
#define PPS_SYNC_ATTR __attribute__((amdgpu_waves_per_eu(PPS_SYNC_WAVES_PER_EU)))
#else
#define PPS_SYNC_ATTR
#endif

// Row addressing of the SoA state.  x, v and the best positions share one
// row stride, so coordinate j of particle i is at the same byte offset in all
// three.  With OFF32 (every array under 4 GiB) that offset is a 32-bit value
// added to the array's base held in scalar registers, so a thread keeps one
// 32-bit register per row instead of a 64-bit address per row and array.
template <bool OFF32>
struct RowAddr;
template <>
struct RowAddr<true> {
  uint32_t i4, row4;
  __device__ __forceinline__ RowAddr(int64_t i, int64_t npad)
      : i4((uint32_t)i * 4u), row4((uint32_t)npad * 4u) {}
  __device__ __forceinline__ float* at(float* base, int j) const {
    return (float*)((char*)base + (i4 + (uint32_t)j * row4));
  }
};
template <>
struct RowAddr<false> {
  int64_t i, npad;
  __device__ __forceinline__ RowAddr(int64_t i_, int64_t np) : i(i_), npad(np) {}
  __device__ __forceinline__ float* at(float* base, int j) const {
    return base + (int64_t)j * npad + i;
  }
};

// State streams of the sync step, with nontemporal hints (NT) when the swarm
// state outgrows the 256 MB Infinity Cache: nothing of it is reused within a
// launch, and beyond the cache the hinted loads and stores stream 4-10 %
// faster (measured 2^22 x 10-D: 159 vs 177 us), while a cache-resident state
// is 7-13 % slower with them (2^20 x 10-D: 42 vs 39 us).  Loads alone or
// stores alone gain nothing.
#define PPS_NT_MIN_STATE_BYTES (288ull << 20)
template <bool NT>
__device__ __forceinline__ float ld_state(const float* p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}
template <bool NT>
__device__ __forceinline__ void st_state(float* p, float v) {
  if constexpr (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

// Start of a launch: the global best of iteration t-1 (a minimum over the G
// block keys, then the winner's position from its block's candidate) is two
// dependent round trips to memory.  Nothing of the first tile depends on it
// but the update itself, so the block keys are requested first, the first
// tile's state loads are issued behind them, and (MAXD <= 16) the tile's
// random draws are computed while both are in flight.  At the end the block's
// winner publishes its position from registers instead of reading back what
// it stored.
// Inertia weight of launch k: the w argument, or in a replayed chunk graph
// the run's table entry for this launch's iteration.
__device__ __forceinline__ float step_w(const SwarmDev& S, const StepArgs& A, float w,
                                        uint32_t k) {
  return A.w_table ? S.wtab[S.tbase[2] + k - 1u] : w;
}

template <int MAXD, bool EXACT, int OBJ, int BLOCK, bool OFF32, bool NT>
__global__ __launch_bounds__(BLOCK) PPS_SYNC_ATTR void k_sync_step(SwarmDev S, ProbArgs P,
                                                                   StepArgs A, float w,
                                                                   uint32_t k) {
  // Both cost registers: they are compiled where the workgroup size leaves
  // room for them without spilling.
  constexpr bool PRE = MAXD * BLOCK <= 16 * 256;   // draws computed ahead of the state
  constexpr bool HOLD = MAXD * BLOCK <= 32 * 256;  // block winner's position in registers
  __shared__ uint64_t s_red[BLOCK / 64];
  __shared__ float s_gb[MAXD];
  const uint32_t t = S.tbase[0] + k;
  const uint32_t epoch = S.tbase[1];
  w = step_w(S, A, w, k);
  const int d = EXACT ? MAXD : P.d;
  const int G = S.G;
  const int tid = threadIdx.x;
  const bool cur = A.gbest_from == PPS_GBEST_CURRENT;
  const uint64_t* keys = S.bkey + (int64_t)((t - 1u) & 1u) * G;
  uint64_t gk = tid < G ? keys[tid] : KEY_NONE;
  // The key request goes out before the state loads: loads return in issue
  // order, so the block reduction can start before the state has arrived.
  __builtin_amdgcn_sched_barrier(0);

  int64_t tile = blockIdx.x;
  int64_t i = tile * BLOCK + tid;
  float x[MAXD], v[MAXD], pb[MAXD], r1[MAXD], r2[MAXD], pc = 0.0f;
#pragma unroll
  for (int j = 0; j < MAXD; ++j) x[j] = v[j] = pb[j] = r1[j] = r2[j] = 0.0f;
  auto load_tile = [&]() {
    const RowAddr<OFF32> R(i, S.npad);
#pragma unroll
    for (int j = 0; j < MAXD; ++j) {
      if (j < d) {
        x[j] = ld_state<NT>(R.at(S.x, j));
        v[j] = ld_state<NT>(R.at(S.v, j));
        pb[j] = ld_state<NT>(R.at(S.pb, j));
      }
    }
    pc = ld_state<NT>(S.pc + i);
  };
  if (i < S.n) load_tile();
  // Keep the loads ahead of everything below.
  __builtin_amdgcn_sched_barrier(0);
  if (PRE && i < S.n) step_draws<MAXD>(r1, r2, (uint32_t)i, t, epoch, S.key0, S.key1, d);
  for (int j = tid + BLOCK; j < G; j += BLOCK) gk = umin64(gk, keys[j]);
  // (the barrier after s_gb below orders these slot reads before the end's
  // slot writes)
  gk = block_min_u64_once<BLOCK>(gk, s_red);
  {
    const uint32_t gidx = (uint32_t)gk;
    const int gblk = (int)((gidx / (uint32_t)S.block) % (uint32_t)G);
    if (tid < d) s_gb[tid] = S.cand[((int64_t)((t - 1u) & 1u) * G + gblk) * d + tid];
  }
  __syncthreads();
  float gb[MAXD];
#pragma unroll
  for (int j = 0; j < MAXD; ++j) gb[j] = j < d ? s_gb[j] : 0.0f;

  uint64_t mykey = KEY_NONE;
  float hold[MAXD];
#pragma unroll
  for (int j = 0; j < MAXD; ++j) hold[j] = 0.0f;
  while (tile * BLOCK < S.n) {
    if (i < S.n) {
      const RowAddr<OFF32> R(i, S.npad);
      if constexpr (PRE) {
        velocity_position_drawn<MAXD, OBJ>(x, v, pb, gb, w, A.c1, A.c2, r1, r2, d, P);
      } else {
        velocity_position<MAXD, OBJ>(x, v, pb, gb, w, A.c1, A.c2, (uint32_t)i, t, epoch,
                                     S.key0, S.key1, d, P);
      }
      const float c = cost_at<MAXD, EXACT, OBJ>(x, d, P, (float)(i + 2), pc);
#pragma unroll
      for (int j = 0; j < MAXD; ++j) {
        if (j < d) {
          st_state<NT>(R.at(S.x, j), x[j]);
          st_state<NT>(R.at(S.v, j), v[j]);
        }
      }
      // Best positions are written in whole 64-byte segments (16 lanes of a
      // row) whenever any particle of the segment improved (the others
      // rewrite the values they loaded): a segment written by a few lanes is
      // a partial write, which HBM turns into a read-modify-write of the
      // whole ECC word.  Measured at 2^22 x 10-D Rosenbrock, 200 iterations
      // after 1000 (us per launch): 170.1 with whole 64-lane rows, 164.1 with
      // 16-lane segments, 168.1 with 8- or 4-lane segments.
      const bool imp = c < pc;
      const uint32_t seg = (uint32_t)(threadIdx.x & 63) & ~15u;
      if ((__ballot(imp) >> seg) & 0xFFFFull) {
#pragma unroll
        for (int j = 0; j < MAXD; ++j)
          if (j < d) st_state<NT>(R.at(S.pb, j), imp ? x[j] : pb[j]);
        st_state<NT>(S.pc + i, imp ? c : pc);
      }
      if (imp) pc = c;
      const uint64_t key = make_key(pc, (uint32_t)i);
      if constexpr (HOLD) {
        // Particles of one thread come in increasing index order, so a
        // strictly smaller key is the thread's new argmin.
        const bool win = key < mykey;
#pragma unroll
        for (int j = 0; j < MAXD; ++j)
          if (win) hold[j] = (cur || imp) ? x[j] : pb[j];
      }
      mykey = umin64(mykey, key);
    }
    tile += gridDim.x;
    i = tile * BLOCK + tid;
    if (tile * BLOCK < S.n && i < S.n) {
      load_tile();
      // Issue the whole tile's loads before any arithmetic: left alone, the
      // scheduler interleaves them with the arithmetic pair by pair and the
      // wave waits for memory several times per tile.
      __builtin_amdgcn_sched_barrier(0);
      if (PRE) step_draws<MAXD>(r1, r2, (uint32_t)i, t, epoch, S.key0, S.key1, d);
    }
  }
  if constexpr (HOLD) {
    const uint64_t bk = block_min_u64_once<BLOCK>(mykey, s_red);
    if (bk != KEY_NONE && mykey == bk) {
#pragma unroll
      for (int j = 0; j < MAXD; ++j)
        if (j < d) S.cand[((int64_t)(t & 1u) * G + blockIdx.x) * d + j] = hold[j];
    }
    if (tid == 0) S.bkey[(int64_t)(t & 1u) * G + blockIdx.x] = bk;
  } else {
    publish_block<BLOCK>(S, t & 1u, d, mykey, cur ? S.x : S.pb, s_red);
  }
}

// Sync step of the ODE estimation (either model, any of the five solvers)
// with two lanes per particle: a 64-particle tile (the tiling of k_init and
// k_sync_step, so block keys and candidates mean the same) runs on 128
// threads, lanes 2m and 2m+1 holding particle m.  Both lanes load the
// particle and compute the same update; the loss, a chain of hundreds to
// thousands of dependent solver steps per evaluation, is split between them
// (pair_objective), which halves the chain each lane issues and doubles
// the waves over the SIMDs.  The even lane writes the state.
template <int OBJ, bool OFF32>
__global__ __launch_bounds__(128) void k_sync_step_pair(SwarmDev S, ProbArgs P, StepArgs A,
                                                           float w, uint32_t k) {
  constexpr int MAXD = 4, TP = 64, BLOCK = 2 * TP;
  __shared__ uint64_t s_red[BLOCK / 64];
  __shared__ float s_gb[MAXD];
  const uint32_t t = S.tbase[0] + k;
  const uint32_t epoch = S.tbase[1];
  w = step_w(S, A, w, k);
  const int d = MAXD;
  const int G = S.G;
  const int tid = threadIdx.x;
  const bool cur = A.gbest_from == PPS_GBEST_CURRENT;
  const uint64_t* keys = S.bkey + (int64_t)((t - 1u) & 1u) * G;
  uint64_t gk = KEY_NONE;
  for (int j = tid; j < G; j += BLOCK) gk = umin64(gk, keys[j]);
  gk = block_min_u64<BLOCK>(gk, s_red);
  {
    const uint32_t gidx = (uint32_t)gk;
    const int gblk = (int)((gidx / (uint32_t)S.block) % (uint32_t)G);
    if (tid < d) s_gb[tid] = S.cand[((int64_t)((t - 1u) & 1u) * G + gblk) * d + tid];
  }
  __syncthreads();
  float gb[MAXD];
#pragma unroll
  for (int j = 0; j < MAXD; ++j) gb[j] = s_gb[j];
  const bool even = (tid & 1) == 0;
  uint64_t mykey = KEY_NONE;
  float hold[MAXD];
#pragma unroll
  for (int j = 0; j < MAXD; ++j) hold[j] = 0.0f;
  for (int64_t tile = blockIdx.x; tile * TP < S.n; tile += gridDim.x) {
    const int64_t i = tile * TP + (tid >> 1);
    if (i < S.n) {  // the same for both lanes of a pair
      const RowAddr<OFF32> R(i, S.npad);
      float x[MAXD], v[MAXD], pb[MAXD];
#pragma unroll
      for (int j = 0; j < MAXD; ++j) {
        x[j] = *R.at(S.x, j);
        v[j] = *R.at(S.v, j);
        pb[j] = *R.at(S.pb, j);
      }
      float pc = S.pc[i];
      velocity_position<MAXD, OBJ>(x, v, pb, gb, w, A.c1, A.c2, (uint32_t)i, t, epoch,
                                   S.key0, S.key1, d, P);
      float z[MAXD];
#pragma unroll
      for (int j = 0; j < MAXD; ++j) z[j] = x[j] - P.shift[j];  // as objective()
      // as cost_at(): the particle's best cost bounds an adaptive loss
      const float c = pair_objective<OBJ>(z, P, P.ode_stop ? pc : NO_BOUND);
      const bool imp = c < pc;
      if (even) {
#pragma unroll
        for (int j = 0; j < MAXD; ++j) {
          *R.at(S.x, j) = x[j];
          *R.at(S.v, j) = v[j];
          if (imp) *R.at(S.pb, j) = x[j];
        }
        if (imp) S.pc[i] = c;
      }
      if (imp) pc = c;
      const uint64_t key = make_key(pc, (uint32_t)i);
      const bool win = key < mykey;
#pragma unroll
      for (int j = 0; j < MAXD; ++j)
        if (win) hold[j] = (cur || imp) ? x[j] : pb[j];
      mykey = umin64(mykey, key);
    }
  }
  const uint64_t bk = block_min_u64<BLOCK>(mykey, s_red);
  if (bk != KEY_NONE && mykey == bk && even) {
#pragma unroll
    for (int j = 0; j < MAXD; ++j) S.cand[((int64_t)(t & 1u) * G + blockIdx.x) * d + j] = hold[j];
  }
  if (tid == 0) S.bkey[(int64_t)(t & 1u) * G + blockIdx.x] = bk;
}

// Whole sync loop in one workgroup of nw = ceil(n / 64) wavefronts (at most
// BLOCK threads), one particle per thread held in registers.  The reduction
// slots alternate by iteration parity, so a slot is rewritten only after the
// barrier of the following iteration.  For MAXD <= 8 every thread also writes
// its candidate position to a parity buffer in LDS before the reduction and
// every thread reads the winner's row after it: one barrier per iteration.
// Wider problems hand the winner's position over through s_gb (two barriers).
// Swarms of one wavefront run k_sync_wave instead (a one-wavefront branch in
// this loop, even untaken, slowed 16-wavefront swarms by up to a third).
template <int MAXD, bool EXACT, int OBJ, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_sync_wg(SwarmDev S, ProbArgs P, StepArgs A,
                                                   float w, float wdamp, uint32_t t0,
                                                   int32_t T) {
  constexpr int NW = BLOCK / 64;
  static_assert(NW == 16, "the slot reduction reads 16 slots");
  constexpr bool POS_LDS = MAXD <= 8;
  __shared__ uint64_t s_red[2][NW];
  __shared__ float s_gb[MAXD];
  __shared__ float s_pos[POS_LDS ? 2 : 1][POS_LDS ? MAXD : 1][POS_LDS ? BLOCK : 1];
  const int d = EXACT ? MAXD : P.d;
  const int nw = (int)(blockDim.x >> 6);
  const int64_t i = threadIdx.x;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const bool act = i < S.n;
  // Global best handed over in slot (t0 & 1) of the one workgroup.
  float gb[MAXD];
#pragma unroll
  for (int j = 0; j < MAXD; ++j) gb[j] = j < d ? S.cand[(int64_t)(t0 & 1u) * d + j] : 0.0f;
  float x[MAXD], v[MAXD], pb[MAXD], pc = 0.0f;
#pragma unroll
  for (int j = 0; j < MAXD; ++j) x[j] = v[j] = pb[j] = 0.0f;
  if (act) {
    load_vec<MAXD>(x, S.x, S.npad, i, d);
    load_vec<MAXD>(v, S.v, S.npad, i, d);
    load_vec<MAXD>(pb, S.pb, S.npad, i, d);
    pc = S.pc[i];
  }
  const bool cur = A.gbest_from == PPS_GBEST_CURRENT;
  // Slots of wavefronts not launched hold KEY_NONE, so every iteration reads
  // all NW slots in one unrolled round of LDS loads.
  if (nw > 1) {
    if ((int)threadIdx.x < 2 * NW && (int)threadIdx.x % NW >= nw)
      s_red[threadIdx.x / NW][threadIdx.x % NW] = KEY_NONE;
    __syncthreads();
  }
  for (int32_t k = 1; k <= T; ++k) {
    const int b = k & 1;
    if (act) {
      velocity_position<MAXD, OBJ>(x, v, pb, gb, w, A.c1, A.c2, (uint32_t)i, t0 + (uint32_t)k,
                                   S.epoch, S.key0, S.key1, d, P);
      const float c = cost_at<MAXD, EXACT, OBJ>(x, d, P, (float)(i + 2), pc);
      if (c < pc) {
#pragma unroll
        for (int j = 0; j < MAXD; ++j) pb[j] = x[j];
        pc = c;
      }
      if constexpr (POS_LDS) {
#pragma unroll
        for (int j = 0; j < MAXD; ++j)
          if (j < d) s_pos[b][j][i] = cur ? x[j] : pb[j];
      }
    }
    // The wavefront's winner as a key (32-bit DPP minimum of the cost order
    // and a ballot), then the workgroup's from the 16 slots.
    uint32_t om;
    const int wl = wave_argmin_first(cost_order(pc), act, om);
    if (lane == 0)
      s_red[b][wid] = wl < 0 ? KEY_NONE : ((uint64_t)om << 32) | (uint32_t)(wid * 64 + wl);
    __syncthreads();
    const uint64_t r = slots16_min(s_red[b]);
    const uint32_t wi = (uint32_t)r;
    if constexpr (POS_LDS) {
#pragma unroll
      for (int j = 0; j < MAXD; ++j) gb[j] = j < d ? s_pos[b][j][wi] : 0.0f;
    } else {
      if (act && (uint32_t)i == wi) {
#pragma unroll
        for (int j = 0; j < MAXD; ++j)
          if (j < d) s_gb[j] = cur ? x[j] : pb[j];
      }
      __syncthreads();
#pragma unroll
      for (int j = 0; j < MAXD; ++j) gb[j] = j < d ? s_gb[j] : 0.0f;
    }
    w = w * wdamp;
  }
  if (act) {
    store_vec<MAXD>(x, S.x, S.npad, i, d);
    store_vec<MAXD>(v, S.v, S.npad, i, d);
    store_vec<MAXD>(pb, S.pb, S.npad, i, d);
    S.pc[i] = pc;
  }
  // Hand-over state for whatever runs next (parity of t0 + T): the last
  // winner's position (held by every thread in gb) and its key.  The
  // reduction uses the slot set the last iteration did not.
  const uint64_t bk = block_min_u64_n(act ? make_key(pc, (uint32_t)i) : KEY_NONE,
                                      s_red[(T + 1) & 1], nw);
  const uint32_t par = (t0 + (uint32_t)T) & 1u;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int j = 0; j < MAXD; ++j)
      if (j < d) S.cand[(int64_t)par * d + j] = gb[j];
    S.bkey[par] = bk;
  }
}

// Whole sync loop of a swarm of at most 64 particles in one wavefront, one
// particle per lane held in registers: each iteration's winner (smallest
// best cost, then lowest index) comes from a DPP minimum and a ballot, and
// its position from the winner's lane, so nothing goes through LDS and no
// barrier is taken.  Hand-over state as k_sync_wg.
template <int MAXD, bool EXACT, int OBJ>
__global__ __launch_bounds__(64) void k_sync_wave(SwarmDev S, ProbArgs P, StepArgs A, float w,
                                                  float wdamp, uint32_t t0, int32_t T) {
  const int d = EXACT ? MAXD : P.d;
  const int64_t i = threadIdx.x;
  const bool act = i < S.n;
  float gb[MAXD];
#pragma unroll
  for (int j = 0; j < MAXD; ++j) gb[j] = j < d ? S.cand[(int64_t)(t0 & 1u) * d + j] : 0.0f;
  float x[MAXD], v[MAXD], pb[MAXD], pc = 0.0f;
#pragma unroll
  for (int j = 0; j < MAXD; ++j) x[j] = v[j] = pb[j] = 0.0f;
  if (act) {
    load_vec<MAXD>(x, S.x, S.npad, i, d);
    load_vec<MAXD>(v, S.v, S.npad, i, d);
    load_vec<MAXD>(pb, S.pb, S.npad, i, d);
    pc = S.pc[i];
  }
  const bool cur = A.gbest_from == PPS_GBEST_CURRENT;
  for (int32_t k = 1; k <= T; ++k) {
    if (act) {
      velocity_position<MAXD, OBJ>(x, v, pb, gb, w, A.c1, A.c2, (uint32_t)i, t0 + (uint32_t)k,
                                   S.epoch, S.key0, S.key1, d, P);
      const float c = cost_at<MAXD, EXACT, OBJ>(x, d, P, (float)(i + 2), pc);
      if (c < pc) {
#pragma unroll
        for (int j = 0; j < MAXD; ++j) pb[j] = x[j];
        pc = c;
      }
    }
    const int wi = wave_argmin_first(cost_order(pc), act);
#pragma unroll
    for (int j = 0; j < MAXD; ++j) gb[j] = j < d ? read_lane(cur ? x[j] : pb[j], wi) : 0.0f;
    w = w * wdamp;
  }
  if (act) {
    store_vec<MAXD>(x, S.x, S.npad, i, d);
    store_vec<MAXD>(v, S.v, S.npad, i, d);
    store_vec<MAXD>(pb, S.pb, S.npad, i, d);
    S.pc[i] = pc;
  }
  const uint64_t bk = wave_min_u64(act ? make_key(pc, (uint32_t)i) : KEY_NONE);
  const uint32_t par = (t0 + (uint32_t)T) & 1u;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int j = 0; j < MAXD; ++j)
      if (j < d) S.cand[(int64_t)par * d + j] = gb[j];
    S.bkey[par] = bk;
  }
}

// Async: reference semantics per particle (src/kernels.jl:205-221):
//   gbest = gbest_ref (at launch); for iter: step; if best_cost < gbest.cost:
//   gbest = (best_position, best_cost); w *= wdamp.
// Since best_cost never increases, after the first switch the running best
// is the particle's own best, so gbest = follow ? best : gbest_ref with
// follow = (best_cost < gbest_ref.cost) evaluated after each step.
template <int MAXD, bool EXACT, int OBJ, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_async(SwarmDev S, ProbArgs P, StepArgs A,
                                                 float w0, float wdamp, uint32_t t0,
                                                 int32_t T) {
  __shared__ uint64_t s_red[BLOCK / 64];
  __shared__ float s_g0[MAXD];
  const int d = EXACT ? MAXD : P.d;
  if ((int)threadIdx.x < d) s_g0[threadIdx.x] = S.gbrec->pos[threadIdx.x];
  const float g0c = S.gbrec->cost;
  __syncthreads();
  float g0[MAXD];
#pragma unroll
  for (int j = 0; j < MAXD; ++j) g0[j] = j < d ? s_g0[j] : 0.0f;
  uint64_t mykey = KEY_NONE;
  for (int64_t tile = blockIdx.x; tile * BLOCK < S.n; tile += gridDim.x) {
    const int64_t i = tile * BLOCK + threadIdx.x;
    if (i < S.n) {
      float x[MAXD], v[MAXD], pb[MAXD];
      load_vec<MAXD>(x, S.x, S.npad, i, d);
      load_vec<MAXD>(v, S.v, S.npad, i, d);
      load_vec<MAXD>(pb, S.pb, S.npad, i, d);
      float pc = S.pc[i];
      bool follow = false;
      float w = w0;
      for (int32_t k = 1; k <= T; ++k) {
        float g[MAXD];
#pragma unroll
        for (int j = 0; j < MAXD; ++j) g[j] = follow ? pb[j] : g0[j];
        velocity_position<MAXD, OBJ>(x, v, pb, g, w, A.c1, A.c2, (uint32_t)i, t0 + (uint32_t)k,
                                S.epoch, S.key0, S.key1, d, P);
        const float c = cost_at<MAXD, EXACT, OBJ>(x, d, P, (float)(k + 1), pc);
        if (c < pc) {
#pragma unroll
          for (int j = 0; j < MAXD; ++j) pb[j] = x[j];
          pc = c;
        }
        follow = follow || (pc < g0c);
        w = w * wdamp;
      }
      store_vec<MAXD>(x, S.x, S.npad, i, d);
      store_vec<MAXD>(v, S.v, S.npad, i, d);
      store_vec<MAXD>(pb, S.pb, S.npad, i, d);
      S.pc[i] = pc;
      mykey = umin64(mykey, make_key(pc, (uint32_t)i));
    }
  }
  publish_block<BLOCK>(S, (t0 + (uint32_t)T) & 1u, d, mykey, S.pb, s_red);
}

__device__ __forceinline__ float ld_acq(const float* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Queue-lock iteration (src/kernels.jl:38-125) with BLOCK = 1024 particles
// per workgroup, one tile per workgroup.  Each workgroup snapshots the live
// global best (position and cost) at its start, steps, counts the particles
// whose best cost beats the snapshot, and, if at least min_improvers did,
// publishes its best under a bounded spin lock when it still beats the live
// global best.  Workgroups race on the record as the reference's blocks do;
// which publications a block sees depends on the schedule either way.
template <int MAXD, bool EXACT, int OBJ, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_qlock(SwarmDev S, ProbArgs P, StepArgs A, float w,
                                                 uint32_t k) {
  const uint32_t t = S.tbase[0] + k;
  const uint32_t epoch = S.tbase[1];
  w = step_w(S, A, w, k);
  static_assert(BLOCK == 1024, "the slot reductions read 16 slots");
  __shared__ uint64_t s_red[BLOCK / 64];
  __shared__ int s_cnt[BLOCK / 64];
  __shared__ float s_gb[MAXD];
  __shared__ float s_gcost;
  const int d = EXACT ? MAXD : P.d;
  GbRec* R = S.gbrec;
  // The live global best as this block finds it at its start: position and
  // cost in one round trip (the publication re-reads the cost under the lock).
  if ((int)threadIdx.x < d) s_gb[threadIdx.x] = ld_acq(&R->pos[threadIdx.x]);
  if (threadIdx.x == 64) s_gcost = ld_acq(&R->cost);
  __syncthreads();
  const float gcost = s_gcost;
  float gb[MAXD];
#pragma unroll
  for (int j = 0; j < MAXD; ++j) gb[j] = j < d ? s_gb[j] : 0.0f;
  const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  const bool act = i < S.n;
  float pc = 0.0f;
  float pb[MAXD];
  if (act) {
    float x[MAXD], v[MAXD];
    load_vec<MAXD>(x, S.x, S.npad, i, d);
    load_vec<MAXD>(v, S.v, S.npad, i, d);
    load_vec<MAXD>(pb, S.pb, S.npad, i, d);
    pc = S.pc[i];
    velocity_position<MAXD, OBJ>(x, v, pb, gb, w, A.c1, A.c2, (uint32_t)i, t, epoch, S.key0,
                            S.key1, d, P);
    const float c = cost_at<MAXD, EXACT, OBJ>(x, d, P, (float)(i + 2), pc);
    store_vec<MAXD>(x, S.x, S.npad, i, d);
    store_vec<MAXD>(v, S.v, S.npad, i, d);
    // best positions in whole 64-byte segments, as k_sync_step
    const bool better = c < pc;
    if (better) {
#pragma unroll
      for (int j = 0; j < MAXD; ++j) pb[j] = x[j];
      pc = c;
    }
    const uint32_t seg = (uint32_t)(threadIdx.x & 63) & ~15u;
    if ((__ballot(better) >> seg) & 0xFFFFull) {
      store_vec<MAXD>(pb, S.pb, S.npad, i, d);
      S.pc[i] = pc;
    }
  }
  // count the particles whose best cost beats it and find the first lowest of
  // them: per wavefront a ballot and a 32-bit DPP minimum, then 16 slots
  const bool imp = act && (pc < gcost);
  uint32_t om;
  const int wl = wave_argmin_first(cost_order(pc), imp, om);
  const uint64_t bal = __ballot(imp);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
    s_red[wid] = wl < 0 ? KEY_NONE
                        : ((uint64_t)om << 32) | (uint32_t)((int64_t)blockIdx.x * BLOCK + wid * 64 + wl);
    s_cnt[wid] = (int)__popcll(bal);
  }
  __syncthreads();
  const uint64_t bk = slots16_min(s_red);
  const int count = slots16_sum(s_cnt);
  if (count > 0 && count >= A.min_improvers && imp && (uint32_t)i == (uint32_t)bk) {
    // Bounded spin: a holder is a running wave that releases within a few
    // hundred cycles; give up (and flag it) rather than wait forever.
    bool got = false;
    for (int spin = 0; spin < (1 << 22); ++spin) {
      if (atomicCAS(&R->lock, 0u, 1u) == 0u) {
        got = true;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    if (got) {
      __atomic_thread_fence(__ATOMIC_ACQUIRE);
      if (pc < ld_acq(&R->cost)) {
        for (int j = 0; j < d; ++j)
          __hip_atomic_store(&R->pos[j], pb[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&R->idx, (uint32_t)i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&R->cost, pc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      __hip_atomic_exchange(&R->lock, 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      atomicOr(&R->err, 1u);
    }
  }
}

// Queue-lock iterations t = tbase[0] + 1 .. tbase[0] + T of a swarm that fits
// one workgroup (n <= BLOCK): the reference's queue-lock kernel launched T
// times with its one block, as one launch of nw = ceil(n / 64) wavefronts.
// With one block nothing else publishes, so the live global best is the
// block's own and stays on chip between iterations (in every thread's
// registers below 16 coordinates, else in LDS, re-read each iteration: the
// 128-register budget of a 1024-thread workgroup); each iteration is
// k_qlock's: every particle steps against it, the particles whose best cost
// beats it are counted, and when at least min_improvers did, the first of the
// lowest of them is published.  One wavefront counts with a ballot and takes
// the winner's row from its lane (no barrier); more exchange per-wavefront
// counts and keys through LDS (two barriers: the second orders the winner's
// LDS writes before the reads and every read of this iteration's slots before
// the next iteration's writes).  The swarm stays in registers; state and
// record are written back at the end.
template <int MAXD, bool EXACT, int OBJ, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_qlock_wg(SwarmDev S, ProbArgs P, StepArgs A, float w,
                                                    float wdamp, int32_t T) {
  constexpr bool GB_REG = MAXD < 16;
  static_assert(BLOCK == 1024, "the slot reductions read 16 slots");
  __shared__ uint64_t s_red[BLOCK / 64];
  __shared__ int s_cnt[BLOCK / 64];
  __shared__ float s_gb[MAXD];
  __shared__ float s_gcost;
  const uint32_t t0 = S.tbase[0];
  const uint32_t epoch = S.tbase[1];
  const int d = EXACT ? MAXD : P.d;
  const int nw = (int)(blockDim.x >> 6);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  GbRec* R = S.gbrec;
  float gb[MAXD];
  if constexpr (GB_REG) {
#pragma unroll
    for (int j = 0; j < MAXD; ++j) gb[j] = j < d ? R->pos[j] : 0.0f;
  } else {
    for (int j = threadIdx.x; j < d; j += blockDim.x) s_gb[j] = R->pos[j];
    __syncthreads();
  }
  float gcost = R->cost;
  uint32_t gidx = R->idx;
  const int64_t i = threadIdx.x;
  const bool act = i < S.n;
  float x[MAXD], v[MAXD], pb[MAXD], pc = 0.0f;
#pragma unroll
  for (int j = 0; j < MAXD; ++j) x[j] = v[j] = pb[j] = 0.0f;
  if (act) {
    load_vec<MAXD>(x, S.x, S.npad, i, d);
    load_vec<MAXD>(v, S.v, S.npad, i, d);
    load_vec<MAXD>(pb, S.pb, S.npad, i, d);
    pc = S.pc[i];
  }
  // Slots of wavefronts not launched hold KEY_NONE and 0, so every iteration
  // reads all slots in one unrolled round of LDS loads.
  if (nw > 1) {
    if ((int)threadIdx.x < BLOCK / 64 && (int)threadIdx.x >= nw) {
      s_red[threadIdx.x] = KEY_NONE;
      s_cnt[threadIdx.x] = 0;
    }
    __syncthreads();
  }
  for (int32_t k = 1; k <= T; ++k) {
    const uint32_t t = t0 + (uint32_t)k;
    if constexpr (!GB_REG) {
#pragma unroll
      for (int j = 0; j < MAXD; ++j) gb[j] = j < d ? s_gb[j] : 0.0f;
    }
    if (act) {
      velocity_position<MAXD, OBJ>(x, v, pb, gb, w, A.c1, A.c2, (uint32_t)i, t, epoch, S.key0,
                                   S.key1, d, P);
      const float c = cost_at<MAXD, EXACT, OBJ>(x, d, P, (float)(i + 2), pc);
      if (c < pc) {
#pragma unroll
        for (int j = 0; j < MAXD; ++j) pb[j] = x[j];
        pc = c;
      }
    }
    w = w * wdamp;
    const bool imp = act && (pc < gcost);
    const uint64_t bal = __ballot(imp);
    // one wavefront: count with a ballot, take the winner from its lane
    if (nw == 1) {
      const int count = (int)__popcll(bal);
      if (count > 0 && count >= A.min_improvers) {
        const int wi = wave_argmin_first(cost_order(pc), imp);
        if constexpr (GB_REG) {
#pragma unroll
          for (int j = 0; j < MAXD; ++j) gb[j] = j < d ? read_lane(pb[j], wi) : 0.0f;
        } else if (lane == wi) {
          // this wavefront's next reads of s_gb follow in program order
#pragma unroll
          for (int j = 0; j < MAXD; ++j)
            if (j < d) s_gb[j] = pb[j];
        }
        gcost = read_lane(pc, wi);
        gidx = (uint32_t)wi;
      }
      continue;
    }
    // each wavefront's count and first lowest improver, then the
    // workgroup's from the 16 slots
    uint32_t om;
    const int wl = wave_argmin_first(cost_order(pc), imp, om);
    if (lane == 0) {
      s_red[wid] = wl < 0 ? KEY_NONE : ((uint64_t)om << 32) | (uint32_t)(wid * 64 + wl);
      s_cnt[wid] = (int)__popcll(bal);
    }
    __syncthreads();
    const uint64_t r = slots16_min(s_red);
    const int count = slots16_sum(s_cnt);
    const bool pub = count > 0 && count >= A.min_improvers;
    if (pub && (uint32_t)i == (uint32_t)r) {
#pragma unroll
      for (int j = 0; j < MAXD; ++j)
        if (j < d) s_gb[j] = pb[j];
      s_gcost = pc;
    }
    __syncthreads();
    if (pub) {
      if constexpr (GB_REG) {
#pragma unroll
        for (int j = 0; j < MAXD; ++j) gb[j] = j < d ? s_gb[j] : 0.0f;
      }
      gcost = s_gcost;
      gidx = (uint32_t)r;
    }
  }
  if (act) {
    store_vec<MAXD>(x, S.x, S.npad, i, d);
    store_vec<MAXD>(v, S.v, S.npad, i, d);
    store_vec<MAXD>(pb, S.pb, S.npad, i, d);
    S.pc[i] = pc;
  }
  if (threadIdx.x == 0) {
#pragma unroll
    for (int j = 0; j < MAXD; ++j)
      if (j < d) R->pos[j] = GB_REG ? gb[j] : s_gb[j];
    R->cost = gcost;
    R->idx = gidx;
  }
}

// ---------------------------------------------------------------------------
// Launchers.
constexpr int QBLK = 1024; // queue-lock block (the reference's workgroup size)
constexpr int WGB = 1024;  // single-workgroup sync loop
// Wavefronts launched for a one-workgroup swarm of n particles.
inline unsigned waves_of(int64_t n) { return (unsigned)((n + 63) / 64); }

template <int MAXD, bool EXACT, int OBJ>
struct Launch {
  static constexpr int BLK = block_of(OBJ);  // workgroup of the multi-workgroup kernels
  // Also compiled for workgroups of 512 and (up to 16 coordinates; at 32 a
  // 1024-thread workgroup's 128 registers per thread spill) 1024 particles,
  // the autotuner's choices: the streaming objectives at their exact
  // dimensions.
  static constexpr bool MULTI = EXACT && (OBJ == PPS_OBJ_ROSENBROCK ||
                                          OBJ == PPS_OBJ_RASTRIGIN || OBJ == PPS_OBJ_SPHERE);
  static constexpr int MAX_BLOCK = !MULTI ? BLK : (MAXD <= 16 ? 1024 : 512);
  template <int B>
  static hipError_t init_b(const SwarmDev& S, const ProbArgs& P, uint32_t par, int grid,
                           hipStream_t st) {
    hipLaunchKernelGGL((k_init<MAXD, EXACT, OBJ, B>), dim3(grid), dim3(B), 0, st, S, P, par);
    return hipGetLastError();
  }
  static hipError_t init(const SwarmDev& S, const ProbArgs& P, uint32_t par, hipStream_t st) {
    if (S.block == WGB && S.G == 1) return init_b<WGB>(S, P, par, 1, st);
    if constexpr (MULTI) {
      if (S.block == 512) return init_b<512>(S, P, par, S.G, st);
      if constexpr (MAX_BLOCK >= 1024)
        if (S.block == 1024) return init_b<1024>(S, P, par, S.G, st);
    }
    return init_b<BLK>(S, P, par, S.G, st);
  }
  template <int B>
  static hipError_t rekey_b(const SwarmDev& S, const ProbArgs& P, uint32_t par, int32_t gf,
                            int grid, hipStream_t st) {
    hipLaunchKernelGGL((k_rekey<MAXD, EXACT, OBJ, B>), dim3(grid), dim3(B), 0, st, S, P, par, gf);
    return hipGetLastError();
  }
  static hipError_t rekey(const SwarmDev& S, const ProbArgs& P, uint32_t par, int32_t gf,
                          hipStream_t st) {
    if (S.block == WGB && S.G == 1) return rekey_b<WGB>(S, P, par, gf, 1, st);
    if constexpr (MULTI) {
      if (S.block == 512) return rekey_b<512>(S, P, par, gf, S.G, st);
      if constexpr (MAX_BLOCK >= 1024)
        if (S.block == 1024) return rekey_b<1024>(S, P, par, gf, S.G, st);
    }
    return rekey_b<BLK>(S, P, par, gf, S.G, st);
  }
  // Multi-workgroup step with B-particle workgroups (64-bit row offsets only
  // in the objective's default size).
  template <int B>
  static hipError_t step_b(const SwarmDev& S, const ProbArgs& P, const StepArgs& A, float w,
                           uint32_t k, bool off32, bool nt, hipStream_t st) {
    if (off32 && !nt)
      hipLaunchKernelGGL((k_sync_step<MAXD, EXACT, OBJ, B, true, false>), dim3(S.G), dim3(B), 0,
                         st, S, P, A, w, k);
    else if (off32)
      hipLaunchKernelGGL((k_sync_step<MAXD, EXACT, OBJ, B, true, true>), dim3(S.G), dim3(B), 0,
                         st, S, P, A, w, k);
    else if constexpr (B == BLK)
      hipLaunchKernelGGL((k_sync_step<MAXD, EXACT, OBJ, B, false, true>), dim3(S.G), dim3(B), 0,
                         st, S, P, A, w, k);
    else
      return hipErrorInvalidValue;
    return hipGetLastError();
  }
  static hipError_t sync_step(const SwarmDev& S, const ProbArgs& P, const StepArgs& A, float w,
                              uint32_t k, hipStream_t st) {
    // 32-bit row offsets while every state array is below 4 GiB; nontemporal
    // streams once the whole state (3 rows of d plus the best costs) is past
    // the cache (always so when an array exceeds 4 GiB)
    const uint64_t d = (uint64_t)(EXACT ? MAXD : P.d);
    const bool off32 = (uint64_t)S.npad * d * 4u < (1ull << 32);
    const bool nt = (uint64_t)S.npad * (3u * d + 1u) * 4u > PPS_NT_MIN_STATE_BYTES;
    if constexpr (is_ode_objective(OBJ)) {
      if (!(S.block == WGB && S.G == 1) && S.block == 64) {
        if (off32)
          hipLaunchKernelGGL((k_sync_step_pair<OBJ, true>), dim3(S.G), dim3(128), 0, st, S, P,
                             A, w, k);
        else
          hipLaunchKernelGGL((k_sync_step_pair<OBJ, false>), dim3(S.G), dim3(128), 0, st, S,
                             P, A, w, k);
        return hipGetLastError();
      }
    }
    if (S.block == WGB && S.G == 1) {
      hipLaunchKernelGGL((k_sync_step<MAXD, EXACT, OBJ, WGB, true, false>), dim3(1), dim3(WGB), 0,
                         st, S, P, A, w, k);
      return hipGetLastError();
    }
    if constexpr (MULTI) {
      if (S.block == 512) return step_b<512>(S, P, A, w, k, off32, nt, st);
      if constexpr (MAX_BLOCK >= 1024)
        if (S.block == 1024) return step_b<1024>(S, P, A, w, k, off32, nt, st);
    }
    return step_b<BLK>(S, P, A, w, k, off32, nt, st);
  }
  static hipError_t sync_wg(const SwarmDev& S, const ProbArgs& P, const StepArgs& A, float w0,
                            float wdamp, uint32_t t0, int32_t T, hipStream_t st) {
    if (!(S.G == 1 && S.block == WGB && S.n <= WGB)) return hipErrorInvalidValue;
    if (S.n <= 64)
      hipLaunchKernelGGL((k_sync_wave<MAXD, EXACT, OBJ>), dim3(1), dim3(64), 0, st, S, P, A, w0,
                         wdamp, t0, T);
    else
      hipLaunchKernelGGL((k_sync_wg<MAXD, EXACT, OBJ, WGB>), dim3(1), dim3(waves_of(S.n) * 64),
                         0, st, S, P, A, w0, wdamp, t0, T);
    return hipGetLastError();
  }
  static hipError_t async_run(const SwarmDev& S, const ProbArgs& P, const StepArgs& A, float w0,
                              float wdamp, uint32_t t0, int32_t T, hipStream_t st) {
    if (S.block == WGB && S.G == 1)
      hipLaunchKernelGGL((k_async<MAXD, EXACT, OBJ, WGB>), dim3(1), dim3(WGB), 0, st, S, P, A, w0,
                         wdamp, t0, T);
    else
      hipLaunchKernelGGL((k_async<MAXD, EXACT, OBJ, BLK>), dim3(S.G), dim3(BLK), 0, st, S, P, A, w0,
                         wdamp, t0, T);
    return hipGetLastError();
  }
  static hipError_t qlock_step(const SwarmDev& S, const ProbArgs& P, const StepArgs& A, float w,
                               uint32_t k, hipStream_t st) {
    const int64_t nb = (S.n + QBLK - 1) / QBLK;
    hipLaunchKernelGGL((k_qlock<MAXD, EXACT, OBJ, QBLK>), dim3((unsigned)nb), dim3(QBLK), 0, st, S,
                       P, A, w, k);
    return hipGetLastError();
  }
  static hipError_t qlock_wg(const SwarmDev& S, const ProbArgs& P, const StepArgs& A, float w0,
                             float wdamp, int32_t T, hipStream_t st) {
    if (!(S.n <= QBLK)) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_qlock_wg<MAXD, EXACT, OBJ, QBLK>), dim3(1), dim3(waves_of(S.n) * 64), 0,
                       st, S, P, A, w0, wdamp, T);
    return hipGetLastError();
  }
  static const KernelSet* set() {
    static const KernelSet ks = {MAXD,       MAX_BLOCK,  &init,      &rekey,    &sync_step,
                                 &sync_wg,   &async_run, &qlock_step, &qlock_wg};
    return &ks;
  }
};

// (objective, dimension) -> compiled kernel set.  Dimensions used by the
// benchmark shapes get exact-size instantiations; others, and every
// constrained problem, run in the next power-of-two bucket with a runtime
// dimension.
template <int OBJ>
const KernelSet* generic_set(int d, bool constrained) {
  if (!constrained) switch (d) {
    case 1: return Launch<1, true, OBJ>::set();
    case 2: return Launch<2, true, OBJ>::set();
    case 3: return Launch<3, true, OBJ>::set();
    case 4: return Launch<4, true, OBJ>::set();
    case 8: return Launch<8, true, OBJ>::set();
    case 10: return Launch<10, true, OBJ>::set();
    case 16: return Launch<16, true, OBJ>::set();
    case 32: return Launch<32, true, OBJ>::set();
    default: break;
  }
  if (d <= 8) return Launch<8, false, OBJ>::set();
  if (d <= 16) return Launch<16, false, OBJ>::set();
  if (d <= 32) return Launch<32, false, OBJ>::set();
  if (d <= 64) return Launch<64, false, OBJ>::set();
  return nullptr;
}

}  // namespace pps
