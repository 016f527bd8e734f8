// pps_batch.hip — packed launches of many independent sync-PSO problems.
//
// Each admitted problem has a device descriptor (BatchProb): its objective,
// bounds, Philox key, iteration counter, damped inertia, the attractor of its
// next iteration, and pointers to its SoA state in a shared arena.  One launch
// ("chunk") gives every listed problem one workgroup, which runs up to `chunk`
// iterations of that problem with the workgroup barrier as the per-iteration
// synchronisation; the winner's position is handed over through LDS.  A
// problem's result depends only on its own seed, particle indices and
// iteration numbers, so it is the same whether it runs alone (the per-problem
// launch loop of pps_swarm_run) or packed with any other problems, in any
// chunking.  A launch is time-sliced: each workgroup stops its problem after
// the first iteration that ends past a device-clock deadline, so the
// workgroups of one launch finish together whatever their problem sizes.
#include <hip/hip_runtime.h>

#include "pps_batch.h"
#include "pps_device.h"

namespace pps {

namespace {

constexpr uint64_t KEY_NONE = ~0ull;

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

// Run iterations of one problem in this workgroup until it is finished, or
// `chunk` iterations are done, or the device clock passes `deadline` (checked
// by thread 0 once per iteration and broadcast through LDS, so the whole
// workgroup stops after the same iteration).
//
// The swarm is streamed one coordinate pair at a time (one Philox call per
// pair, as in velocity_position()), so register use does not grow with d; a
// particle whose cost improves re-reads its new position to store its best.
template <int OBJ, int BLOCK, int PPT>
__device__ void run_problem(BatchProb* B, int32_t chunk, uint64_t deadline, uint64_t* s_red,
                            ProbLds* L, int* s_stop) {
  const int d = B->P.d;
  const int64_t n = B->n, np = B->npad;
  const uint32_t k0 = B->key0, k1 = B->key1;
  const float p1 = B->P.par[0], p2 = B->P.par[1];
  const bool has_shift = B->P.has_shift != 0;
  float* __restrict__ X = B->x;
  float* __restrict__ V = B->v;
  float* __restrict__ PB = B->pb;
  float* __restrict__ PC = B->pc;
  if (threadIdx.x < d) {
    L->lb[threadIdx.x] = B->P.lb[threadIdx.x];
    L->ub[threadIdx.x] = B->P.ub[threadIdx.x];
    L->shift[threadIdx.x] = has_shift ? B->P.shift[threadIdx.x] : 0.0f;
  }
  __syncthreads();
  if (!B->initialized) {
    // uniform init (as init_position + objective) and the initial global best
    uint64_t mykey = KEY_NONE;
    for (int64_t i = threadIdx.x; i < n; i += BLOCK) {
      Accum<OBJ> acc;
      acc.start(d);
      u4 r{};
      for (int j = 0; j < d; ++j) {
        if ((j & 3) == 0)
          r = philox4x32_10(u4{(uint32_t)i, 0u, (uint32_t)(j >> 2) | (PPS_TAG_INIT << 16), 0u},
                            k0, k1);
        const uint32_t rr = (j & 3) == 0 ? r.x : (j & 3) == 1 ? r.y : (j & 3) == 2 ? r.z : r.w;
        const float x = u01(rr) * (L->ub[j] - L->lb[j]) + L->lb[j];
        X[(int64_t)j * np + i] = x;
        V[(int64_t)j * np + i] = 0.0f;
        PB[(int64_t)j * np + i] = x;
        acc.add(j, has_shift ? x - L->shift[j] : x, p1, p2);
      }
      PC[i] = acc.res;
      mykey = umin64(mykey, make_key(acc.res, (uint32_t)i));
    }
    const uint64_t bk = block_min<BLOCK>(mykey, s_red);
    if (mykey == bk) {
      const int64_t wi = (int64_t)(uint32_t)bk;
      for (int j = 0; j < d; ++j) L->gb[j] = PB[(int64_t)j * np + wi];
    }
    __syncthreads();
    if (threadIdx.x < d) B->gb[threadIdx.x] = L->gb[threadIdx.x];
    if (threadIdx.x == 0) {
      B->gb_cost = key_cost(bk);
      B->gb_idx = (uint32_t)bk;
      B->initialized = 1;
    }
    __syncthreads();
  }
  const int32_t steps = chunk < B->remaining ? chunk : B->remaining;
  if (steps <= 0) return;
  if (threadIdx.x < d) L->gb[threadIdx.x] = B->gb[threadIdx.x];
  const uint32_t t0 = B->iter;
  float w = B->w;
  const float wdamp = B->wdamp, c1 = B->c1, c2 = B->c2;
  if (threadIdx.x == 0) *s_stop = 0;
  __syncthreads();
  uint64_t bk = KEY_NONE;
  int32_t done = 0;
  while (done < steps) {
    const uint32_t t = t0 + (uint32_t)(++done);
    uint64_t mykey = KEY_NONE;
    // PPT particles per thread in flight (i, i + BLOCK, ...): their loads of
    // one coordinate pair are issued together, which hides memory latency
    // when a workgroup streams a large swarm.
    for (int64_t i0 = threadIdx.x; i0 < n; i0 += (int64_t)PPT * BLOCK) {
      Accum<OBJ> acc[PPT];
#pragma unroll
      for (int q = 0; q < PPT; ++q) acc[q].start(d);
      for (int k = 0; 2 * k < d; ++k) {
        float x[PPT][2], v[PPT][2], pbv[PPT][2];
#pragma unroll
        for (int q = 0; q < PPT; ++q) {
          const int64_t i = i0 + (int64_t)q * BLOCK;
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            const int j = 2 * k + h;
            const int64_t o = (int64_t)j * np + i;
            const bool ok = i < n && j < d;
            x[q][h] = ok ? X[o] : 0.0f;
            v[q][h] = ok ? V[o] : 0.0f;
            pbv[q][h] = ok ? PB[o] : 0.0f;
          }
        }
#pragma unroll
        for (int q = 0; q < PPT; ++q) {
          const int64_t i = i0 + (int64_t)q * BLOCK;
          if (i >= n) continue;
          const u4 r = philox4x32_10(u4{(uint32_t)i, t, (uint32_t)k | (PPS_TAG_STEP << 16), 0u}, k0, k1);
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            const int j = 2 * k + h;
            if (j < d) {
              const float r1 = u01(h == 0 ? r.x : r.z);
              const float r2 = u01(h == 0 ? r.y : r.w);
              const float a = w * v[q][h];
              const float b = (c1 * r1) * (pbv[q][h] - x[q][h]);
              const float c = (c2 * r2) * (L->gb[j] - x[q][h]);
              const float vn = (a + b) + c;
              float xn = x[q][h] + vn;
              xn = jl_min(jl_max(xn, L->lb[j]), L->ub[j]);
              const int64_t o = (int64_t)j * np + i;
              X[o] = xn;
              V[o] = vn;
              acc[q].add(j, has_shift ? xn - L->shift[j] : xn, p1, p2);
            }
          }
        }
      }
#pragma unroll
      for (int q = 0; q < PPT; ++q) {
        const int64_t i = i0 + (int64_t)q * BLOCK;
        if (i >= n) continue;
        float pc = PC[i];
        if (acc[q].res < pc) {
          for (int j = 0; j < d; ++j) PB[(int64_t)j * np + i] = X[(int64_t)j * np + i];
          pc = acc[q].res;
          PC[i] = pc;
        }
        mykey = umin64(mykey, make_key(pc, (uint32_t)i));
      }
    }
    bk = block_min<BLOCK>(mykey, s_red);
    if (mykey == bk) {
      const int64_t wi = (int64_t)(uint32_t)bk;
      for (int j = 0; j < d; ++j) L->gb[j] = PB[(int64_t)j * np + wi];
    }
    if (threadIdx.x == 0) *s_stop = wall_clock64() > deadline;
    __syncthreads();
    w = w * wdamp;
    if (*s_stop) break;
  }
  if (threadIdx.x < d) B->gb[threadIdx.x] = L->gb[threadIdx.x];
  if (threadIdx.x == 0) {
    B->gb_cost = key_cost(bk);
    B->gb_idx = (uint32_t)bk;
    B->iter = t0 + (uint32_t)done;
    B->remaining -= done;
    B->w = w;
  }
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_batch(BatchProb* probs, const BatchProb* fresh,
                                                 const uint32_t* fresh_in, uint32_t launch_no,
                                                 const int32_t* ids, int32_t chunk,
                                                 uint64_t budget_ticks, BatchOut* out) {
  __shared__ uint64_t s_red[BLOCK / 64];
  __shared__ ProbLds s_lds;
  __shared__ int s_stop;
  const uint64_t deadline = wall_clock64() + budget_ticks;
  const int32_t id = ids[blockIdx.x];
  BatchProb* B = probs + id;
  if (fresh && fresh_in[id] == launch_no) {
    // newly admitted: pull the descriptor from host-visible memory
    const uint32_t* src = reinterpret_cast<const uint32_t*>(fresh + id);
    uint32_t* dst = reinterpret_cast<uint32_t*>(B);
    for (int k = threadIdx.x; k < (int)(sizeof(BatchProb) / 4); k += BLOCK) dst[k] = src[k];
    __syncthreads();
  }
  if (B->P.obj == PPS_OBJ_RASTRIGIN)
    run_problem<PPS_OBJ_RASTRIGIN, BLOCK, 4>(B, chunk, deadline, s_red, &s_lds, &s_stop);
  else
    run_problem<PPS_OBJ_ROSENBROCK, BLOCK, 4>(B, chunk, deadline, s_red, &s_lds, &s_stop);
  if (out) {
    __syncthreads();
    if (threadIdx.x == 0) {
      // visible to the host after the state writes of this workgroup
      __threadfence();
      if (B->remaining <= 0) {
        BatchOut* o = out + id;
        o->gb_cost = B->gb_cost;
        o->gb_idx = B->gb_idx;
        o->iter = B->iter;
        o->remaining = B->remaining;
        for (int j = 0; j < B->P.d && j < PPS_BMAXD; ++j) o->gb[j] = B->gb[j];
        __threadfence_system();
        __hip_atomic_store(&o->done, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
      }
    }
  }
}

}  // namespace

bool batch_supported(int obj, int d) {
  if (obj != PPS_OBJ_ROSENBROCK && obj != PPS_OBJ_RASTRIGIN) return false;
  return d >= 1 && d <= PPS_BMAXD;
}

hipError_t launch_batch(BatchProb* probs, const BatchProb* fresh, const uint32_t* fresh_in,
                        uint32_t launch_no, const int32_t* ids, int32_t count, int32_t chunk,
                        uint64_t budget_ticks, BatchOut* out, int block, hipStream_t st) {
  if (count <= 0) return hipSuccess;
  if (block == BATCH_BLOCK_LARGE)
    hipLaunchKernelGGL(k_batch<BATCH_BLOCK_LARGE>, dim3(count), dim3(BATCH_BLOCK_LARGE), 0, st,
                       probs, fresh, fresh_in, launch_no, ids, chunk, budget_ticks, out);
  else
    hipLaunchKernelGGL(k_batch<BATCH_BLOCK>, dim3(count), dim3(BATCH_BLOCK), 0, st, probs, fresh,
                       fresh_in, launch_no, ids, chunk, budget_ticks, out);
  return hipGetLastError();
}

}  // namespace pps
