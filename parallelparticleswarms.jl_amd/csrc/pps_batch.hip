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
// chunking.
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

template <int D>
__device__ __forceinline__ void ld(float (&r)[D], const float* b, int64_t np, int64_t i) {
#pragma unroll
  for (int j = 0; j < D; ++j) r[j] = b[(int64_t)j * np + i];
}
template <int D>
__device__ __forceinline__ void st(const float (&r)[D], float* b, int64_t np, int64_t i) {
#pragma unroll
  for (int j = 0; j < D; ++j) b[(int64_t)j * np + i] = r[j];
}

// Run up to `chunk` iterations of one problem in this workgroup.
template <int D, int OBJ, int BLOCK>
__device__ void run_problem(BatchProb* B, int32_t chunk, uint64_t* s_red, float* s_gb) {
  const ProbArgs& P = B->P;
  const int64_t n = B->n, np = B->npad;
  const uint32_t k0 = B->key0, k1 = B->key1;
  float* X = B->x;
  float* V = B->v;
  float* PB = B->pb;
  float* PC = B->pc;
  if (!B->initialized) {
    // uniform init (as k_init) and the initial global best
    uint64_t mykey = KEY_NONE;
    for (int64_t i = threadIdx.x; i < n; i += BLOCK) {
      float x[D], z[D];
      init_position<D>(x, (uint32_t)i, 0u, k0, k1, D, P);
#pragma unroll
      for (int j = 0; j < D; ++j) z[j] = 0.0f;
      const float c = objective<D, OBJ>(x, D, P);
      st<D>(x, X, np, i);
      st<D>(z, V, np, i);
      st<D>(x, PB, np, i);
      PC[i] = c;
      mykey = umin64(mykey, make_key(c, (uint32_t)i));
    }
    const uint64_t bk = block_min<BLOCK>(mykey, s_red);
    if (mykey == bk) {
      const int64_t wi = (int64_t)(uint32_t)bk;
#pragma unroll
      for (int j = 0; j < D; ++j) s_gb[j] = PB[(int64_t)j * np + wi];
    }
    __syncthreads();
    if (threadIdx.x < D) B->gb[threadIdx.x] = s_gb[threadIdx.x];
    if (threadIdx.x == 0) {
      B->gb_cost = key_cost(bk);
      B->gb_idx = (uint32_t)bk;
      B->initialized = 1;
    }
    __syncthreads();
  }
  const int32_t steps = chunk < B->remaining ? chunk : B->remaining;
  if (steps <= 0) return;
  if (threadIdx.x < D) s_gb[threadIdx.x] = B->gb[threadIdx.x];
  const uint32_t t0 = B->iter;
  float w = B->w;
  const float wdamp = B->wdamp, c1 = B->c1, c2 = B->c2;
  __syncthreads();
  uint64_t bk = KEY_NONE;
  if (n <= BLOCK) {
    // swarm in registers for the whole chunk
    const int64_t i = threadIdx.x;
    const bool act = i < n;
    float x[D], v[D], pb[D], pc = 0.0f;
#pragma unroll
    for (int j = 0; j < D; ++j) x[j] = v[j] = pb[j] = 0.0f;
    if (act) {
      ld<D>(x, X, np, i);
      ld<D>(v, V, np, i);
      ld<D>(pb, PB, np, i);
      pc = PC[i];
    }
    for (int32_t k = 1; k <= steps; ++k) {
      float gb[D];
#pragma unroll
      for (int j = 0; j < D; ++j) gb[j] = s_gb[j];
      uint64_t mykey = KEY_NONE;
      if (act) {
        velocity_position<D>(x, v, pb, gb, w, c1, c2, (uint32_t)i, t0 + (uint32_t)k, 0u, k0, k1,
                             D, P);
        const float c = objective<D, OBJ>(x, D, P);
        if (c < pc) {
#pragma unroll
          for (int j = 0; j < D; ++j) pb[j] = x[j];
          pc = c;
        }
        mykey = make_key(pc, (uint32_t)i);
      }
      bk = block_min<BLOCK>(mykey, s_red);
      if (act && mykey == bk) {
#pragma unroll
        for (int j = 0; j < D; ++j) s_gb[j] = pb[j];
      }
      __syncthreads();
      w = w * wdamp;
    }
    if (act) {
      st<D>(x, X, np, i);
      st<D>(v, V, np, i);
      st<D>(pb, PB, np, i);
      PC[i] = pc;
    }
  } else {
    // swarm streamed from the arena (L2/MALL resident for these sizes)
    for (int32_t k = 1; k <= steps; ++k) {
      float gb[D];
#pragma unroll
      for (int j = 0; j < D; ++j) gb[j] = s_gb[j];
      uint64_t mykey = KEY_NONE;
      for (int64_t i = threadIdx.x; i < n; i += BLOCK) {
        float x[D], v[D], pb[D];
        ld<D>(x, X, np, i);
        ld<D>(v, V, np, i);
        ld<D>(pb, PB, np, i);
        float pc = PC[i];
        velocity_position<D>(x, v, pb, gb, w, c1, c2, (uint32_t)i, t0 + (uint32_t)k, 0u, k0, k1,
                             D, P);
        const float c = objective<D, OBJ>(x, D, P);
        st<D>(x, X, np, i);
        st<D>(v, V, np, i);
        if (c < pc) {
          st<D>(x, PB, np, i);
          pc = c;
          PC[i] = c;
        }
        mykey = umin64(mykey, make_key(pc, (uint32_t)i));
      }
      bk = block_min<BLOCK>(mykey, s_red);
      if (mykey == bk) {
        const int64_t wi = (int64_t)(uint32_t)bk;
#pragma unroll
        for (int j = 0; j < D; ++j) s_gb[j] = PB[(int64_t)j * np + wi];
      }
      __syncthreads();
      w = w * wdamp;
    }
  }
  if (threadIdx.x < D) B->gb[threadIdx.x] = s_gb[threadIdx.x];
  if (threadIdx.x == 0) {
    B->gb_cost = key_cost(bk);
    B->gb_idx = (uint32_t)bk;
    B->iter = t0 + (uint32_t)steps;
    B->remaining -= steps;
    B->w = w;
  }
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_batch(BatchProb* probs, const int32_t* ids,
                                                 int32_t chunk, BatchOut* out) {
  __shared__ uint64_t s_red[BLOCK / 64];
  __shared__ float s_gb[PPS_BMAXD];
  const int32_t id = ids[blockIdx.x];
  BatchProb* B = probs + id;
  const int sel = B->P.obj * 64 + B->P.d;
  switch (sel) {
#define PPS_CASE(OBJ, D) \
  case OBJ * 64 + D: run_problem<D, OBJ, BLOCK>(B, chunk, s_red, s_gb); break;
    PPS_CASE(PPS_OBJ_ROSENBROCK, 2)
    PPS_CASE(PPS_OBJ_ROSENBROCK, 4)
    PPS_CASE(PPS_OBJ_ROSENBROCK, 8)
    PPS_CASE(PPS_OBJ_ROSENBROCK, 16)
    PPS_CASE(PPS_OBJ_ROSENBROCK, 32)
    PPS_CASE(PPS_OBJ_RASTRIGIN, 2)
    PPS_CASE(PPS_OBJ_RASTRIGIN, 4)
    PPS_CASE(PPS_OBJ_RASTRIGIN, 8)
    PPS_CASE(PPS_OBJ_RASTRIGIN, 16)
    PPS_CASE(PPS_OBJ_RASTRIGIN, 32)
#undef PPS_CASE
    default: break;
  }
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
  return d == 2 || d == 4 || d == 8 || d == 16 || d == 32;
}

hipError_t launch_batch(BatchProb* probs, const int32_t* ids, int32_t count, int32_t chunk,
                        BatchOut* out, hipStream_t st) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_batch<BATCH_BLOCK>, dim3(count), dim3(BATCH_BLOCK), 0, st, probs, ids,
                     chunk, out);
  return hipGetLastError();
}

}  // namespace pps
