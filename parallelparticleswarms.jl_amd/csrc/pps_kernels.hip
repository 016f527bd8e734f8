// pps_kernels.hip — objective-independent kernels (grid pass, global-best
// writes) and the (objective, dimension) -> kernel-set dispatch.
#include <hip/hip_runtime.h>

#include "pps_device.h"
#include "pps_kernels.h"

namespace pps {

constexpr uint64_t KEY_NONE_ = ~0ull;

template <int BLOCK>
__device__ __forceinline__ uint64_t block_min_u64_(uint64_t k, uint64_t* s_red) {
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

// Grid pass: minimum of the G block keys of parity `par` -> gbrec.
__global__ __launch_bounds__(256) void k_finalize(SwarmDev S, int d, uint32_t par) {
  __shared__ uint64_t s_red[4];
  uint64_t k = KEY_NONE_;
  for (int j = threadIdx.x; j < S.G; j += 256) k = umin64(k, S.bkey[(int64_t)par * S.G + j]);
  k = block_min_u64_<256>(k, s_red);
  const uint32_t gidx = (uint32_t)k;
  const int gblk = (int)((gidx / (uint32_t)S.block) % (uint32_t)S.G);
  if ((int)threadIdx.x < d)
    S.gbrec->pos[threadIdx.x] = S.cand[((int64_t)par * S.G + gblk) * d + threadIdx.x];
  if (threadIdx.x == 0) {
    S.gbrec->cost = key_cost(k);
    S.gbrec->idx = gidx;
  }
}

__global__ __launch_bounds__(64) void k_set_gbest(SwarmDev S, int d, uint32_t par,
                                                  const float* pos, float cost) {
  const int tid = threadIdx.x;
  if (tid < d) {
    S.gbrec->pos[tid] = pos[tid];
    S.cand[(int64_t)par * S.G * d + tid] = pos[tid];  // candidate of block 0
  }
  for (int b = tid; b < S.G; b += 64) S.bkey[(int64_t)par * S.G + b] = b == 0 ? make_key(cost, 0u) : KEY_NONE_;
  if (tid == 0) {
    S.gbrec->cost = cost;
    S.gbrec->idx = 0u;
  }
}

hipError_t launch_finalize(const SwarmDev& S, int d, uint32_t par, hipStream_t st) {
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, st, S, d, par);
  return hipGetLastError();
}

__global__ void k_set_tbase(uint32_t* p, uint32_t t0, uint32_t epoch) {
  if (threadIdx.x == 0) {
    p[0] = t0;
    p[1] = epoch;
    p[2] = 0u;
  }
}

hipError_t launch_set_tbase(uint32_t* tbase, uint32_t t0, uint32_t epoch, hipStream_t st) {
  hipLaunchKernelGGL(k_set_tbase, dim3(1), dim3(64), 0, st, tbase, t0, epoch);
  return hipGetLastError();
}

__global__ void k_advance_tbase(uint32_t* p, uint32_t K) {
  if (threadIdx.x == 0) {
    p[0] += K;
    p[2] += K;
  }
}

hipError_t launch_advance_tbase(uint32_t* tbase, uint32_t K, hipStream_t st) {
  hipLaunchKernelGGL(k_advance_tbase, dim3(1), dim3(64), 0, st, tbase, K);
  return hipGetLastError();
}

hipError_t launch_set_gbest(const SwarmDev& S, int d, uint32_t par, const float* pos_dev,
                            float cost, hipStream_t st) {
  hipLaunchKernelGGL(k_set_gbest, dim3(1), dim3(64), 0, st, S, d, par, pos_dev, cost);
  return hipGetLastError();
}

__global__ __launch_bounds__(1024) void k_argmin_f32(const float* v, int64_t n, uint64_t* out) {
  __shared__ uint64_t s_red[16];
  uint64_t k = KEY_NONE_;
  for (int64_t i = threadIdx.x; i < n; i += 1024) k = umin64(k, make_key(v[i], (uint32_t)i));
  k = block_min_u64_<1024>(k, s_red);
  if (threadIdx.x == 0) *out = k;
}

hipError_t launch_argmin_f32(const float* vals, int64_t n, uint64_t* out_key, hipStream_t st) {
  hipLaunchKernelGGL(k_argmin_f32, dim3(1), dim3(1024), 0, st, vals, n, out_key);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_soa_to_aos(const float* src, int64_t npad, int64_t n,
                                                    int d, float* dst) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k < n * d) {
    const int64_t i = k / d, j = k - i * d;
    dst[k] = src[j * npad + i];
  }
}

hipError_t launch_soa_to_aos(const float* src, int64_t npad, int64_t n, int d, float* dst,
                             hipStream_t st) {
  const int64_t blocks = (n * d + 255) / 256;
  if (blocks <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_soa_to_aos, dim3((unsigned)blocks), dim3(256), 0, st, src, npad, n, d, dst);
  return hipGetLastError();
}

const KernelSet* kernels_rosenbrock(int d, bool constrained);
const KernelSet* kernels_rastrigin(int d, bool constrained);
const KernelSet* kernels_sphere(int d, bool constrained);
const KernelSet* kernels_two_dim(int obj, int d);
const KernelSet* kernels_ode(int obj, int d);

const KernelSet* find_kernels(int obj, int d, bool constrained) {
  if (d < 1 || d > PPS_MAXD) return nullptr;
  switch (obj) {
    case PPS_OBJ_ROSENBROCK: return kernels_rosenbrock(d, constrained);
    case PPS_OBJ_RASTRIGIN: return kernels_rastrigin(d, constrained);
    case PPS_OBJ_SPHERE: return kernels_sphere(d, constrained);
    default:
      if (constrained) return nullptr;
      return is_ode_objective(obj) ? kernels_ode(obj, d) : kernels_two_dim(obj, d);
  }
}

}  // namespace pps
