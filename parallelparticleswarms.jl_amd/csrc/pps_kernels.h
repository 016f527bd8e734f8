// pps_kernels.h — host-side launch interface of the libpps kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "pps_shared.h"

namespace pps {

struct StepArgs {
  float c1, c2;
  int32_t gbest_from;  // PPS_GBEST_*
  int32_t min_improvers;
  // Nonzero in the launches of a replayed chunk graph: launch k takes its
  // inertia weight from S.wtab[tbase[2] + k - 1] instead of its w argument.
  int32_t w_table;
};

// One set of launchers per (objective, dimension bucket).
struct KernelSet {
  int maxd;
  // Largest workgroup (particles) the multi-workgroup sync kernels (init,
  // rekey, sync_step) are compiled for: 1024 or 512 (then also 512 and 256),
  // or the objective's block_of() size alone.
  int max_block;
  // Uniform init of every particle, then block keys/candidates for parity
  // `par` (G workgroups of `block` threads).
  hipError_t (*init)(const SwarmDev&, const ProbArgs&, uint32_t par, hipStream_t);
  // Recompute block keys/candidates from the current state (no step).
  hipError_t (*rekey)(const SwarmDev&, const ProbArgs&, uint32_t par, int32_t gbest_from,
                      hipStream_t);
  // One sync iteration t = tbase[0] + k of epoch tbase[1] (reads parity
  // (t-1)&1, writes t&1).
  hipError_t (*sync_step)(const SwarmDev&, const ProbArgs&, const StepArgs&, float w,
                          uint32_t k, hipStream_t);
  // Sync iterations t0+1 .. t0+T in one workgroup with the swarm in registers
  // (n <= 1024, G == 1).
  hipError_t (*sync_wg)(const SwarmDev&, const ProbArgs&, const StepArgs&, float w0,
                        float wdamp, uint32_t t0, int32_t T, hipStream_t);
  // Async: every particle runs T iterations against its running best.
  hipError_t (*async_run)(const SwarmDev&, const ProbArgs&, const StepArgs&, float w0,
                          float wdamp, uint32_t t0, int32_t T, hipStream_t);
  // Queue-lock: one iteration t = tbase[0] + k against the live global best
  // (workgroups of 1024 particles).
  hipError_t (*qlock_step)(const SwarmDev&, const ProbArgs&, const StepArgs&, float w,
                           uint32_t k, hipStream_t);
  // Queue-lock iterations tbase[0]+1 .. tbase[0]+T of a swarm of at most one
  // such workgroup, in one launch.
  hipError_t (*qlock_wg)(const SwarmDev&, const ProbArgs&, const StepArgs&, float w0,
                         float wdamp, int32_t T, hipStream_t);
};

// nullptr when (obj, d) has no compiled kernels (constrained: no kernels that
// evaluate the constraint penalty).
const KernelSet* find_kernels(int obj, int d, bool constrained = false);

// Reduce the block keys of parity `par` into swarm.gbrec (cost, idx, pos).
hipError_t launch_finalize(const SwarmDev&, int d, uint32_t par, hipStream_t);
// p[0] = v0, p[1] = v1 on the stream (a one-lane kernel, so the write can sit
// in a graph).
// Stream position of a run: tbase = {t0, epoch, 0}.
hipError_t launch_set_tbase(uint32_t* tbase, uint32_t t0, uint32_t epoch, hipStream_t);
// End of a chunk of K launches: tbase[0] += K (iteration base), tbase[2] += K
// (the weight table's offset).
hipError_t launch_advance_tbase(uint32_t* tbase, uint32_t K, hipStream_t);
// Write gbrec and a single-candidate parity `par` from host values.
hipError_t launch_set_gbest(const SwarmDev&, int d, uint32_t par, const float* pos_dev,
                            float cost, hipStream_t);

// Projected L-BFGS from every particle's best position into out_x (SoA,
// d x npad) and out_f (n).
bool lbfgs_supported(int obj, int d);
hipError_t launch_lbfgs(const SwarmDev&, const ProbArgs&, const LbfgsArgs&, float* out_x,
                        float* out_f, hipStream_t);
// dst[i * d + j] = src[j * npad + i] for i < n, j < d (SoA rows -> AoS).
hipError_t launch_soa_to_aos(const float* src, int64_t npad, int64_t n, int d, float* dst,
                             hipStream_t);
// Lowest-index argmin key of vals[0..n) (NaN never wins) into *out_key.
hipError_t launch_argmin_f32(const float* vals, int64_t n, uint64_t* out_key, hipStream_t);

}  // namespace pps
