// pps_batch.h — device descriptors of packed (batched / served) problems.
#pragma once
#include <hip/hip_runtime.h>

#include "pps_shared.h"

#define PPS_BMAXD 32  // largest dimension of a packed problem

namespace pps {

constexpr int BATCH_BLOCK = 256;         // workgroup of a small problem
constexpr int BATCH_BLOCK_LARGE = 1024;  // workgroup of a large problem

// One admitted problem.  Written by the host at admission, then owned by the
// device until the problem finishes (remaining == 0).
struct BatchProb {
  ProbArgs P;         // objective, bounds, shift, parameters
  int64_t n, npad;
  uint32_t key0, key1;
  uint32_t iter;      // completed iterations
  int32_t remaining;  // iterations still to run
  float w, wdamp, c1, c2;
  float* x;
  float* v;
  float* pb;
  float* pc;
  float gb[PPS_BMAXD];  // attractor of the next iteration
  float gb_cost;
  uint32_t gb_idx;
  int32_t initialized;
  int32_t pad_;
};

// Result of one problem, written by its workgroup into host-visible pinned
// memory; `done` is stored last (system-scope release) so the host may read
// the other fields and reuse the problem's arena space once it sees done != 0.
struct BatchOut {
  float gb_cost;
  uint32_t gb_idx;
  uint32_t iter;
  int32_t remaining;
  float gb[PPS_BMAXD];
  uint32_t done;
  uint32_t pad_[3];
};

bool batch_supported(int obj, int d);
// Problems probs[ids[0..count)] each run in one workgroup until finished, or
// `chunk` iterations, or `budget_ticks` of the device wall clock (checked
// after every iteration).  When fresh_in[id] == launch_no the workgroup first
// copies its descriptor fresh[id] (host-visible memory) -> probs[id].  If
// `out` is non-null each problem's result and done flag are published to
// out[id] when its remaining count reaches zero.
// `block` selects the workgroup size (BATCH_BLOCK or BATCH_BLOCK_LARGE).
hipError_t launch_batch(BatchProb* probs, const BatchProb* fresh, const uint32_t* fresh_in,
                        uint32_t launch_no, const int32_t* ids, int32_t count, int32_t chunk,
                        uint64_t budget_ticks, BatchOut* out, int block, hipStream_t st);

}  // namespace pps
