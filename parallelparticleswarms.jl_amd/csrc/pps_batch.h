// pps_batch.h — device descriptors of packed (batched / served) problems.
#pragma once
#include <hip/hip_runtime.h>

#include "pps_shared.h"

#define PPS_BMAXD 32  // largest dimension of a packed problem

namespace pps {

constexpr int BATCH_BLOCK = 256;   // threads per work item
constexpr int BATCH_MAXPARTS = 64; // workgroups per iteration of a large problem

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

// Exchange area of a large problem: the key (order(cost) << 32 | index) and
// best position of each part's best particle, by iteration parity.
struct BatchXchg {
  uint64_t key[2][BATCH_MAXPARTS];
  float cand[2][BATCH_MAXPARTS][PPS_BMAXD];
};

enum { BATCH_SMALL = 0, BATCH_PART = 1, BATCH_FINAL = 2 };

// One workgroup of a slice.
struct BatchItem {
  int32_t slot;
  int32_t kind;    // BATCH_*
  int32_t part;    // PART: q in 0..nparts-1
  int32_t nparts;  // PART/FINAL: k
  uint32_t t;      // PART: iteration (0 = init); FINAL: last iteration
  float w;         // PART: inertia of iteration t
  int32_t fresh;   // SMALL: copy the descriptor from host memory first
  int32_t pad_;
};

bool batch_supported(int obj, int d);
// Parts of a large problem of n particles: tiles of 256 particles dealt
// round robin, tiles_per_part of them per part, at most BATCH_MAXPARTS parts.
int batch_parts(int64_t n, int tiles_per_part);
// One slice: items[0..count) each run in one workgroup of BATCH_BLOCK threads.
// SMALL items stop after the first iteration that ends past budget_ticks of
// the device wall clock.  Finished problems publish result + done flag to
// out[slot] (host-visible memory).
hipError_t launch_slice(BatchProb* probs, const BatchProb* fresh, const BatchItem* items,
                        int32_t count, uint64_t budget_ticks, BatchOut* out, BatchXchg* xchg,
                        hipStream_t st);

}  // namespace pps
