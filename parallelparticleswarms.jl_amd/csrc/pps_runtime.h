// pps_runtime.h — internals shared by the runtime translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "pps_shared.h"

struct pps_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int num_cus = 0;
  int clock_khz = 0;
  size_t total_mem = 0;
  std::string arch;
};

namespace pps {

// Message of the last failure on this thread (pps_last_error).
extern thread_local std::string g_err;
// Record a failure message for pps_last_error() and return `code`.
int fail(int code, const char* fmt, ...);
int hip_fail(hipError_t e, const char* what, const char* file, int line);
// Validate a pps_problem and convert it to kernel arguments (obs pointer left
// unset: the caller uploads observations).
int fill_prob_args(const pps_problem* prob, ProbArgs* P);
// The step's c1 and c2: 0, not finite, or at least 2^-102 in magnitude (the
// kernels fold the draws' 2^-24 scale into them, exactly for these values).
int check_step_coefs(float c1, float c2);

}  // namespace pps
