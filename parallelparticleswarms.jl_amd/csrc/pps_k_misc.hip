// pps_k_misc.hip — kernel instantiations for the two-dimensional known-answer
// functions.
#include "pps_kernels_impl.h"

namespace pps {
const KernelSet* kernels_two_dim(int obj, int d) {
  switch (obj) {
    case PPS_OBJ_BEALE: return d == 2 ? Launch<2, true, PPS_OBJ_BEALE>::set() : nullptr;
    case PPS_OBJ_BOOTH: return d == 2 ? Launch<2, true, PPS_OBJ_BOOTH>::set() : nullptr;
    case PPS_OBJ_HIMMELBLAU: return d == 2 ? Launch<2, true, PPS_OBJ_HIMMELBLAU>::set() : nullptr;
    case PPS_OBJ_QUADRATIC: return d == 2 ? Launch<2, true, PPS_OBJ_QUADRATIC>::set() : nullptr;
    case PPS_OBJ_BOUNDARY_QUADRATIC:
      return d == 1 ? Launch<1, true, PPS_OBJ_BOUNDARY_QUADRATIC>::set() : nullptr;
    case PPS_OBJ_ROSEN2: return d == 2 ? Launch<2, true, PPS_OBJ_ROSEN2>::set() : nullptr;
    default: return nullptr;
  }
}
}  // namespace pps
