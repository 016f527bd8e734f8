// pps_k_sphere.hip — kernel instantiations for PPS_OBJ_SPHERE.
#include "pps_kernels_impl.h"

namespace pps {
const KernelSet* kernels_sphere(int d, bool constrained) {
  return generic_set<PPS_OBJ_SPHERE>(d, constrained);
}
}  // namespace pps
