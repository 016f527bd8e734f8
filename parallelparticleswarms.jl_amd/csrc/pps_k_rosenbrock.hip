// pps_k_rosenbrock.hip — kernel instantiations for PPS_OBJ_ROSENBROCK.
#include "pps_kernels_impl.h"

namespace pps {
const KernelSet* kernels_rosenbrock(int d, bool constrained) {
  return generic_set<PPS_OBJ_ROSENBROCK>(d, constrained);
}
}  // namespace pps
