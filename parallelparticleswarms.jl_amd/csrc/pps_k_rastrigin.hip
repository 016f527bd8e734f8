// pps_k_rastrigin.hip — kernel instantiations for PPS_OBJ_RASTRIGIN.
#include "pps_kernels_impl.h"

namespace pps {
const KernelSet* kernels_rastrigin(int d, bool constrained) {
  return generic_set<PPS_OBJ_RASTRIGIN>(d, constrained);
}
}  // namespace pps
