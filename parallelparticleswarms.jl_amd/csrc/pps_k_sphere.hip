// pps_k_sphere.hip — kernel instantiations for PPS_OBJ_SPHERE.
#include "pps_kernels_impl.h"

namespace pps {
const KernelSet* kernels_sphere(int d) { return generic_set<PPS_OBJ_SPHERE>(d); }
}  // namespace pps
