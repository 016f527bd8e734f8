// pps_k_ode.hip — kernel instantiations for Lotka-Volterra estimation with
// Tsit5 under saveat semantics (continuous-extension reads between steps).
#include "pps_kernels_impl.h"

namespace pps {
const KernelSet* kernels_lotka_volterra_saveat(int obj) {
  switch (obj) {
    case PPS_OBJ_LOTKA_VOLTERRA_TSIT5_SAVEAT:
      return Launch<4, true, PPS_OBJ_LOTKA_VOLTERRA_TSIT5_SAVEAT>::set();
    case PPS_OBJ_LOTKA_VOLTERRA_ATSIT5_SAVEAT:
      return Launch<4, true, PPS_OBJ_LOTKA_VOLTERRA_ATSIT5_SAVEAT>::set();
    default: return nullptr;
  }
}
}  // namespace pps
