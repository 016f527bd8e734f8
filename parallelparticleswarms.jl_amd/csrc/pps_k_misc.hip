// pps_k_misc.hip — kernel instantiations for the fixed-dimension objectives
// (the two-dimensional known-answer functions and Lotka-Volterra estimation).
#include "pps_kernels_impl.h"

namespace pps {
const KernelSet* kernels_lotka_volterra_saveat(int obj);

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
const KernelSet* kernels_lotka_volterra(int obj, int d) {
  if (d != 4) return nullptr;
  switch (obj) {
    case PPS_OBJ_LOTKA_VOLTERRA: return Launch<4, true, PPS_OBJ_LOTKA_VOLTERRA>::set();
    case PPS_OBJ_LOTKA_VOLTERRA_TSIT5: return Launch<4, true, PPS_OBJ_LOTKA_VOLTERRA_TSIT5>::set();
    case PPS_OBJ_LOTKA_VOLTERRA_ATSIT5:
      return Launch<4, true, PPS_OBJ_LOTKA_VOLTERRA_ATSIT5>::set();
    default: return kernels_lotka_volterra_saveat(obj);
  }
}
}  // namespace pps
