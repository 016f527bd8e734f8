// pps_k_ode.hip — kernel instantiations for ODE parameter estimation: two
// models (Lotka-Volterra, FitzHugh-Nagumo) times five in-kernel solvers (RK4;
// Tsit5 fixed and adaptive, each with samples on step ends or read from the
// continuous extension).
#include "pps_kernels_impl.h"

namespace pps {
const KernelSet* kernels_ode(int obj, int d) {
  if (d != 4) return nullptr;
  switch (obj) {
#define PPS_ODE_SET(o) \
    case o: return Launch<4, true, o>::set();
    PPS_ODE_SET(PPS_OBJ_LOTKA_VOLTERRA)
    PPS_ODE_SET(PPS_OBJ_LOTKA_VOLTERRA_TSIT5)
    PPS_ODE_SET(PPS_OBJ_LOTKA_VOLTERRA_ATSIT5)
    PPS_ODE_SET(PPS_OBJ_LOTKA_VOLTERRA_TSIT5_SAVEAT)
    PPS_ODE_SET(PPS_OBJ_LOTKA_VOLTERRA_ATSIT5_SAVEAT)
    PPS_ODE_SET(PPS_OBJ_FITZHUGH_NAGUMO)
    PPS_ODE_SET(PPS_OBJ_FITZHUGH_NAGUMO_TSIT5)
    PPS_ODE_SET(PPS_OBJ_FITZHUGH_NAGUMO_ATSIT5)
    PPS_ODE_SET(PPS_OBJ_FITZHUGH_NAGUMO_TSIT5_SAVEAT)
    PPS_ODE_SET(PPS_OBJ_FITZHUGH_NAGUMO_ATSIT5_SAVEAT)
#undef PPS_ODE_SET
    default: return nullptr;
  }
}
}  // namespace pps
