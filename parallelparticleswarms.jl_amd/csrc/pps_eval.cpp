// pps_eval.cpp — FP64 host evaluation of the objective set, used to
// re-evaluate FP32 minima found on the device (the reference evaluates the
// final objective on the host in the problem's element type,
// src/solve.jl:28-33).
#include <cmath>
#include <cstdint>
#include <vector>

#include "pps.h"

namespace {

double cos2pi_f64(double x) {
  const double r = x - std::nearbyint(x);  // exact
  return std::cos(2.0 * M_PI * r);
}

double lv_sse_f64(const double* th, const pps_problem* P) {
  double u1 = P->params[0], u2 = P->params[1];
  const double dt = P->params[2];
  const int spo = (int)P->params[3];
  auto rhs = [&](double a, double b, double& f1, double& f2) {
    f1 = th[0] * a - th[1] * a * b;
    f2 = -th[2] * b + th[3] * a * b;
  };
  double loss = 0.0;
  for (int k = 0; k < P->n_obs; ++k) {
    if (k > 0) {
      for (int s = 0; s < spo; ++s) {
        double k1a, k1b, k2a, k2b, k3a, k3b, k4a, k4b;
        rhs(u1, u2, k1a, k1b);
        rhs(u1 + 0.5 * dt * k1a, u2 + 0.5 * dt * k1b, k2a, k2b);
        rhs(u1 + 0.5 * dt * k2a, u2 + 0.5 * dt * k2b, k3a, k3b);
        rhs(u1 + dt * k3a, u2 + dt * k3b, k4a, k4b);
        u1 += dt / 6.0 * (k1a + 2.0 * k2a + 2.0 * k3a + k4a);
        u2 += dt / 6.0 * (k1b + 2.0 * k2b + 2.0 * k3b + k4b);
      }
    }
    const double e1 = (double)P->obs[2 * k] - u1, e2 = (double)P->obs[2 * k + 1] - u2;
    loss += e1 * e1 + e2 * e2;
  }
  return loss;
}

double eval_one(const pps_problem* P, const double* xin) {
  const int d = P->d;
  std::vector<double> z(d);
  for (int j = 0; j < d; ++j) z[j] = P->shift ? xin[j] - (double)P->shift[j] : xin[j];
  const float* p = P->params;
  switch (P->objective) {
    case PPS_OBJ_ROSENBROCK: {
      double r = 0.0;
      for (int i = 0; i + 1 < d; ++i) {
        const double a = z[i + 1] - z[i] * z[i], b = p[0] - z[i];
        r += p[1] * a * a + b * b;
      }
      return r;
    }
    case PPS_OBJ_RASTRIGIN: {
      double r = 10.0 * d;
      for (int i = 0; i < d; ++i) r += z[i] * z[i] - 10.0 * cos2pi_f64(z[i]);
      return r;
    }
    case PPS_OBJ_SPHERE: {
      double r = 0.0;
      for (int i = 0; i < d; ++i) r += z[i] * z[i];
      return r;
    }
    case PPS_OBJ_BEALE: {
      const double x1 = z[0], x2 = z[1];
      const double a = 1.5 - x1 + x1 * x2, b = 2.25 - x1 + x1 * x2 * x2,
                   c = 2.625 - x1 + x1 * x2 * x2 * x2;
      return a * a + b * b + c * c;
    }
    case PPS_OBJ_BOOTH: {
      const double a = z[0] + 2 * z[1] - 7, b = 2 * z[0] + z[1] - 5;
      return a * a + b * b;
    }
    case PPS_OBJ_HIMMELBLAU: {
      const double a = z[0] * z[0] + z[1] - 11, b = z[0] + z[1] * z[1] - 7;
      return a * a + b * b;
    }
    case PPS_OBJ_QUADRATIC: {
      const double a = z[0] - 1, b = z[1] + 2;
      return p[0] * a * a + b * b;
    }
    case PPS_OBJ_BOUNDARY_QUADRATIC: {
      const double a = z[0] + 1;
      return a * a;
    }
    case PPS_OBJ_ROSEN2: {
      const double a = p[0] - z[0], b = z[1] - z[0] * z[0];
      return a * a + p[1] * b * b;
    }
    case PPS_OBJ_LOTKA_VOLTERRA:
      return lv_sse_f64(z.data(), P);
  }
  return NAN;
}

}  // namespace

extern "C" int pps_eval_f64(const pps_problem* prob, int64_t n, const double* x, double* f) {
  if (!prob || (n > 0 && (!x || !f)) || n < 0 || prob->d < 1 || prob->d > 64) return PPS_E_INVALID;
  if (prob->objective == PPS_OBJ_LOTKA_VOLTERRA &&
      (prob->d != 4 || !prob->obs || prob->n_obs < 1 || !prob->params || prob->n_params < 4))
    return PPS_E_INVALID;
  for (int64_t i = 0; i < n; ++i) f[i] = eval_one(prob, x + i * prob->d);
  return PPS_OK;
}
