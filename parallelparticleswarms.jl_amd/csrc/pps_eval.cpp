// pps_eval.cpp — FP64 host evaluation of the objective set, used to
// re-evaluate FP32 minima found on the device (the reference evaluates the
// final objective on the host in the problem's element type,
// src/solve.jl:28-33).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "pps_shared.h"

namespace {

double cos2pi_f64(double x) {
  const double r = x - std::nearbyint(x);  // exact
  return std::cos(2.0 * M_PI * r);
}

// Right-hand sides of the ODE estimation models (Lotka-Volterra,
// FitzHugh-Nagumo with rates (a, b, tau_inv, l)).
void ode_rhs_f64(int model, const double* th, double u1, double u2, double& f1, double& f2) {
  if (model == PPS_ODE_FITZHUGH_NAGUMO) {
    f1 = u1 - u1 * u1 * u1 / 3.0 - u2 + th[3];
    f2 = th[2] * (u1 + th[0] - th[1] * u2);
  } else {
    f1 = th[0] * u1 - th[1] * u1 * u2;
    f2 = -th[2] * u2 + th[3] * u1 * u2;
  }
}

double ode_sse_rk4_f64(const double* th, const pps_problem* P) {
  double u1 = P->params[0], u2 = P->params[1];
  const double dt = P->params[2];
  const int spo = (int)P->params[3];
  const int model = ode_model(P->objective);
  auto rhs = [&](double a, double b, double& f1, double& f2) { ode_rhs_f64(model, th, a, b, f1, f2); };
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

// Tsitouras 5(4) tableau (Comput. Math. Appl. 62 (2011) 770-775): stage
// weights (row 6 = solution weights), error weights b - bhat, and the
// continuous extension's coefficients b_i(theta) = sum_j R[i][j] theta^(j+1).
const double TSA[7][6] = {
    {0, 0, 0, 0, 0, 0},
    {0.161, 0, 0, 0, 0, 0},
    {-0.008480655492356989, 0.335480655492357, 0, 0, 0, 0},
    {2.897153057105493, -6.359448489975075, 4.3622954328695815, 0, 0, 0},
    {5.325864828439257, -11.748883564062828, 7.4955393428898365, -0.09249506636175525, 0, 0},
    {5.86145544294642, -12.92096931784711, 8.159367898576159, -0.071584973281401,
     -0.028269050394068383, 0},
    {0.09646076681806523, 0.01, 0.4798896504144996, 1.379008574103742, -3.290069515436081,
     2.324710524099774}};
const double TSE[7] = {-0.00178001105222577714, -0.0008164344596567469, 0.007880878010261995,
                       -0.1447110071732629,     0.5823571654525552,      -0.45808210592918697,
                       0.015151515151515152};
const double TSR[7][4] = {{1.0, -2.763706197274826, 2.9132554618219126, -1.0530884977290216},
                          {0.0, 0.13169999999999998, -0.2234, 0.1017},
                          {0.0, 3.9302962368947516, -5.941033872131505, 2.490627285651253},
                          {0.0, -12.411077166933676, 30.33818863028232, -16.548102889244902},
                          {0.0, 37.50931341651104, -88.1789048947664, 47.37952196281928},
                          {0.0, -27.896526289197286, 65.09189467479366, -34.87065786149661},
                          {0.0, 1.5, -4.0, 2.5}};

// Tsitouras 5(4) in FP64, fixed or adaptive (same controller as the device's
// lv_sse_tsit5: RMS scaled error, factor EEst^(1/5)/0.9 in [0.1, 5], steps
// ending on the observation times).
double ode_sse_tsit5_f64(const double* th, const pps_problem* P, bool adaptive) {
  const int model = ode_model(P->objective);
  auto rhs = [&](double a, double b, double& f1, double& f2) { ode_rhs_f64(model, th, a, b, f1, f2); };
  double u[2] = {P->params[0], P->params[1]}, k[7][2];
  rhs(u[0], u[1], k[0][0], k[0][1]);
  auto step = [&](double h, double* un, double* err) {
    for (int i = 1; i < 7; ++i) {
      double y[2];
      for (int c = 0; c < 2; ++c) {
        double acc = 0.0;
        for (int j = 0; j < i; ++j) acc += TSA[i][j] * k[j][c];
        y[c] = u[c] + h * acc;
      }
      if (i == 6) un[0] = y[0], un[1] = y[1];
      rhs(y[0], y[1], k[i][0], k[i][1]);
    }
    for (int c = 0; c < 2; ++c) {
      double acc = 0.0;
      for (int j = 0; j < 7; ++j) acc += TSE[j] * k[j][c];
      err[c] = h * acc;
    }
  };
  double loss = 0.0, t = 0.0, dt = adaptive ? P->params[3] : P->params[2];
  long steps = 0;
  for (int kk = 0; kk < P->n_obs; ++kk) {
    if (kk > 0) {
      const double t_next = kk * (double)P->params[2];
      int fixed_left = adaptive ? 0 : (int)P->params[3];
      while (adaptive ? t < t_next : fixed_left-- > 0) {
        const bool hit = adaptive && !(t + dt < t_next);
        const double h = hit ? t_next - t : dt;
        const long max_steps = P->n_params > 6 && P->params[6] > 0 ? (long)P->params[6] : 10000;
        if (adaptive && (++steps > max_steps || !(t + h > t))) return INFINITY;
        double un[2] = {0.0, 0.0}, err[2] = {0.0, 0.0};
        step(h, un, err);
        if (!adaptive) {
          u[0] = un[0], u[1] = un[1];
          k[0][0] = k[6][0], k[0][1] = k[6][1];
          t += h;
          continue;
        }
        const double s1 = err[0] / (P->params[4] + std::max(std::fabs(u[0]), std::fabs(un[0])) * P->params[5]);
        const double s2 = err[1] / (P->params[4] + std::max(std::fabs(u[1]), std::fabs(un[1])) * P->params[5]);
        const double eest = std::sqrt((s1 * s1 + s2 * s2) / 2.0);
        double q = 5.0;
        if (std::isfinite(eest)) q = std::min(5.0, std::max(0.1, std::pow(eest, 0.2) / 0.9));
        const double hnew = h / q;
        if (eest <= 1.0) {
          t = hit ? t_next : t + h;
          u[0] = un[0], u[1] = un[1];
          k[0][0] = k[6][0], k[0][1] = k[6][1];
          dt = hit ? std::max(dt, hnew) : hnew;
        } else {
          dt = hnew;
        }
      }
    }
    const double e1 = (double)P->obs[2 * kk] - u[0], e2 = (double)P->obs[2 * kk + 1] - u[1];
    loss += e1 * e1 + e2 * e2;
  }
  return loss;
}

// Tsit5 in FP64 with saveat semantics (as the device's lv_sse_tsit5_saveat):
// free steps from t = 0, the last ending on t_end = (n_obs - 1) * params[2],
// observations inside a step read from the continuous extension.
double ode_sse_tsit5_saveat_f64(const double* th, const pps_problem* P, bool adaptive) {
  const int model = ode_model(P->objective);
  auto rhs = [&](double a, double b, double& f1, double& f2) { ode_rhs_f64(model, th, a, b, f1, f2); };
  double u[2] = {P->params[0], P->params[1]}, k[7][2];
  rhs(u[0], u[1], k[0][0], k[0][1]);
  const double tau = P->params[2], t_end = (P->n_obs - 1) * tau;
  double e1 = (double)P->obs[0] - u[0], e2 = (double)P->obs[1] - u[1];
  double loss = e1 * e1 + e2 * e2, t = 0.0, dt = P->params[3];
  const long max_steps = !adaptive ? (1L << 40)
                         : P->n_params > 6 && P->params[6] > 0 ? (long)P->params[6] : 10000;
  long steps = 0;
  int kk = 1;
  while (kk < P->n_obs) {
    const bool last = !(t + dt < t_end);
    const double h = last ? t_end - t : dt;
    if (++steps > max_steps || !(t + h > t)) return INFINITY;
    double un[2] = {0.0, 0.0}, err[2] = {0.0, 0.0};
    for (int i = 1; i < 7; ++i) {
      double y[2];
      for (int c = 0; c < 2; ++c) {
        double acc = 0.0;
        for (int j = 0; j < i; ++j) acc += TSA[i][j] * k[j][c];
        y[c] = u[c] + h * acc;
      }
      if (i == 6) un[0] = y[0], un[1] = y[1];
      rhs(y[0], y[1], k[i][0], k[i][1]);
    }
    if (adaptive) {
      for (int c = 0; c < 2; ++c) {
        double acc = 0.0;
        for (int j = 0; j < 7; ++j) acc += TSE[j] * k[j][c];
        err[c] = h * acc / (P->params[4] + std::max(std::fabs(u[c]), std::fabs(un[c])) * P->params[5]);
      }
      const double eest = std::sqrt((err[0] * err[0] + err[1] * err[1]) / 2.0);
      double q = 5.0;
      if (std::isfinite(eest)) q = std::min(5.0, std::max(0.1, std::pow(eest, 0.2) / 0.9));
      dt = h / q;
      if (!(eest <= 1.0)) continue;
    }
    const double t_new = last ? t_end : t + h;
    for (; kk < P->n_obs && kk * tau <= t_new; ++kk) {
      const double th_ = (kk * tau - t) / h;
      double v[2] = {u[0], u[1]};
      for (int i = 0; i < 7; ++i) {
        double b = 0.0;
        for (int j = 3; j >= 0; --j) b = (b + TSR[i][j]) * th_;
        v[0] += h * b * k[i][0];
        v[1] += h * b * k[i][1];
      }
      e1 = (double)P->obs[2 * kk] - v[0];
      e2 = (double)P->obs[2 * kk + 1] - v[1];
      loss += e1 * e1 + e2 * e2;
    }
    t = t_new;
    u[0] = un[0], u[1] = un[1];
    k[0][0] = k[6][0], k[0][1] = k[6][1];
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
    default:
      if (!is_ode_objective(P->objective)) break;
      switch (ode_solver(P->objective)) {
        case PPS_ODE_RK4: return ode_sse_rk4_f64(z.data(), P);
        case PPS_ODE_TSIT5: return ode_sse_tsit5_f64(z.data(), P, false);
        case PPS_ODE_ATSIT5: return ode_sse_tsit5_f64(z.data(), P, true);
        case PPS_ODE_TSIT5_SAVEAT: return ode_sse_tsit5_saveat_f64(z.data(), P, false);
        default: return ode_sse_tsit5_saveat_f64(z.data(), P, true);
      }
  }
  return NAN;
}

}  // namespace

extern "C" int pps_eval_f64(const pps_problem* prob, int64_t n, const double* x, double* f) {
  if (!prob || (n > 0 && (!x || !f)) || n < 0 || prob->d < 1 || prob->d > 64) return PPS_E_INVALID;
  const bool ode = is_ode_objective(prob->objective);
  const int need = ode && ode_adaptive(prob->objective) ? 6 : 4;
  if (ode && (prob->d != 4 || !prob->obs || prob->n_obs < 1 || !prob->params ||
             prob->n_params < need))
    return PPS_E_INVALID;
  for (int64_t i = 0; i < n; ++i) f[i] = eval_one(prob, x + i * prob->d);
  return PPS_OK;
}
