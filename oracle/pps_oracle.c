/*
 * pps_oracle.c — CPU restatement of ParallelParticleSwarms.jl's particle-swarm
 * hot path.  TEST INFRASTRUCTURE ONLY: the parity tests and the smoke check
 * compare the kernels' results with it, and the bench's CPU-baseline leg times
 * it.  Nothing in the product links or calls it.
 *
 * It follows the reference (v1.4.0) as text, function by function:
 *   - particle layout: SPSOParticle{SVector{d,Float32},Float32} as an array of
 *     structs, (3d+2) floats per particle: position, velocity, cost,
 *     best_position, best_cost (src/ParallelParticleSwarms.jl:34-40);
 *   - the per-particle step update_particle_state (src/kernels.jl:1-26) in
 *     Julia's evaluation order, FP32, no FMA contraction;
 *   - the sync CPU kernel + driver (src/kernels.jl:182-194,
 *     src/lowerlevel_solve.jl:38-65) and the GPU sync driver's gbest choice
 *     (src/kernels.jl:127-177, src/lowerlevel_solve.jl:1-36);
 *   - the async kernel (src/kernels.jl:196-225, src/lowerlevel_solve.jl:95-121);
 *   - the queue-lock kernel (src/kernels.jl:38-125, src/lowerlevel_solve.jl:67-93)
 *     with its blocks run one after another in index order (one legal schedule);
 *   - uniform initialisation (src/utils.jl:1-5,57-84; src/init.jl:38-54);
 *   - the ODE clamp quirk (src/ode_pso.jl:18-19) and loss (src/ode_pso.jl:47-52).
 * Julia's task-local RNG is replaced by Philox4x32-10 under the stream layout
 * that the kernels use (the reference's own RNG cannot be shared with a GPU
 * kernel written elsewhere).
 *
 * Compile with -ffp-contract=off and without -ffast-math.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORA_MAXD 64
#define ORA_MAXPAR 8

enum {
  ORA_ROSENBROCK = 1, ORA_RASTRIGIN = 2, ORA_SPHERE = 3, ORA_BEALE = 4,
  ORA_BOOTH = 5, ORA_HIMMELBLAU = 6, ORA_QUADRATIC = 7,
  ORA_BOUNDARY_QUADRATIC = 8, ORA_ROSEN2 = 9, ORA_LOTKA_VOLTERRA = 16,
  ORA_LOTKA_VOLTERRA_TSIT5 = 17, ORA_LOTKA_VOLTERRA_ATSIT5 = 18,
  ORA_LOTKA_VOLTERRA_TSIT5_SAVEAT = 19, ORA_LOTKA_VOLTERRA_ATSIT5_SAVEAT = 20,
  ORA_FITZHUGH_NAGUMO = 21, ORA_FITZHUGH_NAGUMO_ATSIT5_SAVEAT = 25
};
enum { ORA_GBEST_PBEST = 0, ORA_GBEST_CURRENT = 1 };
enum { ORA_CLAMP_COMPONENT = 0, ORA_CLAMP_LEXICOGRAPHIC = 1 };

typedef struct ora_problem {
  int32_t obj, d, has_shift, clamp;
  float lb[ORA_MAXD], ub[ORA_MAXD], shift[ORA_MAXD], par[ORA_MAXPAR];
  const float* obs;
  int32_t n_obs, steps_per_obs;
  /* constraints g_k(x) <= 0: n_cons rows [a_0..a_{d-1}, q_0..q_{d-1}, c] */
  const float* cons;
  int32_t n_cons;
  float u0[ORA_MAXD]; /* the unbounded sampler's centre */
} ora_problem;

enum { ORA_SAMPLER_UNIFORM = 0, ORA_SAMPLER_UNBOUNDED = 1, ORA_SAMPLER_LATIN_HYPERCUBE = 2 };

/* ---------------------------------------------------------------- Philox */
/* Philox4x32-10, Salmon et al. SC'11: round multipliers 0xD2511F53,
 * 0xCD9E8D57, Weyl increments 0x9E3779B9, 0xBB67AE85. */
void ora_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    uint64_t a = (uint64_t)0xD2511F53u * c0;
    uint64_t b = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(b >> 32) ^ c1 ^ k0;
    uint32_t n1 = (uint32_t)b;
    uint32_t n2 = (uint32_t)(a >> 32) ^ c3 ^ k1;
    uint32_t n3 = (uint32_t)a;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static inline float u01(uint32_t u) { return (float)(u >> 8) * (1.0f / 16777216.0f); }

static inline void seed_key(uint64_t seed, uint32_t key[2]) {
  key[0] = (uint32_t)seed;
  key[1] = (uint32_t)(seed >> 32);
}

/* ------------------------------------------------- Julia scalar semantics */
/* Base.max/Base.min on Float32 (NaN-propagating, max(-0,+0) = +0). */
static inline float jl_max(float x, float y) {
  float diff = x - y;
  float pick = signbit(diff) ? y : x;
  return (isnan(x) || isnan(y)) ? diff : pick;
}
static inline float jl_min(float x, float y) {
  float diff = x - y;
  float pick = signbit(diff) ? x : y;
  return (isnan(x) || isnan(y)) ? diff : pick;
}
/* Base.isless on Float32: total order with -0 < +0 and NaN last. */
static inline int jl_isless(float a, float b) {
  if (isnan(a) || isnan(b)) return !isnan(a) && isnan(b);
  if (a == 0.0f && b == 0.0f) return signbit(a) && !signbit(b);
  return a < b;
}
static inline int jl_isequal(float a, float b) {
  if (isnan(a) && isnan(b)) return 1;
  if (a == 0.0f && b == 0.0f) return signbit(a) == signbit(b);
  return a == b;
}
/* Base.cmp(::AbstractVector, ::AbstractVector): lexicographic with isequal /
 * isless on the first differing component. */
static int jl_cmp_vec(const float* a, const float* b, int d) {
  for (int j = 0; j < d; ++j)
    if (!jl_isequal(a[j], b[j])) return jl_isless(a[j], b[j]) ? -1 : 1;
  return 0;
}

/* ------------------------------------------------------------ objectives */
/* cos(2 pi x): reduce r = x - rint(x) (exact), a = |r|, fold a > 1/4 to
 * 1/2 - a with a sign flip, then the degree-14 Taylor polynomial of
 * cos(2 pi a) in s = a^2 (coefficients (-1)^k (2pi)^2k/(2k)! rounded to FP32),
 * Horner with fmaf. */
float ora_cos2pi(float x) {
  static const float C[8] = {1.0f,          -19.739208802178716f, 64.93939402266828f,
                             -85.45681720669371f, 60.24464137187664f, -26.426256783374388f,
                             7.903536371318465f,  -1.7143907110886711f};
  float r = x - rintf(x);
  float a = fabsf(r);
  int flip = a > 0.25f;
  if (flip) a = 0.5f - a;
  float s = a * a;
  float p = C[7];
  for (int k = 6; k >= 0; --k) p = fmaf(p, s, C[k]);
  return flip ? -p : p;
}
float ora_sin2pi(float x) {
  float r = x - rintf(x);
  return ora_cos2pi(r - 0.25f);
}

/* ODE estimation models, objective codes 16..20 (Lotka-Volterra,
 * benchmarks/Lotka_Volterra/lotka_volterra.jl:9-13) and 21..25
 * (FitzHugh-Nagumo, benchmarks/Fitzhugh_Nagumo/ode_estimation.jl:6-12, rates
 * (a, b, tau_inv, l): dv = v - v^3/3 - w + l, dw = tau_inv (v + a - b w), v^3
 * as Julia's literal power (v*v)*v), each with five solvers in the order RK4,
 * Tsit5, adaptive Tsit5, and the two saveat forms; and the fixed-step
 * classical RK4 SSE loss against observations (src/ode_pso.jl:47-52: per
 * sample sum(x.^2) of the 2-vector error, samples summed in order). */
static int ode_model(int obj) { return obj >= ORA_FITZHUGH_NAGUMO; }
static int ode_solver(int obj) {
  return obj - (obj >= ORA_FITZHUGH_NAGUMO ? ORA_FITZHUGH_NAGUMO : ORA_LOTKA_VOLTERRA);
}
static void ode_rhs(int model, float u1, float u2, const float* p, float* f1, float* f2) {
  if (model) {
    *f1 = ((u1 - ((u1 * u1) * u1) / 3.0f) - u2) + p[3];
    *f2 = p[2] * ((u1 + p[0]) - p[1] * u2);
  } else {
    *f1 = p[0] * u1 - (p[1] * u1) * u2;
    *f2 = (-p[2]) * u2 + (p[3] * u1) * u2;
  }
}
/* Tsitouras 5(4) (Comput. Math. Appl. 62 (2011) 770-775), the method of the
 * reference's default ode_alg = GPUTsit5() (src/ode_pso.jl:54-170): stage
 * weights a (row 6 = solution weights b, FSAL) and error weights
 * btilde = b - bhat, as FP32. */
static const float TS_A[7][6] = {
    {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f},
    {0.161f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f},
    {-0.008480655492356989f, 0.335480655492357f, 0.0f, 0.0f, 0.0f, 0.0f},
    {2.897153057105493f, -6.359448489975075f, 4.3622954328695815f, 0.0f, 0.0f, 0.0f},
    {5.325864828439257f, -11.748883564062828f, 7.4955393428898365f, -0.09249506636175525f, 0.0f,
     0.0f},
    {5.86145544294642f, -12.92096931784711f, 8.159367898576159f, -0.071584973281401f,
     -0.028269050394068383f, 0.0f},
    {0.09646076681806523f, 0.01f, 0.4798896504144996f, 1.379008574103742f, -3.290069515436081f,
     2.324710524099774f}};
static const float TS_E[7] = {-0.00178001105222577714f, -0.0008164344596567469f,
                              0.007880878010261995f,    -0.1447110071732629f,
                              0.5823571654525552f,      -0.45808210592918697f,
                              0.015151515151515152f};

/* Continuous extension (Tsitouras 2011, Sec. 4, order 4): b_i(theta) =
 * r_i0 theta + r_i1 theta^2 + r_i2 theta^3 + r_i3 theta^4, b_i(1) = b_i. */
static const float TS_R[7][4] = {
    {1.0f, -2.763706197274826f, 2.9132554618219126f, -1.0530884977290216f},
    {0.0f, 0.13169999999999998f, -0.2234f, 0.1017f},
    {0.0f, 3.9302962368947516f, -5.941033872131505f, 2.490627285651253f},
    {0.0f, -12.411077166933676f, 30.33818863028232f, -16.548102889244902f},
    {0.0f, 37.50931341651104f, -88.1789048947664f, 47.37952196281928f},
    {0.0f, -27.896526289197286f, 65.09189467479366f, -34.87065786149661f},
    {0.0f, 1.5f, -4.0f, 2.5f}};

/* The tableau as used (for the order-condition checks). */
void ora_tsit5_tableau(float* a /* 7 x 6 */, float* e /* 7 */) {
  memcpy(a, TS_A, sizeof TS_A);
  memcpy(e, TS_E, sizeof TS_E);
}

/* The continuous extension's weights b_i(theta) as evaluated:
 * b_0 = theta (r_00 + theta (r_01 + theta (r_02 + theta r_03))),
 * b_i = theta (theta (r_i1 + theta (r_i2 + theta r_i3))) for i >= 1. */
void ora_tsit5_interp_weights(float theta, float* b /* 7 */) {
  b[0] = theta * (TS_R[0][0] + theta * (TS_R[0][1] + theta * (TS_R[0][2] + theta * TS_R[0][3])));
  for (int i = 1; i < 7; ++i)
    b[i] = theta * (theta * (TS_R[i][1] + theta * (TS_R[i][2] + theta * TS_R[i][3])));
}

/* One Tsit5 step from u with k[0] = f(u): y_i = u + h * (a_i0 k_0 + ... ),
 * sums left to right; k[i] = f(y_i); the new state is y_6; k[6] = f(y_6)
 * (FSAL); err = h * (e_0 k_0 + ... + e_6 k_6). */
static void tsit5_step(int model, const float u[2], float k[7][2], float h, const float* th,
                          float un[2], float err[2]) {
  for (int i = 1; i < 7; ++i) {
    float y[2];
    for (int c = 0; c < 2; ++c) {
      float acc = TS_A[i][0] * k[0][c];
      for (int j = 1; j < i; ++j) acc = acc + TS_A[i][j] * k[j][c];
      y[c] = u[c] + h * acc;
    }
    ode_rhs(model, y[0], y[1], th, &k[i][0], &k[i][1]);
    if (i == 6) un[0] = y[0], un[1] = y[1];
  }
  for (int c = 0; c < 2; ++c) {
    float acc = TS_E[0] * k[0][c];
    for (int j = 1; j < 7; ++j) acc = acc + TS_E[j] * k[j][c];
    err[c] = h * acc;
  }
}

/* x^(1/5) from IEEE basic operations: exponent-halving seed (the float's
 * bits, as an int32, minus those of 1.0f, divided by 5, plus those of 1.0f)
 * and four Newton steps y <- (4y + x/y^4)/5; 0 for x <= 0 or NaN. */
float ora_root5(float x) {
  if (!(x > 0.0f)) return 0.0f;
  if (isinf(x)) return x;
  int32_t bits;
  memcpy(&bits, &x, 4);
  bits = (bits - 0x3f800000) / 5 + 0x3f800000;
  float y;
  memcpy(&y, &bits, 4);
  for (int it = 0; it < 4; ++it) {
    float y2 = y * y;
    y = (4.0f * y + x / (y2 * y2)) / 5.0f;
  }
  return y;
}

/* Lotka-Volterra SSE with Tsit5, fixed (dt = par[2], par[3] steps per
 * observation; parameter_estim_ode! with Val(false)) or adaptive (observation
 * interval par[2], initial step par[3], abstol par[4], reltol par[5];
 * Val(true)).  DiffEqGPU's controller is not vendored; this one: scaled error
 * s_i = err_i / (abstol + max(|u_i|, |unew_i|) reltol), EEst = sqrt((s_1^2 +
 * s_2^2) * 0.5), accept when EEst <= 1, q = root5(EEst) / 0.9 in [0.1, 5]
 * (5 when EEst is not finite), next step h / q; steps end exactly on the
 * observation times (t_k = k * par[2] in FP32) and a step shortened to meet
 * one keeps the larger of the old and new proposal; more steps than par[6]
 * (10000 when par[6] <= 0), or a step that does not advance t, give +Inf. */
float ora_ode_sse_tsit5(const float* th, const ora_problem* P, int adaptive) {
  const int model = ode_model(P->obj);
  float u[2] = {P->par[0], P->par[1]}, k[7][2];
  ode_rhs(model, u[0], u[1], th, &k[0][0], &k[0][1]);
  float loss = 0.0f, t = 0.0f, dt = adaptive ? P->par[3] : P->par[2];
  int32_t steps = 0;
  for (int kk = 0; kk < P->n_obs; ++kk) {
    if (kk > 0) {
      if (!adaptive) {
        for (int s = 0; s < P->steps_per_obs; ++s) {
          float un[2], err[2];
          tsit5_step(model, u, k, dt, th, un, err);
          u[0] = un[0], u[1] = un[1];
          k[0][0] = k[6][0], k[0][1] = k[6][1];
        }
      } else {
        float t_next = (float)kk * P->par[2];
        float atol = P->par[4], rtol = P->par[5];
        int32_t max_steps = P->par[6] > 0.0f ? (int32_t)P->par[6] : 10000;
        while (t < t_next) {
          int hit = !(t + dt < t_next);
          float h = hit ? t_next - t : dt;
          if (++steps > max_steps || !(t + h > t)) return INFINITY;
          float un[2], err[2];
          tsit5_step(model, u, k, h, th, un, err);
          float s[2];
          for (int c = 0; c < 2; ++c) {
            float a = fabsf(u[c]), b = fabsf(un[c]);
            s[c] = err[c] / (atol + (a > b ? a : b) * rtol);
          }
          float eest = sqrtf((s[0] * s[0] + s[1] * s[1]) * 0.5f);
          float q = 5.0f;
          if (eest <= 3.0e38f) {
            q = ora_root5(eest) / 0.9f;
            q = q < 0.1f ? 0.1f : (q > 5.0f ? 5.0f : q);
          }
          float hnew = h / q;
          if (eest <= 1.0f) {
            t = hit ? t_next : t + h;
            u[0] = un[0], u[1] = un[1];
            k[0][0] = k[6][0], k[0][1] = k[6][1];
            dt = hit ? (hnew > dt ? hnew : dt) : hnew;
          } else {
            dt = hnew;
          }
        }
      }
    }
    float e1 = P->obs[2 * kk] - u[0];
    float e2 = P->obs[2 * kk + 1] - u[1];
    loss = loss + (e1 * e1 + e2 * e2);
  }
  return loss;
}

/* Lotka-Volterra SSE with Tsit5 and saveat semantics (GPUTsit5 filling
 * `saveat` between steps, e.g. benchmarks/Lotka_Volterra/lotka_volterra.jl:
 * 124-127 with dt = 0.1 and saves every 0.01): observation k at t_k = k *
 * par[2] (FP32 product); steps of par[3] (initial step when adaptive) from
 * t = 0, each shortened to end on t_end = t_{n_obs-1} when t + dt does not
 * fall short of it; after an accepted step [t, t + h] every unsaved t_k <=
 * t + h gets u + h * (b_0(theta) k_0 + ... + b_6(theta) k_6), sum left to
 * right, theta = (t_k - t) / h.  Adaptive: the controller of
 * ora_ode_sse_tsit5 with next step h / q after every attempt; step limit par[6]
 * (10000 when <= 0); +Inf on too many steps or a step that does not advance
 * t.  Fixed steps have no limit (the host bounds their number). */
float ora_ode_sse_tsit5_saveat(const float* th, const ora_problem* P, int adaptive) {
  const int model = ode_model(P->obj);
  float u[2] = {P->par[0], P->par[1]}, k[7][2];
  ode_rhs(model, u[0], u[1], th, &k[0][0], &k[0][1]);
  const float tau = P->par[2];
  const float t_end = (float)(P->n_obs - 1) * tau;
  float e1 = P->obs[0] - u[0], e2 = P->obs[1] - u[1];
  float loss = e1 * e1 + e2 * e2, t = 0.0f, dt = P->par[3];
  int32_t steps = 0;
  const int32_t max_steps =
      adaptive ? (P->par[6] > 0.0f ? (int32_t)P->par[6] : 10000) : 0x7fffffff;
  int kk = 1;
  while (kk < P->n_obs) {
    const int last = !(t + dt < t_end);
    const float h = last ? t_end - t : dt;
    if (++steps > max_steps || !(t + h > t)) return INFINITY;
    float un[2], err[2];
    tsit5_step(model, u, k, h, th, un, err);
    if (adaptive) {
      float s[2];
      for (int c = 0; c < 2; ++c) {
        float a = fabsf(u[c]), b = fabsf(un[c]);
        s[c] = err[c] / (P->par[4] + (a > b ? a : b) * P->par[5]);
      }
      float eest = sqrtf((s[0] * s[0] + s[1] * s[1]) * 0.5f);
      float q = 5.0f;
      if (eest <= 3.0e38f) {
        q = ora_root5(eest) / 0.9f;
        q = q < 0.1f ? 0.1f : (q > 5.0f ? 5.0f : q);
      }
      dt = h / q;
      if (!(eest <= 1.0f)) continue;
    }
    const float t_new = last ? t_end : t + h;
    for (; kk < P->n_obs; ++kk) {
      const float tk = (float)kk * tau;
      if (tk > t_new) break;
      float b[7], v[2];
      ora_tsit5_interp_weights((tk - t) / h, b);
      for (int c = 0; c < 2; ++c) {
        float acc = b[0] * k[0][c];
        for (int j = 1; j < 7; ++j) acc = acc + b[j] * k[j][c];
        v[c] = u[c] + h * acc;
      }
      e1 = P->obs[2 * kk] - v[0];
      e2 = P->obs[2 * kk + 1] - v[1];
      loss = loss + (e1 * e1 + e2 * e2);
    }
    t = t_new;
    u[0] = un[0], u[1] = un[1];
    k[0][0] = k[6][0], k[0][1] = k[6][1];
  }
  return loss;
}

float ora_ode_sse_rk4(const float* th, const ora_problem* P) {
  const int model = ode_model(P->obj);
  float u1 = P->par[0], u2 = P->par[1];
  float dt = P->par[2];
  float h2 = dt * 0.5f;
  float h6 = dt / 6.0f;
  float loss = 0.0f;
  for (int k = 0; k < P->n_obs; ++k) {
    if (k > 0) {
      for (int s = 0; s < P->steps_per_obs; ++s) {
        float k1a, k1b, k2a, k2b, k3a, k3b, k4a, k4b;
        ode_rhs(model, u1, u2, th, &k1a, &k1b);
        ode_rhs(model, u1 + h2 * k1a, u2 + h2 * k1b, th, &k2a, &k2b);
        ode_rhs(model, u1 + h2 * k2a, u2 + h2 * k2b, th, &k3a, &k3b);
        ode_rhs(model, u1 + dt * k3a, u2 + dt * k3b, th, &k4a, &k4b);
        u1 = u1 + h6 * (((k1a + 2.0f * k2a) + 2.0f * k3a) + k4a);
        u2 = u2 + h6 * (((k1b + 2.0f * k2b) + 2.0f * k3b) + k4b);
      }
    }
    float e1 = P->obs[2 * k] - u1;
    float e2 = P->obs[2 * k + 1] - u2;
    loss = loss + (e1 * e1 + e2 * e2);
  }
  return loss;
}

/* f(x - shift).  Rosenbrock in the loop form of test/gpu.jl:14-20:
 * res += p[2] * (x[i+1] - x[i]^2)^2 + (p[1] - x[i])^2. */
float ora_objective(const ora_problem* P, const float* xin) {
  float z[ORA_MAXD];
  int d = P->d;
  for (int j = 0; j < d; ++j) z[j] = P->has_shift ? xin[j] - P->shift[j] : xin[j];
  switch (P->obj) {
    case ORA_ROSENBROCK: {
      float res = 0.0f;
      for (int i = 0; i + 1 < d; ++i) {
        float dd = z[i + 1] - z[i] * z[i];
        float e = P->par[0] - z[i];
        res = res + (P->par[1] * (dd * dd) + e * e);
      }
      return res;
    }
    case ORA_RASTRIGIN: {
      float res = 10.0f * (float)d;
      for (int i = 0; i < d; ++i) res = res + (z[i] * z[i] - 10.0f * ora_cos2pi(z[i]));
      return res;
    }
    case ORA_SPHERE: {
      float res = 0.0f;
      for (int i = 0; i < d; ++i) res = res + z[i] * z[i];
      return res;
    }
    case ORA_BEALE: {
      float x1 = z[0], x2 = z[1];
      float a = (1.5f - x1) + x1 * x2;
      float b = (2.25f - x1) + x1 * (x2 * x2);
      float c = (2.625f - x1) + x1 * ((x2 * x2) * x2);
      return (a * a + b * b) + c * c;
    }
    case ORA_BOOTH: {
      float a = (z[0] + 2.0f * z[1]) - 7.0f, b = (2.0f * z[0] + z[1]) - 5.0f;
      return a * a + b * b;
    }
    case ORA_HIMMELBLAU: {
      float a = (z[0] * z[0] + z[1]) - 11.0f, b = (z[0] + z[1] * z[1]) - 7.0f;
      return a * a + b * b;
    }
    case ORA_QUADRATIC: {
      float a = z[0] - 1.0f, b = z[1] + 2.0f;
      return P->par[0] * (a * a) + b * b;
    }
    case ORA_BOUNDARY_QUADRATIC: {
      float a = z[0] + 1.0f;
      return a * a;
    }
    case ORA_ROSEN2: {
      float a = P->par[0] - z[0], b = z[1] - z[0] * z[0];
      return a * a + P->par[1] * (b * b);
    }
    default:
      if (P->obj < ORA_LOTKA_VOLTERRA || P->obj > ORA_FITZHUGH_NAGUMO_ATSIT5_SAVEAT) break;
      switch (ode_solver(P->obj)) {
        case 0: return ora_ode_sse_rk4(z, P);
        case 1: return ora_ode_sse_tsit5(z, P, 0);
        case 2: return ora_ode_sse_tsit5(z, P, 1);
        case 3: return ora_ode_sse_tsit5_saveat(z, P, 0);
        default: return ora_ode_sse_tsit5_saveat(z, P, 1);
      }
  }
  return NAN;
}

/* ---------------------------------------------------------- particle AoS */
#define POS(p, d) (p)
#define VEL(p, d) ((p) + (d))
#define COST(p, d) ((p)[2 * (d)])
#define BPOS(p, d) ((p) + 2 * (d) + 1)
#define BCOST(p, d) ((p)[3 * (d) + 1])
static inline int64_t stride_of(int d) { return 3 * (int64_t)d + 2; }

/* Argmin of best_cost, ties to the lowest index, NaN never below a number
 * (the key the kernels reduce: order(cost) << 32 | index). */
static inline int better(float a, int64_t ia, float b, int64_t ib) {
  int an = isnan(a), bn = isnan(b);
  if (an != bn) return bn;           /* number beats NaN */
  if (!an && a != b) return a < b;   /* +0 == -0 */
  return ia < ib;
}
int64_t ora_argmin(const float* P, int64_t n, int d) {
  int64_t s = stride_of(d), best = 0;
  for (int64_t i = 1; i < n; ++i)
    if (better(BCOST(P + i * s, d), i, BCOST(P + best * s, d), best)) best = i;
  return best;
}

/* update_particle_state, src/kernels.jl:1-26 (velocity_position part):
 *   v = w .* v .+ c1 .* r1 .* (best - x) .+ c2 .* r2 .* (gbest - x)
 *   x = x + v ; x = min.(max.(x, lb), ub)          (per-component clamp)
 *   or min(max(x, lb), ub) on whole vectors        (ODE clamp, ode_pso.jl:18-19)
 * r1[j], r2[j] from philox(ctr = (i, t, floor(j/2) | 0 << 16, epoch)):
 * lanes (0,1) for even j, lanes (2,3) for odd j. */
static void move_particle(float* p, const ora_problem* P, const float* gb, float w,
                          float c1, float c2, uint32_t i, uint32_t t, uint32_t epoch,
                          const uint32_t key[2]) {
  int d = P->d;
  float* x = POS(p, d);
  float* v = VEL(p, d);
  const float* pb = BPOS(p, d);
  uint32_t r[4];
  for (int j = 0; j < d; ++j) {
    if ((j & 1) == 0) {
      uint32_t ctr[4] = {i, t, (uint32_t)(j >> 1), epoch};
      ora_philox(ctr, key, r);
    }
    float r1 = u01(r[(j & 1) ? 2 : 0]);
    float r2 = u01(r[(j & 1) ? 3 : 1]);
    float a = w * v[j];
    float b = (c1 * r1) * (pb[j] - x[j]);
    float c = (c2 * r2) * (gb[j] - x[j]);
    v[j] = (a + b) + c;
  }
  for (int j = 0; j < d; ++j) x[j] = x[j] + v[j];
  if (P->clamp == ORA_CLAMP_COMPONENT) {
    for (int j = 0; j < d; ++j) x[j] = jl_max(x[j], P->lb[j]);
    for (int j = 0; j < d; ++j) x[j] = jl_min(x[j], P->ub[j]);
  } else {
    /* max(x, lb) = isless(lb, x) ? x : lb ; min(y, ub) = isless(ub, y) ? ub : y */
    if (!(jl_cmp_vec(P->lb, x, d) < 0)) memcpy(x, P->lb, sizeof(float) * d);
    if (jl_cmp_vec(P->ub, x, d) < 0) memcpy(x, P->ub, sizeof(float) * d);
  }
}

/* Constraint penalty, calc_penalty (src/utils.jl:353-372) with the default
 * schedules theta_default / gamma_default (src/utils.jl:320-336) and h = sqrt
 * (src/algorithms.jl:88-89):
 *   q = max.(cons(x), 0); penalty = sum(theta.(q));
 *   for i: penalty += theta(q_i) * q_i^gamma(q_i);  penalty = h(iter) * penalty
 * theta's thresholds are Float64 literals (0.001, 0.1, 1.0) compared after
 * promotion; gamma's is the integer 1.  q^1 = q and q^2 = q*q (the correctly
 * rounded power; the reference's @fastmath pow of a Float32 by 1f0 or 2f0).
 * cons(x) is the separable quadratic family: nonzero linear terms in index
 * order, then nonzero quadratic terms, then + c (for the reference's test
 * constraints, test/constraints.jl:11-14, their own evaluation order). */
static float theta_default(float q) {
  double qd = (double)q;
  if (qd < 0.001) return 10.0f;
  if (qd <= 0.1) return 20.0f;
  if (qd <= 1.0) return 100.0f;
  return 300.0f;
}

float ora_constraint(const ora_problem* P, int k, const float* x) {
  int d = P->d;
  const float* row = P->cons + (size_t)k * (2 * d + 1);
  float g = 0.0f;
  for (int j = 0; j < d; ++j)
    if (row[j] != 0.0f) g = g + row[j] * x[j];
  for (int j = 0; j < d; ++j)
    if (row[d + j] != 0.0f) g = g + row[d + j] * (x[j] * x[j]);
  return g + row[2 * d];
}

float ora_penalty(const ora_problem* P, const float* x, float iter) {
  float q[16];
  for (int k = 0; k < P->n_cons; ++k) q[k] = jl_max(ora_constraint(P, k, x), 0.0f);
  float pen = theta_default(q[0]);
  for (int k = 1; k < P->n_cons; ++k) pen = pen + theta_default(q[k]);
  for (int k = 0; k < P->n_cons; ++k) {
    float pw = q[k] < 1.0f ? q[k] : q[k] * q[k];
    pen = pen + theta_default(q[k]) * pw;
  }
  return sqrtf(iter) * pen;
}

/* handle_constraints (src/kernels.jl:28-36): the objective, plus the penalty
 * at h's argument `iter` when the problem has constraints. */
float ora_cost(const ora_problem* P, const float* x, float iter) {
  float c = ora_objective(P, x);
  if (P->n_cons > 0) c = c + ora_penalty(P, x, iter);
  return c;
}

/* cost = f(x); if cost < best_cost: best = x (src/kernels.jl:19-24). */
static void evaluate_particle(float* p, const ora_problem* P, float iter) {
  int d = P->d;
  COST(p, d) = ora_cost(P, POS(p, d), iter);
  if (COST(p, d) < BCOST(p, d)) {
    memcpy(BPOS(p, d), POS(p, d), sizeof(float) * d);
    BCOST(p, d) = COST(p, d);
  }
}

/* `iter` is h's argument as the reference computes it at the call site:
 * 1-based particle index + 1 in the sync and queue-lock kernels
 * (src/kernels.jl:77,151,191 pass the index i as `iter`), 1-based iteration of
 * the launch + 1 in the async kernel (src/kernels.jl:213-214). */
static void step_particle(float* p, const ora_problem* P, const float* gb, float w,
                          float c1, float c2, uint32_t i, uint32_t t, uint32_t epoch,
                          const uint32_t key[2], float iter) {
  move_particle(p, P, gb, w, c1, c2, i, t, epoch, key);
  evaluate_particle(p, P, iter);
}

/* Keyed permutation of [0, n) of the Latin hypercube sampler (the design's
 * definition, restated by the library's lhs_permute): balanced 4-round
 * Feistel network on b bits (b even, 2^b >= n), round function
 * philox(ctr = (right half, round, j | 2 << 16, epoch)) word 0 masked to b/2
 * bits, (L, R) -> (R, L ^ F(R)), iterated until the value is below n. */
uint32_t ora_lhs_permute(uint32_t i, uint32_t n, uint32_t j, uint32_t epoch, uint64_t seed) {
  uint32_t key[2];
  seed_key(seed, key);
  uint32_t b = 2;
  while (b < 32 && ((uint64_t)1 << b) < (uint64_t)n) b += 2;
  uint32_t h = b / 2, mask = (uint32_t)(((uint64_t)1 << h) - 1);
  uint32_t x = i;
  do {
    uint32_t L = x >> h, R = x & mask;
    for (uint32_t r = 0; r < 4; ++r) {
      uint32_t ctr[4] = {R, r, j | (2u << 16), epoch}, out[4];
      ora_philox(ctr, key, out);
      uint32_t t = L ^ (out[0] & mask);
      L = R;
      R = t;
    }
    x = (L << h) | R;
  } while (x >= n);
  return x;
}

/* Initial position of particle i under a sampler, random words from
 * philox(ctr = (i, 0, floor(j/4) | tag << 16, epoch)) lane j mod 4:
 *   uniform (src/utils.jl:1-5,57-84), tag 1: x_j = r * (ub_j - lb_j) + lb_j;
 *   unbounded (src/utils.jl:23-55), tag 1: x_j = u0_j + r * |u0_j| if
 *     |u0_j| > 0, else r (the reference's ifelse draws twice; one counter-based
 *     draw per coordinate here);
 *   Latin hypercube, tag 3: t = (pi_j(i) + r) / n, x_j = t * (ub_j - lb_j) + lb_j. */
static void sample_position(const ora_problem* P, int sampler, int64_t n, int64_t i,
                            uint32_t epoch, uint64_t seed, const uint32_t key[2], float* x) {
  int d = P->d;
  uint32_t tag = sampler == ORA_SAMPLER_LATIN_HYPERCUBE ? 3u : 1u;
  uint32_t r[4];
  for (int j = 0; j < d; ++j) {
    if ((j & 3) == 0) {
      uint32_t ctr[4] = {(uint32_t)i, 0u, (uint32_t)(j >> 2) | (tag << 16), epoch};
      ora_philox(ctr, key, r);
    }
    float u = u01(r[j & 3]);
    if (sampler == ORA_SAMPLER_UNBOUNDED) {
      float c = P->u0[j], a = fabsf(c);
      x[j] = a > 0.0f ? c + u * a : u;
    } else if (sampler == ORA_SAMPLER_LATIN_HYPERCUBE) {
      uint32_t pi = ora_lhs_permute((uint32_t)i, (uint32_t)n, (uint32_t)j, epoch, seed);
      float t = ((float)pi + u) / (float)n;
      x[j] = t * (P->ub[j] - P->lb[j]) + P->lb[j];
    } else {
      x[j] = u * (P->ub[j] - P->lb[j]) + P->lb[j];
    }
  }
}

/* init (src/init.jl:15-54, src/utils.jl:23-112): positions from the sampler,
 * or `positions` (n x d, particle-major; the QuasiMonteCarlo path) when not
 * NULL; v = 0; cost = f(x) [+ penalty at h(1)]; best = x.  The init global
 * best is (best_position, best_cost) of the argmin (src/init.jl:49-51). */
void ora_init(const ora_problem* P, int64_t n, uint64_t seed, uint32_t epoch, float* parts,
              float* gb_pos, float* gb_cost, int64_t* gb_idx, int nthreads, int32_t sampler,
              const float* positions) {
  int d = P->d;
  int64_t s = stride_of(d);
  uint32_t key[2];
  seed_key(seed, key);
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
  for (int64_t i = 0; i < n; ++i) {
    float* p = parts + i * s;
    if (positions)
      memcpy(POS(p, d), positions + i * d, sizeof(float) * d);
    else
      sample_position(P, sampler, n, i, epoch, seed, key, POS(p, d));
    for (int j = 0; j < d; ++j) VEL(p, d)[j] = 0.0f;
    COST(p, d) = ora_cost(P, POS(p, d), 1.0f); /* calc_penalty(.., 1, ..), src/utils.jl:72-73 */
    memcpy(BPOS(p, d), POS(p, d), sizeof(float) * d);
    BCOST(p, d) = COST(p, d);
  }
  int64_t b = ora_argmin(parts, n, d);
  memcpy(gb_pos, BPOS(parts + b * s, d), sizeof(float) * d);
  *gb_cost = BCOST(parts + b * s, d);
  *gb_idx = b;
}

/* Sync PSO (ParallelSyncPSOKernel): every iteration steps all particles
 * against the previous global best (src/kernels.jl:182-194 on the CPU
 * backend, :127-177 on a GPU backend), then takes the argmin of best_cost and
 * sets gbest = (position, best_cost) on the CPU backend
 * (src/lowerlevel_solve.jl:59-60) or (best_position, best_cost) on a GPU
 * backend (src/kernels.jl:153,171); w = w * wdamp after each iteration.
 * Iteration k of this call uses random-stream iteration t0 + k (k = 1..).
 * gb_pos holds the attractor on entry and the final global best on exit. */
void ora_sync_run(const ora_problem* P, int64_t n, uint64_t seed, uint32_t epoch, uint32_t t0,
                  int32_t iters, float w, float wdamp, float c1, float c2, int32_t gbest_from,
                  float* parts, float* gb_pos, float* gb_cost, int64_t* gb_idx, int nthreads) {
  int d = P->d;
  int64_t s = stride_of(d);
  uint32_t key[2];
  seed_key(seed, key);
  float gb[ORA_MAXD];
  memcpy(gb, gb_pos, sizeof(float) * d);
  for (int32_t k = 1; k <= iters; ++k) {
    uint32_t t = t0 + (uint32_t)k;
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
    for (int64_t i = 0; i < n; ++i)
      step_particle(parts + i * s, P, gb, w, c1, c2, (uint32_t)i, t, epoch, key,
                    (float)(i + 2));
    int64_t b = ora_argmin(parts, n, d);
    const float* src = gbest_from == ORA_GBEST_CURRENT ? POS(parts + b * s, d)
                                                       : BPOS(parts + b * s, d);
    memcpy(gb, src, sizeof(float) * d);
    *gb_cost = BCOST(parts + b * s, d);
    *gb_idx = b;
    w = w * wdamp;
  }
  memcpy(gb_pos, gb, sizeof(float) * d);
}

/* Async PSO (ParallelPSOKernel, global_update = false; src/kernels.jl:196-225):
 * each particle runs all iterations against its own running best, seeded with
 * the global best at launch and replaced whenever best_cost < its cost; the
 * result is the argmin over particles of (best_position, best_cost)
 * (src/lowerlevel_solve.jl:119-120). */
void ora_async_run(const ora_problem* P, int64_t n, uint64_t seed, uint32_t epoch, uint32_t t0,
                   int32_t iters, float w0, float wdamp, float c1, float c2, float* parts,
                   float* gb_pos, float* gb_cost, int64_t* gb_idx, int nthreads) {
  int d = P->d;
  int64_t s = stride_of(d);
  uint32_t key[2];
  seed_key(seed, key);
  float g0[ORA_MAXD];
  memcpy(g0, gb_pos, sizeof(float) * d);
  float g0c = *gb_cost;
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
  for (int64_t i = 0; i < n; ++i) {
    float* p = parts + i * s;
    float g[ORA_MAXD];
    memcpy(g, g0, sizeof(float) * d);
    float gc = g0c;
    float w = w0;
    for (int32_t k = 1; k <= iters; ++k) {
      step_particle(p, P, g, w, c1, c2, (uint32_t)i, t0 + (uint32_t)k, epoch, key,
                    (float)(k + 1));
      if (BCOST(p, d) < gc) {
        memcpy(g, BPOS(p, d), sizeof(float) * d);
        gc = BCOST(p, d);
      }
      w = w * wdamp;
    }
  }
  int64_t b = ora_argmin(parts, n, d);
  memcpy(gb_pos, BPOS(parts + b * s, d), sizeof(float) * d);
  *gb_cost = BCOST(parts + b * s, d);
  *gb_idx = b;
}

/* Queue-lock PSO (ParallelPSOKernel, global_update = true; src/kernels.jl:38-125):
 * per iteration, particles are processed in blocks of `block`; each block
 * steps against the live global best, queues the particles whose best_cost
 * beats it, and when at least `min_improvers` queued (the reference's
 * `queue_num[1] > 1`) publishes the queue minimum if it still beats the global
 * best.  Blocks run one after another in index order; the queue is in particle
 * order and its minimum keeps the first of equal costs. */
void ora_qlock_run(const ora_problem* P, int64_t n, uint64_t seed, uint32_t epoch, uint32_t t0,
                   int32_t iters, float w, float wdamp, float c1, float c2, int32_t block,
                   int32_t min_improvers, float* parts, float* gb_pos, float* gb_cost,
                   int64_t* gb_idx) {
  int d = P->d;
  int64_t s = stride_of(d);
  uint32_t key[2];
  seed_key(seed, key);
  for (int32_t k = 1; k <= iters; ++k) {
    uint32_t t = t0 + (uint32_t)k;
    for (int64_t b0 = 0; b0 < n; b0 += block) {
      int64_t b1 = b0 + block < n ? b0 + block : n;
      float gb[ORA_MAXD];
      memcpy(gb, gb_pos, sizeof(float) * d);
      for (int64_t i = b0; i < b1; ++i)
        step_particle(parts + i * s, P, gb, w, c1, c2, (uint32_t)i, t, epoch, key,
                      (float)(i + 2));
      int64_t count = 0, best = -1;
      for (int64_t i = b0; i < b1; ++i) {
        float bc = BCOST(parts + i * s, d);
        if (bc < *gb_cost) {
          ++count;
          if (best < 0 || bc < BCOST(parts + best * s, d)) best = i;
        }
      }
      if (count >= min_improvers && BCOST(parts + best * s, d) < *gb_cost) {
        memcpy(gb_pos, BPOS(parts + best * s, d), sizeof(float) * d);
        *gb_cost = BCOST(parts + best * s, d);
        *gb_idx = best;
      }
    }
    w = w * wdamp;
  }
}

/* Evaluate the objective over AoS positions (n x d). */
void ora_eval(const ora_problem* P, int64_t n, const float* x, float* f) {
  for (int64_t i = 0; i < n; ++i) f[i] = ora_objective(P, x + i * P->d);
}

/* One velocity/position move without evaluation (the ODE path's
 * ode_update_particle_states!, src/ode_pso.jl:1-24), exposed for unit tests. */
void ora_move(const ora_problem* P, float* particle, const float* gb, float w, float c1,
              float c2, uint32_t i, uint32_t t, uint32_t epoch, uint64_t seed) {
  uint32_t key[2];
  seed_key(seed, key);
  move_particle(particle, P, gb, w, c1, c2, i, t, epoch, key);
}

int ora_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ===================================================================== */
/* HybridPSO local phase: projected L-BFGS with a Strong-Wolfe or an Armijo
 * line search.
 *
 * The reference runs SimpleOptimization._lbfgs (projected L-BFGS with a
 * Strong-Wolfe line search from LineSearch.jl, whose c1, c2, alpha_init,
 * alpha_max, maxiters and zoom_maxiters it forwards, src/hybrid.jl:129-136)
 * from every particle's best_position (src/lbfgs.jl:9-21,
 * src/hybrid.jl:100-158); neither package is vendored under the reference,
 * so this restates the published algorithms (Nocedal & Wright, Numerical
 * Optimization: Alg. 7.4 two-loop recursion, Alg. 7.5 L-BFGS, Alg. 3.5 line
 * search for the strong Wolfe conditions with its zoom, Alg. 3.6) on the
 * projected path x(a) = P(x + a p), with an analytic gradient.  The Armijo
 * backtracking search is the benchmark configuration's.  Non-finite final
 * values become +Inf (src/lbfgs.jl:20).
 *
 * The kernel gives each particle a group of G = 2^ceil(log2 d) lanes, one
 * coordinate per lane; every dot product and objective sum is an xor-butterfly
 * over the group (step s: v[l] = v[l] + v[l ^ s] for s = G/2 .. 1), which this
 * restatement reproduces operation for operation.
 * ===================================================================== */
#define ORA_LB_MAXH 16

#define ORA_LS_ARMIJO 0
#define ORA_LS_STRONG_WOLFE 1

typedef struct ora_lbfgs_params {
  int32_t maxiters, history;
  float abstol, armijo_c, armijo_rho;
  int32_t max_backtracks;
  int32_t linesearch;    /* ORA_LS_* */
  float wolfe_c2, alpha_init, alpha_max;
  int32_t ls_maxiters, zoom_maxiters;
  float reltol;          /* stop after a step with |f - f_new| <= reltol |f| (0: off) */
} ora_lbfgs_params;

static int group_size(int d) {
  int G = 1;
  while (G < d) G <<= 1;
  return G;
}

/* xor-butterfly sum over G lanes (every lane ends with this value) */
static float tsum(const float* v, int G) {
  float t[ORA_MAXD], u[ORA_MAXD];
  memcpy(t, v, sizeof(float) * G);
  for (int off = G / 2; off >= 1; off >>= 1) {
    for (int l = 0; l < G; ++l) u[l] = t[l] + t[l ^ off];
    memcpy(t, u, sizeof(float) * G);
  }
  return t[0];
}
static float tdot(const float* a, const float* b, int G) {
  float t[ORA_MAXD];
  for (int l = 0; l < G; ++l) t[l] = a[l] * b[l];
  return tsum(t, G);
}
static float tmaxabs(const float* a, int G) {
  float m = 0.0f;
  for (int l = 0; l < G; ++l) m = fmaxf(m, fabsf(a[l]));
  return m;
}

/* Objective in lane-group form: separable / neighbour objectives add their
 * per-lane terms with the butterfly; fixed two-dimensional objectives use the
 * sequential formula.  Gradient per lane (analytic). */
static void lane_fg(const ora_problem* P, const float* xin, int G, float* f, float* g) {
  int d = P->d;
  float z[ORA_MAXD], term[ORA_MAXD];
  for (int l = 0; l < G; ++l) {
    z[l] = l < d ? (P->has_shift ? xin[l] - P->shift[l] : xin[l]) : 0.0f;
    term[l] = 0.0f;
    g[l] = 0.0f;
  }
  const float p1 = P->par[0], p2 = P->par[1];
  switch (P->obj) {
    case ORA_ROSENBROCK:
      for (int l = 0; l < d; ++l) {
        float gl = 0.0f;
        if (l + 1 < d) {
          float dd = z[l + 1] - z[l] * z[l];
          float e = p1 - z[l];
          term[l] = p2 * (dd * dd) + e * e;
          gl = (-4.0f * p2) * z[l] * dd - 2.0f * e;
        }
        if (l > 0) {
          float dp = z[l] - z[l - 1] * z[l - 1];
          gl = gl + (2.0f * p2) * dp;
        }
        g[l] = gl;
      }
      *f = tsum(term, G);
      return;
    case ORA_RASTRIGIN:
      for (int l = 0; l < d; ++l) {
        term[l] = z[l] * z[l] - 10.0f * ora_cos2pi(z[l]);
        g[l] = 2.0f * z[l] + 62.83185307179586f * ora_sin2pi(z[l]);
      }
      *f = 10.0f * (float)d + tsum(term, G);
      return;
    case ORA_SPHERE:
      for (int l = 0; l < d; ++l) {
        term[l] = z[l] * z[l];
        g[l] = 2.0f * z[l];
      }
      *f = tsum(term, G);
      return;
    default:
      break;
  }
  /* fixed-dimension objectives: value by the sequential formula */
  *f = ora_objective(P, xin);
  float x1 = z[0], x2 = d > 1 ? z[1] : 0.0f;
  switch (P->obj) {
    case ORA_BEALE: {
      float a = (1.5f - x1) + x1 * x2;
      float b = (2.25f - x1) + x1 * (x2 * x2);
      float c = (2.625f - x1) + x1 * ((x2 * x2) * x2);
      g[0] = 2.0f * a * (x2 - 1.0f) + 2.0f * b * (x2 * x2 - 1.0f) + 2.0f * c * ((x2 * x2) * x2 - 1.0f);
      g[1] = 2.0f * a * x1 + 2.0f * b * (2.0f * x1 * x2) + 2.0f * c * (3.0f * x1 * (x2 * x2));
      break;
    }
    case ORA_BOOTH: {
      float a = (x1 + 2.0f * x2) - 7.0f, b = (2.0f * x1 + x2) - 5.0f;
      g[0] = 2.0f * a + 4.0f * b;
      g[1] = 4.0f * a + 2.0f * b;
      break;
    }
    case ORA_HIMMELBLAU: {
      float a = (x1 * x1 + x2) - 11.0f, b = (x1 + x2 * x2) - 7.0f;
      g[0] = 4.0f * x1 * a + 2.0f * b;
      g[1] = 2.0f * a + 4.0f * x2 * b;
      break;
    }
    case ORA_QUADRATIC:
      g[0] = 2.0f * p1 * (x1 - 1.0f);
      g[1] = 2.0f * (x2 + 2.0f);
      break;
    case ORA_BOUNDARY_QUADRATIC:
      g[0] = 2.0f * (x1 + 1.0f);
      break;
    case ORA_ROSEN2: {
      float a = p1 - x1, b = x2 - x1 * x1;
      g[0] = -2.0f * a - 4.0f * p2 * x1 * b;
      g[1] = 2.0f * p2 * b;
      break;
    }
    default:
      for (int l = 0; l < G; ++l) g[l] = NAN;
  }
}

static inline float proj(float x, float lo, float hi) { return jl_min(jl_max(x, lo), hi); }

/* Line search on phi(a) = f(P(x + a p)).  phi'(0) counts the lanes the
 * projection lets move along p (not on a bound that p points out of);
 * phi'(a), a > 0, the lanes whose unprojected x + a p lies strictly inside
 * the box (the one-sided derivative of the projected path). */
static float dphi_at0(const float* x, const float* g, const float* p, const float* lo,
                      const float* hi, int G) {
  float t[ORA_MAXD];
  for (int l = 0; l < G; ++l)
    t[l] = ((x[l] <= lo[l] && p[l] < 0.0f) || (x[l] >= hi[l] && p[l] > 0.0f)) ? 0.0f
                                                                              : g[l] * p[l];
  return tsum(t, G);
}
static float dphi_at(const float* x, const float* gt, const float* p, const float* lo,
                     const float* hi, float a, int G) {
  float t[ORA_MAXD];
  for (int l = 0; l < G; ++l) {
    float u = x[l] + a * p[l];
    t[l] = (lo[l] < u && u < hi[l]) ? gt[l] * p[l] : 0.0f;
  }
  return tsum(t, G);
}
static void ls_eval(const ora_problem* P, const float* x, const float* p, const float* lo,
                    const float* hi, float a, int G, float* xt, float* ft, float* gt) {
  for (int l = 0; l < G; ++l) xt[l] = proj(x[l] + a * p[l], lo[l], hi[l]);
  lane_fg(P, xt, G, ft, gt);
}

/* Strong-Wolfe search (Nocedal & Wright Alg. 3.5/3.6): accept a with
 *   phi(a) <= phi(0) + c1 a phi'(0)   and   |phi'(a)| <= c2 |phi'(0)|.
 * Trial steps alpha_init, doubled up to alpha_max (at most ls_maxiters of
 * them) until one of them brackets an acceptable step; zoom's trial is the
 * minimiser of the quadratic through phi(lo), phi'(lo) and phi(hi) when it
 * lies in the middle 80 % of the bracket, else its midpoint.  When no step
 * meets both conditions within the iteration limits, the last step of
 * sufficient decrease found is taken (none: the search fails, returns 0).
 * Returns 1 with the accepted point in xn, fn, gn. */
int ora_wolfe_search(const ora_problem* P, const ora_lbfgs_params* lp, int G, const float* x,
                     float f, const float* g, const float* p, const float* lo, const float* hi,
                     float* xn, float* fn, float* gn, float* alpha) {
  const float d0 = dphi_at0(x, g, p, lo, hi, G);
  if (!(d0 < 0.0f)) return 0;
  const float c1 = lp->armijo_c, curv = -(lp->wolfe_c2 * d0);
  float xt[ORA_MAXD], gt[ORA_MAXD], ft;
  float xl[ORA_MAXD], gl[ORA_MAXD];   /* point at a_lo (a_prev before a bracket) */
  memcpy(xl, x, sizeof(float) * G);
  memcpy(gl, g, sizeof(float) * G);
  float a_lo = 0.0f, f_lo = f, d_lo = d0, a_hi = 0.0f, f_hi = 0.0f;
  float a = lp->alpha_init;
  int bracket = 0;
  for (int i = 0; i < lp->ls_maxiters && !bracket; ++i) {
    ls_eval(P, x, p, lo, hi, a, G, xt, &ft, gt);
    if (ft > f + (c1 * a) * d0 || (i > 0 && ft >= f_lo)) {
      a_hi = a;
      f_hi = ft;
      bracket = 1;
      break;
    }
    const float da = dphi_at(x, gt, p, lo, hi, a, G);
    if (fabsf(da) <= curv) goto accept_t;
    if (da >= 0.0f) {
      a_hi = a_lo;
      f_hi = f_lo;
      a_lo = a;
      f_lo = ft;
      d_lo = da;
      memcpy(xl, xt, sizeof(float) * G);
      memcpy(gl, gt, sizeof(float) * G);
      bracket = 1;
      break;
    }
    a_lo = a;
    f_lo = ft;
    d_lo = da;
    memcpy(xl, xt, sizeof(float) * G);
    memcpy(gl, gt, sizeof(float) * G);
    if (!(a < lp->alpha_max)) break;
    a = a + a;
    if (a > lp->alpha_max) a = lp->alpha_max;
  }
  if (bracket) {
    for (int j = 0; j < lp->zoom_maxiters; ++j) {
      const float delta = a_hi - a_lo;
      const float den = 2.0f * ((f_hi - f_lo) - d_lo * delta);
      a = a_lo - ((d_lo * delta) * delta) / den;
      const float b0 = fminf(a_lo, a_hi), b1 = fmaxf(a_lo, a_hi), w = b1 - b0;
      if (!(a >= b0 + 0.1f * w && a <= b1 - 0.1f * w)) a = a_lo + 0.5f * delta;
      ls_eval(P, x, p, lo, hi, a, G, xt, &ft, gt);
      if (ft > f + (c1 * a) * d0 || ft >= f_lo) {
        a_hi = a;
        f_hi = ft;
      } else {
        const float da = dphi_at(x, gt, p, lo, hi, a, G);
        if (fabsf(da) <= curv) goto accept_t;
        if (da * (a_hi - a_lo) >= 0.0f) {
          a_hi = a_lo;
          f_hi = f_lo;
        }
        a_lo = a;
        f_lo = ft;
        d_lo = da;
        memcpy(xl, xt, sizeof(float) * G);
        memcpy(gl, gt, sizeof(float) * G);
      }
    }
  }
  if (!(a_lo > 0.0f)) return 0;
  memcpy(xn, xl, sizeof(float) * G);
  memcpy(gn, gl, sizeof(float) * G);
  *fn = f_lo;
  if (alpha) *alpha = a_lo;
  return 1;
accept_t:
  memcpy(xn, xt, sizeof(float) * G);
  memcpy(gn, gt, sizeof(float) * G);
  *fn = ft;
  if (alpha) *alpha = a;
  return 1;
}

/* Armijo backtracking: the first a = rho^k (k <= max_backtracks) with
 * f(P(x + a p)) <= f + c g'(P(x + a p) - x). */
static int armijo_search(const ora_problem* P, const ora_lbfgs_params* lp, int G, const float* x,
                         float f, const float* g, const float* p, const float* lo,
                         const float* hi, float* xn, float* fn, float* gn) {
  float alpha = 1.0f;
  for (int bt = 0; bt <= lp->max_backtracks; ++bt) {
    float dx[ORA_MAXD];
    for (int l = 0; l < G; ++l) {
      xn[l] = proj(x[l] + alpha * p[l], lo[l], hi[l]);
      dx[l] = xn[l] - x[l];
    }
    lane_fg(P, xn, G, fn, gn);
    float decr = tdot(g, dx, G);
    if (*fn <= f + lp->armijo_c * decr) return 1;
    alpha = alpha * lp->armijo_rho;
  }
  return 0;
}

/* One particle: x (d floats, in/out), returns the final value. */
float ora_lbfgs_one(const ora_problem* P, const ora_lbfgs_params* lp, float* xio) {
  int d = P->d, G = group_size(d), m = lp->history;
  float lo[ORA_MAXD], hi[ORA_MAXD], x[ORA_MAXD], g[ORA_MAXD], f;
  float S[ORA_LB_MAXH][ORA_MAXD], Y[ORA_LB_MAXH][ORA_MAXD], RHO[ORA_LB_MAXH];
  for (int l = 0; l < G; ++l) {
    lo[l] = l < d ? P->lb[l] : 0.0f;
    hi[l] = l < d ? P->ub[l] : 0.0f;
    x[l] = proj(l < d ? xio[l] : 0.0f, lo[l], hi[l]);
  }
  lane_fg(P, x, G, &f, g);
  int k = 0; /* stored pairs, newest at index 0 */
  for (int it = 0; it < lp->maxiters; ++it) {
    float pg[ORA_MAXD];
    for (int l = 0; l < G; ++l) pg[l] = proj(x[l] - g[l], lo[l], hi[l]) - x[l];
    float pgn = tmaxabs(pg, G);
    if (!(pgn > lp->abstol)) break;
    float q[ORA_MAXD], r[ORA_MAXD], p[ORA_MAXD], a[ORA_LB_MAXH];
    memcpy(q, g, sizeof(float) * G);
    for (int i = 0; i < k; ++i) {
      a[i] = RHO[i] * tdot(S[i], q, G);
      for (int l = 0; l < G; ++l) q[l] = q[l] - a[i] * Y[i][l];
    }
    float gamma = 1.0f;
    if (k > 0) gamma = tdot(S[0], Y[0], G) / tdot(Y[0], Y[0], G);
    for (int l = 0; l < G; ++l) r[l] = gamma * q[l];
    for (int i = k - 1; i >= 0; --i) {
      float b = RHO[i] * tdot(Y[i], r, G);
      for (int l = 0; l < G; ++l) r[l] = r[l] + (a[i] - b) * S[i][l];
    }
    for (int l = 0; l < G; ++l) p[l] = -r[l];
    float gp = tdot(g, p, G);
    if (!(gp < 0.0f)) {
      for (int l = 0; l < G; ++l) p[l] = -g[l];
      k = 0;
    }
    /* line search; if the quasi-Newton direction fails, forget the history
     * and search once more along -g */
    float fn = 0.0f, xn[ORA_MAXD], gn[ORA_MAXD];
    int accepted = 0;
    for (int attempt = 0; attempt < 2 && !accepted; ++attempt) {
      if (attempt == 1) {
        if (k == 0) break;
        for (int l = 0; l < G; ++l) p[l] = -g[l];
        k = 0;
      }
      accepted = lp->linesearch == ORA_LS_STRONG_WOLFE
                     ? ora_wolfe_search(P, lp, G, x, f, g, p, lo, hi, xn, &fn, gn, NULL)
                     : armijo_search(P, lp, G, x, f, g, p, lo, hi, xn, &fn, gn);
    }
    if (!accepted) break;
    float s[ORA_MAXD], y[ORA_MAXD];
    for (int l = 0; l < G; ++l) {
      s[l] = xn[l] - x[l];
      y[l] = gn[l] - g[l];
    }
    float ss = tmaxabs(s, G);
    float sy = tdot(s, y, G), yy = tdot(y, y, G);
    if (sy > 1e-10f * yy && sy > 0.0f) {
      int top = k < m ? k : m - 1;
      for (int i = top; i >= 1; --i) {
        memcpy(S[i], S[i - 1], sizeof(float) * G);
        memcpy(Y[i], Y[i - 1], sizeof(float) * G);
        RHO[i] = RHO[i - 1];
      }
      memcpy(S[0], s, sizeof(float) * G);
      memcpy(Y[0], y, sizeof(float) * G);
      RHO[0] = 1.0f / sy;
      if (k < m) ++k;
    } else {
      k = 0; /* no positive curvature along s: restart from a scaled gradient */
    }
    int stall = lp->reltol > 0.0f && fabsf(f - fn) <= lp->reltol * fabsf(f);
    memcpy(x, xn, sizeof(float) * G);
    memcpy(g, gn, sizeof(float) * G);
    f = fn;
    /* projection pinned the step (no progress possible), or a relative
     * decrease below reltol */
    if (!(ss > 0.0f) || stall) break;
  }
  memcpy(xio, x, sizeof(float) * d);
  return isfinite(f) ? f : INFINITY;
}

/* Value and gradient as the local search evaluates them (lane-group form). */
void ora_value_grad(const ora_problem* P, const float* x, float* f, float* g) {
  int G = group_size(P->d);
  float xs[ORA_MAXD];
  for (int l = 0; l < G; ++l) xs[l] = l < P->d ? x[l] : 0.0f;
  lane_fg(P, xs, G, f, g);
}

/* All particles from their best positions (AoS swarm); out_x n x d, out_f n. */
void ora_lbfgs_run(const ora_problem* P, const ora_lbfgs_params* lp, int64_t n, const float* parts,
                   float* out_x, float* out_f, int nthreads) {
  int d = P->d;
  int64_t s = stride_of(d);
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads > 0 ? nthreads : 1)
  for (int64_t i = 0; i < n; ++i) {
    float x[ORA_MAXD];
    memcpy(x, BPOS(parts + i * s, d), sizeof(float) * d);
    out_f[i] = ora_lbfgs_one(P, lp, x);
    memcpy(out_x + i * d, x, sizeof(float) * d);
  }
}
