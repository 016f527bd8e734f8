// pps_device.h — device-side arithmetic shared by every libpps kernel.
//
// Everything here is FP32 and written so that each operation happens in a
// fixed order with no contraction (the library is compiled with
// -ffp-contract=off): a particle update is then a pure function of its inputs,
// reproducible bit for bit by any IEEE-754 FP32 implementation that follows
// the same order.  The order follows the reference's Julia expressions
// (src/kernels.jl:1-26), which Julia evaluates left to right without FMA.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "pps_shared.h"


namespace pps {

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
// as 1, 2, 3", SC'11).  Ten rounds of the 4x32 S-box/P-box with the published
// multipliers and Weyl key increments.  The key is per-swarm (uniform across a
// wavefront), so the key schedule lives in scalar registers.
// ---------------------------------------------------------------------------
struct u4 {
  uint32_t x, y, z, w;
};

__device__ __forceinline__ u4 philox4x32_10(u4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)PHILOX_M0 * c.x;
    const uint64_t p1 = (uint64_t)PHILOX_M1 * c.z;
    u4 n;
    // three-input xor in one v_bitop3_b32 (truth table 0x96)
    n.x = __builtin_amdgcn_bitop3_b32((uint32_t)(p1 >> 32), c.y, k0, 0x96);
    n.y = (uint32_t)p1;
    n.z = __builtin_amdgcn_bitop3_b32((uint32_t)(p0 >> 32), c.w, k1, 0x96);
    n.w = (uint32_t)p0;
    c = n;
    k0 += PHILOX_W0;
    k1 += PHILOX_W1;
  }
  return c;
}

// Uniform in [0, 1): the top 24 bits scaled by 2^-24 (exact in FP32).
__device__ __forceinline__ float u01(uint32_t u) {
  return (float)(u >> 8) * 0x1.0p-24f;
}
// The step's attraction terms c * u01(u) as (c * 2^-24) * m, m = the top 24
// bits as a float: m * 2^-24 is exact, and so is c * 2^-24 for c = 0, c not
// finite or |c| >= 2^-102 (the library refuses other nonzero c), so both forms
// round the same real number c m 2^-24 once and are equal bit for bit, with
// one multiplication per draw less.  scaled_coef(c) is computed once per
// thread, draw_m per draw.
__device__ __forceinline__ float draw_m(uint32_t u) { return (float)(u >> 8); }
__device__ __forceinline__ float scaled_coef(float c) { return c * 0x1.0p-24f; }

// ---------------------------------------------------------------------------
// Julia's Base.max / Base.min for IEEE floats: NaN-propagating, and
// max(-0.0, +0.0) = +0.0, min(-0.0, +0.0) = -0.0.  These are IEEE 754-2019
// maximum / minimum, which gfx950 has as single instructions
// (v_maximum3_f32 / v_minimum3_f32); its v_max_f32 / v_min_f32 are maxNum /
// minNum (a NaN operand loses) and would not do.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float jl_max(float x, float y) {
  return __builtin_elementwise_maximum(x, y);
}
__device__ __forceinline__ float jl_min(float x, float y) {
  return __builtin_elementwise_minimum(x, y);
}

// Julia's isless for Float32 (total order: -0.0 < +0.0, NaN above +Inf).
__device__ __forceinline__ bool jl_isless(float a, float b) {
  const bool an = __builtin_isnan(a), bn = __builtin_isnan(b);
  if (an | bn) return !an & bn;
  // Signed-integer image of the bits: negative floats have their magnitude
  // bits flipped, so -0.0 maps to -1 and +0.0 to 0.
  int32_t sa = (int32_t)__float_as_uint(a), sb = (int32_t)__float_as_uint(b);
  if (sa < 0) sa ^= 0x7FFFFFFF;
  if (sb < 0) sb ^= 0x7FFFFFFF;
  return sa < sb;
}

// ---------------------------------------------------------------------------
// Argmin keys.  key = order(cost) << 32 | particle index.  order() is monotone
// in the FP32 value with +0 == -0 (the reference compares costs with `<`), and
// NaN maps above +Inf so a NaN cost never wins.  The smallest key is the
// lowest-index particle among those with the smallest cost.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t cost_order(float c) {
  if (__builtin_isnan(c)) return 0xFFFFFFFFu;
  uint32_t u = __float_as_uint(c);
  if ((u & 0x7FFFFFFFu) == 0u) u = 0u;
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ uint64_t make_key(float cost, uint32_t idx) {
  return ((uint64_t)cost_order(cost) << 32) | (uint64_t)idx;
}
__device__ __forceinline__ float key_cost(uint64_t key) {
  const uint32_t o = (uint32_t)(key >> 32);
  if (o == 0xFFFFFFFFu) return __uint_as_float(0x7FC00000u);
  const uint32_t u = (o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o;
  return __uint_as_float(u);
}

__device__ __forceinline__ uint64_t umin64(uint64_t a, uint64_t b) {
  return a < b ? a : b;
}

// 64-bit DPP move: the value of the lane the control selects, or ~0 (the
// identity of min) in rows outside ROW_MASK.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ uint64_t dpp_u64(uint64_t k) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)(uint32_t)k, CTRL, ROW_MASK,
                                                            0xF, false);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)(uint32_t)(k >> 32), CTRL,
                                                            ROW_MASK, 0xF, false);
  return ((uint64_t)hi << 32) | lo;
}

// Minimum over the 64 lanes of a wavefront, returned to every lane (wave-
// uniform).  All 64 lanes must be active.  DPP moves only, no LDS: quad
// permutations and row rotations leave every lane of a 16-lane row with the
// row's minimum, the row broadcasts of lanes 15 and 31 fold the four rows into
// row 3, and lane 63 is read out.  The minimum does not depend on the order
// of the comparisons.
__device__ __forceinline__ uint64_t wave_min_u64(uint64_t k) {
  k = umin64(k, dpp_u64<0xB1, 0xF>(k));   // quad_perm [1,0,3,2]
  k = umin64(k, dpp_u64<0x4E, 0xF>(k));   // quad_perm [2,3,0,1]
  k = umin64(k, dpp_u64<0x124, 0xF>(k));  // row_ror:4
  k = umin64(k, dpp_u64<0x128, 0xF>(k));  // row_ror:8
  k = umin64(k, dpp_u64<0x142, 0xA>(k));  // row_bcast:15 into rows 1 and 3
  k = umin64(k, dpp_u64<0x143, 0xC>(k));  // row_bcast:31 into rows 2 and 3
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)k, 63);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(k >> 32), 63);
  return ((uint64_t)hi << 32) | lo;
}

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t k) {
  return (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)k, CTRL, ROW_MASK, 0xF, false);
}

// Lowest lane among those with `sel` whose `o` is the smallest over them
// (wave-uniform; -1 when no lane has `sel`).  With o = cost_order(cost) and
// lane = particle index this is the lane of the smallest make_key: the same
// DPP minimum as wave_min_u64 on 32 bits, then the first lane of the ballot of
// the lanes holding it.  All 64 lanes must be active.
__device__ __forceinline__ int wave_argmin_first(uint32_t o, bool sel, uint32_t& m) {
  m = sel ? o : 0xFFFFFFFFu;
  m = min(m, dpp_u32<0xB1, 0xF>(m));
  m = min(m, dpp_u32<0x4E, 0xF>(m));
  m = min(m, dpp_u32<0x124, 0xF>(m));
  m = min(m, dpp_u32<0x128, 0xF>(m));
  m = min(m, dpp_u32<0x142, 0xA>(m));
  m = min(m, dpp_u32<0x143, 0xC>(m));
  m = (uint32_t)__builtin_amdgcn_readlane((int)m, 63);
  const uint64_t eq = __ballot(sel && o == m);
  return eq ? (int)__builtin_ctzll(eq) : -1;
}
__device__ __forceinline__ int wave_argmin_first(uint32_t o, bool sel) {
  uint32_t m;
  return wave_argmin_first(o, sel, m);
}

// Smallest of 16 keys key = order << 32 | index held in slots[0..15] (one per
// wavefront of a workgroup, ~0 for wavefronts not launched), ties to the
// lowest slot: every lane reads slot lane % 16, the DPP minimum of the orders
// over each 16-lane row (all rows hold the same slots, so every lane ends
// with it), then the first lane holding it.  Slot q's keys carry the lowest
// indices of wavefront q, so the lowest slot among equal orders holds the
// lowest index.  Wave-uniform; all 64 lanes must be active.
__device__ __forceinline__ uint64_t slots16_min(const uint64_t* slots) {
  const uint64_t k = slots[threadIdx.x & 15];
  const uint32_t o = (uint32_t)(k >> 32);
  uint32_t m = o;
  m = min(m, dpp_u32<0xB1, 0xF>(m));
  m = min(m, dpp_u32<0x4E, 0xF>(m));
  m = min(m, dpp_u32<0x124, 0xF>(m));
  m = min(m, dpp_u32<0x128, 0xF>(m));
  m = (uint32_t)__builtin_amdgcn_readfirstlane((int)m);
  const uint64_t eq = __ballot(o == m) & 0xFFFFull;
  const int q = (int)__builtin_ctzll(eq);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)k, q);
  return ((uint64_t)m << 32) | lo;
}

// Sum of 16 counts in slots[0..15], the same row-wise exchange: pairs, quads,
// then quads 4 and 8 lanes apart (disjoint sets, each count added once).
// Wave-uniform; all 64 lanes must be active.
__device__ __forceinline__ int slots16_sum(const int* slots) {
  int c = slots[threadIdx.x & 15];
  c += __builtin_amdgcn_update_dpp(0, c, 0xB1, 0xF, 0xF, false);
  c += __builtin_amdgcn_update_dpp(0, c, 0x4E, 0xF, 0xF, false);
  c += __builtin_amdgcn_update_dpp(0, c, 0x124, 0xF, 0xF, false);
  c += __builtin_amdgcn_update_dpp(0, c, 0x128, 0xF, 0xF, false);
  return __builtin_amdgcn_readfirstlane(c);
}

// Lane `l` (wave-uniform) of a float.
__device__ __forceinline__ float read_lane(float x, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}

// ---------------------------------------------------------------------------
// cos(2 pi x), the same function on every device and in every build:
//   r = x - rint(x)          exact, r in [-1/2, 1/2]
//   a = |r|; if a > 1/4: a = 1/2 - a (exact), negate
//   cos(2 pi a) = sum_{k=0}^{7} (-1)^k (2 pi)^{2k} a^{2k} / (2k)!
// The coefficients are the Taylor coefficients of cos(2 pi a) rounded to FP32;
// the truncation error on [0, 1/4] is below (pi/2)^16/16! ~ 7e-11.  Horner in
// s = a^2 with explicit FMAs.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float cos2pi(float x) {
  const float r = x - __builtin_rintf(x);
  float a = __builtin_fabsf(r);
  const bool flip = a > 0.25f;
  a = flip ? (0.5f - a) : a;
  const float s = a * a;
  float p = COS2PI_C7;
  p = __builtin_fmaf(p, s, COS2PI_C6);
  p = __builtin_fmaf(p, s, COS2PI_C5);
  p = __builtin_fmaf(p, s, COS2PI_C4);
  p = __builtin_fmaf(p, s, COS2PI_C3);
  p = __builtin_fmaf(p, s, COS2PI_C2);
  p = __builtin_fmaf(p, s, COS2PI_C1);
  p = __builtin_fmaf(p, s, 1.0f);
  return flip ? -p : p;
}
// sin(2 pi x) = cos(2 pi (x - 1/4)); used by the Rastrigin gradient.
__device__ __forceinline__ float sin2pi(float x) {
  const float r = x - __builtin_rintf(x);  // exact
  return cos2pi(r - 0.25f);
}

// Right-hand sides of the ODE estimation models (two states, four rates), in
// the operation order of the reference's benchmark models:
// Lotka-Volterra (benchmarks/Lotka_Volterra/lotka_volterra.jl:9-13)
//   du1 = p1*u1 - p2*u1*u2 ;  du2 = -p3*u2 + p4*u1*u2
// FitzHugh-Nagumo (benchmarks/Fitzhugh_Nagumo/ode_estimation.jl:6-12), rates
// (a, b, tau_inv, l):
//   dv = v - v^3/3 - w + l ;  dw = tau_inv * (v + a - b*w)
// with v^3 = (v*v)*v (Julia's literal power) and the division by 3 exact IEEE.
template <int MODEL>
__device__ __forceinline__ void ode_rhs(float u1, float u2, float p1, float p2, float p3,
                                        float p4, float& f1, float& f2) {
  if constexpr (MODEL == PPS_ODE_FITZHUGH_NAGUMO) {
    f1 = ((u1 - ((u1 * u1) * u1) / 3.0f) - u2) + p4;
    f2 = p3 * ((u1 + p1) - p2 * u2);
  } else {
    f1 = p1 * u1 - (p2 * u1) * u2;
    f2 = (-p3) * u2 + (p4 * u1) * u2;
  }
}

// Two lanes per particle (the ODE estimation's pair kernel): the even lane of
// a lane pair carries u1, the odd lane u2, and each right-hand side reads both
// components of the pair by DPP.  Both lanes of each pair must be active.
__device__ __forceinline__ float pair_even(float x) {  // quad_perm [0,0,2,2]
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0xA0, 0xF, 0xF, true));
}
__device__ __forceinline__ float pair_odd(float x) {  // quad_perm [1,1,3,3]
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0xF5, 0xF, 0xF, true));
}
// Lotka-Volterra on a pair: both derivatives have one form,
// f = X * own + (Y * u1) * u2:
//   even lane: X = p1,  Y = -p2:  p1 u1 + ((-p2) u1) u2 = p1 u1 - (p2 u1) u2
//   odd lane:  X = -p3, Y = p4:   (-p3) u2 + (p4 u1) u2
// (negating a factor negates the rounded product exactly, and a - b is
// a + (-b)), so each lane's derivative equals ode_rhs's bit for bit.
__device__ __forceinline__ float lv_pair_rhs(float X, float Y, float a) {
  return X * a + (Y * pair_even(a)) * pair_odd(a);
}

// ---------------------------------------------------------------------------
// Tsitouras 5(4) (Ch. Tsitouras, Comput. Math. Appl. 62 (2011) 770-775), the
// method of the reference's default ode_alg = GPUTsit5() (src/ode_pso.jl:54-170):
// nodes c, stage weights a, solution weights b (= row 7 of a, FSAL) and the
// error weights btilde = b - bhat of the embedded 4th-order solution, as FP32.
// ---------------------------------------------------------------------------
#define TS_A21 0.161f
#define TS_A31 -0.008480655492356989f
#define TS_A32 0.335480655492357f
#define TS_A41 2.897153057105493f
#define TS_A42 -6.359448489975075f
#define TS_A43 4.3622954328695815f
#define TS_A51 5.325864828439257f
#define TS_A52 -11.748883564062828f
#define TS_A53 7.4955393428898365f
#define TS_A54 -0.09249506636175525f
#define TS_A61 5.86145544294642f
#define TS_A62 -12.92096931784711f
#define TS_A63 8.159367898576159f
#define TS_A64 -0.071584973281401f
#define TS_A65 -0.028269050394068383f
#define TS_B1 0.09646076681806523f
#define TS_B2 0.01f
#define TS_B3 0.4798896504144996f
#define TS_B4 1.379008574103742f
#define TS_B5 -3.290069515436081f
#define TS_B6 2.324710524099774f
#define TS_E1 -0.00178001105222577714f
#define TS_E2 -0.0008164344596567469f
#define TS_E3 0.007880878010261995f
#define TS_E4 -0.1447110071732629f
#define TS_E5 0.5823571654525552f
#define TS_E6 -0.45808210592918697f
#define TS_E7 0.015151515151515152f

// The components an ODE loss integrates in one lane.  OdeLane: both
// components of a two-component model.  LvPair / FhnPair: a model on two lanes
// per particle (pair_even, pair_odd), the lane's own component.  Component c
// of sample k is obs[2 k + first() + c]; sumsq() is e1^2 + e2^2 over the
// particle's two components, in that order (in both lanes of a pair).
template <int MODEL>
struct OdeLane {
  static constexpr int NC = 2;
  float p1, p2, p3, p4;
  __device__ __forceinline__ explicit OdeLane(const float* th)
      : p1(th[0]), p2(th[1]), p3(th[2]), p4(th[3]) {}
  __device__ __forceinline__ int first() const { return 0; }
  __device__ __forceinline__ void start(const ProbArgs& P, float (&u)[NC]) const {
    u[0] = P.par[0];
    u[1] = P.par[1];
  }
  __device__ __forceinline__ void rhs(const float (&u)[NC], float (&f)[NC]) const {
    ode_rhs<MODEL>(u[0], u[1], p1, p2, p3, p4, f[0], f[1]);
  }
  __device__ __forceinline__ float sumsq(const float (&e)[NC]) const {
    return e[0] * e[0] + e[1] * e[1];
  }
  // u(t + tt h) from Tsit5's continuous extension (tsit5_interp)
  __device__ __forceinline__ void interp(const float (&u)[NC], float h, float tt,
                                         const float (&k)[7][NC], float (&v)[NC]) const;
};
struct LvPair {
  static constexpr int NC = 1;
  bool odd;
  float X, Y;
  __device__ __forceinline__ explicit LvPair(const float* th)
      : odd((threadIdx.x & 1u) != 0), X(odd ? -th[2] : th[0]), Y(odd ? th[3] : -th[1]) {}
  __device__ __forceinline__ int first() const { return odd ? 1 : 0; }
  __device__ __forceinline__ void start(const ProbArgs& P, float (&u)[NC]) const {
    u[0] = odd ? P.par[1] : P.par[0];
  }
  __device__ __forceinline__ void rhs(const float (&u)[NC], float (&f)[NC]) const {
    f[0] = lv_pair_rhs(X, Y, u[0]);
  }
  __device__ __forceinline__ float sumsq(const float (&e)[NC]) const {
    const float q = e[0] * e[0];
    return pair_even(q) + pair_odd(q);
  }
  // As OdeLane::interp, with the weights b2..b7 split between the lanes
  __device__ __forceinline__ void interp(const float (&u)[NC], float h, float tt,
                                         const float (&k)[7][NC], float (&v)[NC]) const;
};
// FitzHugh-Nagumo on a pair: the two derivatives differ in form, so both
// lanes evaluate ode_rhs's two expressions on the pair's (u1, u2) and keep
// their own; stage sums, steps, interpolation and loss split as for LvPair.
struct FhnPair : LvPair {
  float p1, p2, p3, p4;
  __device__ __forceinline__ explicit FhnPair(const float* th)
      : LvPair(th), p1(th[0]), p2(th[1]), p3(th[2]), p4(th[3]) {}
  __device__ __forceinline__ void rhs(const float (&u)[NC], float (&f)[NC]) const {
    float f1, f2;
    ode_rhs<PPS_ODE_FITZHUGH_NAGUMO>(pair_even(u[0]), pair_odd(u[0]), p1, p2, p3, p4, f1, f2);
    f[0] = odd ? f2 : f1;
  }
};

// Sum of squared errors of a fixed-step RK4 trajectory (dt = par[2], sampled
// every steps_per_obs steps, starting at u0 = par[0..1] at sample 0) against
// the observations; the per-sample error is e1^2 + e2^2 and samples are added
// in time order (reference src/ode_pso.jl:47-52).
template <class SYS>
__device__ __forceinline__ float ode_sse_rk4(const SYS& sys, const ProbArgs& P) {
  constexpr int NC = SYS::NC;
  float u[NC];
  sys.start(P, u);
  const float dt = P.par[2];
  const float h2 = dt * 0.5f;
  const float h6 = dt / 6.0f;
  float loss = 0.0f;
  const float* obs = P.obs + sys.first();
  for (int o = 0; o < P.n_obs; ++o) {
    if (o > 0) {
      for (int s = 0; s < P.steps_per_obs; ++s) {
        float k1[NC], k2[NC], k3[NC], k4[NC], y[NC];
        sys.rhs(u, k1);
#pragma unroll
        for (int c = 0; c < NC; ++c) y[c] = u[c] + h2 * k1[c];
        sys.rhs(y, k2);
#pragma unroll
        for (int c = 0; c < NC; ++c) y[c] = u[c] + h2 * k2[c];
        sys.rhs(y, k3);
#pragma unroll
        for (int c = 0; c < NC; ++c) y[c] = u[c] + dt * k3[c];
        sys.rhs(y, k4);
#pragma unroll
        for (int c = 0; c < NC; ++c)
          u[c] = u[c] + h6 * (((k1[c] + 2.0f * k2[c]) + 2.0f * k3[c]) + k4[c]);
      }
    }
    float e[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) e[c] = obs[2 * o + c] - u[c];
    loss = loss + sys.sumsq(e);
  }
  return loss;
}

// One Tsit5 step from u with k1 = f(u) in k[0]: stage sums left to right,
// y_i = u + h * sum_j a_ij k_j, per component; fills k[1..6] with k2..k7
// (k7 = f(new state), the next step's k1) and returns the new state and,
// when ERR, the error estimate h * sum_j btilde_j k_j.
template <class SYS, bool ERR>
__device__ __forceinline__ void tsit5_stages(const SYS& sys, const float (&u)[SYS::NC], float h,
                                             float (&k)[7][SYS::NC], float (&n)[SYS::NC],
                                             float (&e)[SYS::NC]) {
  constexpr int NC = SYS::NC;
  float y[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) y[c] = u[c] + h * (TS_A21 * k[0][c]);
  sys.rhs(y, k[1]);
#pragma unroll
  for (int c = 0; c < NC; ++c) y[c] = u[c] + h * (TS_A31 * k[0][c] + TS_A32 * k[1][c]);
  sys.rhs(y, k[2]);
#pragma unroll
  for (int c = 0; c < NC; ++c)
    y[c] = u[c] + h * ((TS_A41 * k[0][c] + TS_A42 * k[1][c]) + TS_A43 * k[2][c]);
  sys.rhs(y, k[3]);
#pragma unroll
  for (int c = 0; c < NC; ++c)
    y[c] = u[c] + h * (((TS_A51 * k[0][c] + TS_A52 * k[1][c]) + TS_A53 * k[2][c]) +
                       TS_A54 * k[3][c]);
  sys.rhs(y, k[4]);
#pragma unroll
  for (int c = 0; c < NC; ++c)
    y[c] = u[c] + h * ((((TS_A61 * k[0][c] + TS_A62 * k[1][c]) + TS_A63 * k[2][c]) +
                        TS_A64 * k[3][c]) + TS_A65 * k[4][c]);
  sys.rhs(y, k[5]);
#pragma unroll
  for (int c = 0; c < NC; ++c)
    n[c] = u[c] + h * (((((TS_B1 * k[0][c] + TS_B2 * k[1][c]) + TS_B3 * k[2][c]) +
                         TS_B4 * k[3][c]) + TS_B5 * k[4][c]) + TS_B6 * k[5][c]);
  sys.rhs(n, k[6]);
  if (ERR) {
#pragma unroll
    for (int c = 0; c < NC; ++c)
      e[c] = h * ((((((TS_E1 * k[0][c] + TS_E2 * k[1][c]) + TS_E3 * k[2][c]) +
                     TS_E4 * k[3][c]) + TS_E5 * k[4][c]) + TS_E6 * k[5][c]) +
                  TS_E7 * k[6][c]);
  }
}

// Tsit5's continuous extension (Tsitouras 2011, Sec. 4; 4th order):
// u(t + theta h) = u + h * sum_i b_i(theta) k_i with
// b_1(theta) = theta (r_11 + theta (r_12 + theta (r_13 + theta r_14))) and
// b_i(theta) = theta (theta (r_i2 + theta (r_i3 + theta r_i4))) for i >= 2;
// b_i(1) = b_i.  Sum over i left to right.
#define TS_R11 1.0f
#define TS_R12 -2.763706197274826f
#define TS_R13 2.9132554618219126f
#define TS_R14 -1.0530884977290216f
#define TS_R22 0.13169999999999998f
#define TS_R23 -0.2234f
#define TS_R24 0.1017f
#define TS_R32 3.9302962368947516f
#define TS_R33 -5.941033872131505f
#define TS_R34 2.490627285651253f
#define TS_R42 -12.411077166933676f
#define TS_R43 30.33818863028232f
#define TS_R44 -16.548102889244902f
#define TS_R52 37.50931341651104f
#define TS_R53 -88.1789048947664f
#define TS_R54 47.37952196281928f
#define TS_R62 -27.896526289197286f
#define TS_R63 65.09189467479366f
#define TS_R64 -34.87065786149661f
#define TS_R72 1.5f
#define TS_R73 -4.0f
#define TS_R74 2.5f

template <int NC>
__device__ __forceinline__ void tsit5_interp(const float (&u)[NC], float h, float tt,
                                             const float (&k)[7][NC], float (&v)[NC]) {
  const float b1 = tt * (TS_R11 + tt * (TS_R12 + tt * (TS_R13 + tt * TS_R14)));
  const float b2 = tt * (tt * (TS_R22 + tt * (TS_R23 + tt * TS_R24)));
  const float b3 = tt * (tt * (TS_R32 + tt * (TS_R33 + tt * TS_R34)));
  const float b4 = tt * (tt * (TS_R42 + tt * (TS_R43 + tt * TS_R44)));
  const float b5 = tt * (tt * (TS_R52 + tt * (TS_R53 + tt * TS_R54)));
  const float b6 = tt * (tt * (TS_R62 + tt * (TS_R63 + tt * TS_R64)));
  const float b7 = tt * (tt * (TS_R72 + tt * (TS_R73 + tt * TS_R74)));
#pragma unroll
  for (int c = 0; c < NC; ++c)
    v[c] = u[c] + h * ((((((b1 * k[0][c] + b2 * k[1][c]) + b3 * k[2][c]) + b4 * k[3][c]) +
                         b5 * k[4][c]) + b6 * k[5][c]) + b7 * k[6][c]);
}

template <int MODEL>
__device__ __forceinline__ void OdeLane<MODEL>::interp(const float (&u)[NC], float h, float tt,
                                                       const float (&k)[7][NC],
                                                       float (&v)[NC]) const {
  tsit5_interp<NC>(u, h, tt, k, v);
}

// The even lane evaluates b2, b3, b4 and the odd lane b5, b6, b7 (the same
// polynomials, in the same form), and each product b_i k_i reads b_i from the
// lane that has it by DPP; the sum is tsit5_interp's, bit for bit.
__device__ __forceinline__ void LvPair::interp(const float (&u)[NC], float h, float tt,
                                               const float (&k)[7][NC], float (&v)[NC]) const {
  const float r02 = odd ? TS_R52 : TS_R22, r03 = odd ? TS_R53 : TS_R23, r04 = odd ? TS_R54 : TS_R24;
  const float r12 = odd ? TS_R62 : TS_R32, r13 = odd ? TS_R63 : TS_R33, r14 = odd ? TS_R64 : TS_R34;
  const float r22 = odd ? TS_R72 : TS_R42, r23 = odd ? TS_R73 : TS_R43, r24 = odd ? TS_R74 : TS_R44;
  const float b1 = tt * (TS_R11 + tt * (TS_R12 + tt * (TS_R13 + tt * TS_R14)));
  const float s0 = tt * (tt * (r02 + tt * (r03 + tt * r04)));  // b2 | b5
  const float s1 = tt * (tt * (r12 + tt * (r13 + tt * r14)));  // b3 | b6
  const float s2 = tt * (tt * (r22 + tt * (r23 + tt * r24)));  // b4 | b7
  v[0] = u[0] + h * ((((((b1 * k[0][0] + pair_even(s0) * k[1][0]) + pair_even(s1) * k[2][0]) +
                        pair_even(s2) * k[3][0]) + pair_odd(s0) * k[4][0]) +
                      pair_odd(s1) * k[5][0]) + pair_odd(s2) * k[6][0]);
}

// x^(1/5) for x > 0 from basic IEEE operations only (an exponent-halving seed
// and four Newton steps y <- (4y + x/y^4) / 5), so that every platform
// computes the same float; 0 for x <= 0.
__device__ __forceinline__ float root5(float x) {
  if (!(x > 0.0f)) return 0.0f;
  if (x == __builtin_inff()) return x;
  const int32_t bits = (int32_t)__float_as_uint(x);
  float y = __uint_as_float((uint32_t)((bits - 0x3f800000) / 5 + 0x3f800000));
#pragma unroll
  for (int it = 0; it < 4; ++it) {
    const float y2 = y * y;
    y = (4.0f * y + x / (y2 * y2)) / 5.0f;
  }
  return y;
}

// Step limit of one adaptive evaluation when par[6] is not given.
#define ODE_MAX_STEPS 10000

// ODE-model SSE with Tsit5: fixed steps (dt = par[2], par[3] steps per
// observation, the reference's parameter_estim_ode! with Val(false)) or
// adaptive steps (par[2] = observation interval, par[3] = initial dt,
// par[4] = abstol, par[5] = reltol; Val(true)).  Adaptive control: scaled
// error s_i = e_i / (abstol + max(|u_i|, |u_new_i|) reltol), EEst =
// sqrt((s_1^2 + s_2^2) / 2); accept when EEst <= 1; factor q = EEst^(1/5) /
// 0.9 clamped to [0.1, 5] (5 when EEst is not finite), next dt = h / q; steps
// end exactly on the observation times, and a step shortened to meet one does
// not shrink the following proposal.  More steps than par[6] (10000 when
// par[6] <= 0), or a step too small to advance t, make the loss +Inf: a
// parameter vector whose trajectory turns stiff (the whole-vector clamp of
// the ODE path admits negative rates) would otherwise hold its wavefront for
// the explicit method's stability-limited steps.  Loss as ode_sse_rk4.
//
// `bound` (adaptive steps only): the loss is a running sum of non-negative
// terms, so once it reaches `bound` (the particle's best cost) the full sum
// cannot fall below it and the caller's `cost < best` test is decided; the
// partial sum is returned.  This is what keeps a particle whose trajectory
// turns stiff from running to the step limit.  A NaN bound never stops the
// sum.  (With fixed steps every lane of a wavefront integrates equally long,
// and stopping early was measured to gain nothing.)
template <class SYS, bool ADAPTIVE>
__device__ __forceinline__ float ode_sse_tsit5(const SYS& sys, const ProbArgs& P, float bound) {
  constexpr int NC = SYS::NC;
  float u[NC], k[7][NC];
  sys.start(P, u);
  sys.rhs(u, k[0]);
  float loss = 0.0f;
  const float* obs = P.obs + sys.first();
  if (P.n_obs <= 0) return loss;
  {
    float e[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) e[c] = obs[c] - u[c];
    loss = loss + sys.sumsq(e);
  }
  if (!ADAPTIVE) {
    const float dt = P.par[2];
    for (int o = 1; o < P.n_obs; ++o) {
      for (int s = 0; s < P.steps_per_obs; ++s) {
        float n[NC], e[NC];
        tsit5_stages<SYS, false>(sys, u, dt, k, n, e);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          u[c] = n[c];
          k[0][c] = k[6][c];
        }
      }
      float e[NC];
#pragma unroll
      for (int c = 0; c < NC; ++c) e[c] = obs[2 * o + c] - u[c];
      loss = loss + sys.sumsq(e);
    }
    return loss;
  }
  // Adaptive: one flat loop over attempted steps in which each lane moves on
  // to its next observation by itself.  (Nested loops, a step loop inside a
  // loop over observations, would make a wavefront run, interval by
  // interval, as many steps as its slowest lane needs in that interval.)
  const float atol = P.par[4], rtol = P.par[5];
  const int32_t max_steps = P.par[6] > 0.0f ? (int32_t)P.par[6] : ODE_MAX_STEPS;
  float t = 0.0f, dt = P.par[3];
  int32_t steps = 0;
  int o = 1;
  while (o < P.n_obs) {
    const float t_next = (float)o * P.par[2];
    if (!(t < t_next)) {
      float e[NC];
#pragma unroll
      for (int c = 0; c < NC; ++c) e[c] = obs[2 * o + c] - u[c];
      loss = loss + sys.sumsq(e);
      if (loss >= bound) return loss;
      ++o;
      continue;
    }
    const bool hit = !(t + dt < t_next);
    const float h = hit ? t_next - t : dt;
    if (++steps > max_steps || !(t + h > t)) return __builtin_inff();
    float kk[7][NC], n[NC], e[NC], sc[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) kk[0][c] = k[0][c];
    tsit5_stages<SYS, true>(sys, u, h, kk, n, e);
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const float a = __builtin_fabsf(u[c]), b = __builtin_fabsf(n[c]);
      sc[c] = e[c] / (atol + (a > b ? a : b) * rtol);
    }
    const float eest = sqrtf(sys.sumsq(sc) * 0.5f);
    float q = 5.0f;
    if (eest <= 3.0e38f) {
      q = root5(eest) / 0.9f;
      q = q < 0.1f ? 0.1f : (q > 5.0f ? 5.0f : q);
    }
    const float hnew = h / q;
    if (eest <= 1.0f) {
      t = hit ? t_next : t + h;
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        u[c] = n[c];
        k[0][c] = kk[6][c];
      }
      dt = hit ? (hnew > dt ? hnew : dt) : hnew;
    } else {
      dt = hnew;
    }
  }
  return loss;
}

// ODE-model SSE with Tsit5 and saveat semantics: observation k at
// t_k = k * par[2] (FP32 product), steps of par[3] (the initial step when
// ADAPTIVE) from t = 0, the last step shortened to end on t_end = t_{n_obs-1};
// after each accepted step [t, t + h] every unsaved t_k <= t + h is read from
// the continuous extension at theta = (t_k - t) / h.  Adaptive control, step
// limit par[6] and the +Inf results as ode_sse_tsit5; fixed steps carry no
// limit (the host bounds their number).  Loss as ode_sse_rk4.
template <class SYS, bool ADAPTIVE>
__device__ __forceinline__ float ode_sse_tsit5_saveat(const SYS& sys, const ProbArgs& P,
                                                      float bound) {
  constexpr int NC = SYS::NC;
  float u[NC], k[7][NC];
  sys.start(P, u);
  sys.rhs(u, k[0]);
  const float* obs = P.obs + sys.first();
  const float tau = P.par[2];
  const float t_end = (float)(P.n_obs - 1) * tau;
  float loss;
  {
    float e[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) e[c] = obs[c] - u[c];
    loss = sys.sumsq(e);
  }
  int o = 1;
  float t = 0.0f, dt = P.par[3];
  int32_t steps = 0;
  const int32_t max_steps =
      ADAPTIVE ? (P.par[6] > 0.0f ? (int32_t)P.par[6] : ODE_MAX_STEPS) : 0x7fffffff;
  while (o < P.n_obs) {
    const bool last = !(t + dt < t_end);
    const float h = last ? t_end - t : dt;
    if (++steps > max_steps || !(t + h > t)) return __builtin_inff();
    float n[NC], e[NC];
    tsit5_stages<SYS, ADAPTIVE>(sys, u, h, k, n, e);
    if (ADAPTIVE) {
      float sc[NC];
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const float a = __builtin_fabsf(u[c]), b = __builtin_fabsf(n[c]);
        sc[c] = e[c] / (P.par[4] + (a > b ? a : b) * P.par[5]);
      }
      const float eest = sqrtf(sys.sumsq(sc) * 0.5f);
      float q = 5.0f;
      if (eest <= 3.0e38f) {
        q = root5(eest) / 0.9f;
        q = q < 0.1f ? 0.1f : (q > 5.0f ? 5.0f : q);
      }
      dt = h / q;
      if (!(eest <= 1.0f)) continue;
    }
    const float t_new = last ? t_end : t + h;
    while (o < P.n_obs) {
      const float tk = (float)o * tau;
      if (tk > t_new) break;
      float v[NC], dv[NC];
      sys.interp(u, h, (tk - t) / h, k, v);
#pragma unroll
      for (int c = 0; c < NC; ++c) dv[c] = obs[2 * o + c] - v[c];
      loss = loss + sys.sumsq(dv);
      if (ADAPTIVE && loss >= bound) return loss;
      ++o;
    }
    t = t_new;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      u[c] = n[c];
      k[0][c] = k[6][c];
    }
  }
  return loss;
}

// ODE loss of objective OBJ on two lanes per particle (LvPair or FhnPair):
// the value objective<4, OBJ> returns for z, bit for bit.  Both lanes of each
// pair must be active and hold the same z.
template <int OBJ>
__device__ __forceinline__ float pair_objective(const float (&z)[4], const ProbArgs& P,
                                                float bound) {
  static_assert(is_ode_objective(OBJ), "ODE objectives only");
  using SYS = typename std::conditional<ode_model(OBJ) == PPS_ODE_LOTKA_VOLTERRA, LvPair,
                                        FhnPair>::type;
  const SYS sys(z);
  constexpr int S = ode_solver(OBJ);
  if constexpr (S == PPS_ODE_RK4) return ode_sse_rk4<SYS>(sys, P);
  else if constexpr (S == PPS_ODE_TSIT5) return ode_sse_tsit5<SYS, false>(sys, P, bound);
  else if constexpr (S == PPS_ODE_ATSIT5) return ode_sse_tsit5<SYS, true>(sys, P, bound);
  else if constexpr (S == PPS_ODE_TSIT5_SAVEAT) return ode_sse_tsit5_saveat<SYS, false>(sys, P, bound);
  else return ode_sse_tsit5_saveat<SYS, true>(sys, P, bound);
}

// Objective f(x - shift) over a register array (compile-time capacity MAXD,
// runtime dimension d <= MAXD).  OBJ is a compile-time PPS_OBJ_* constant.
// `bound` is passed to the adaptive ODE losses (see ode_sse_tsit5); other
// objectives ignore it.
template <int MAXD, int OBJ>
__device__ __forceinline__ float objective(const float (&xin)[MAXD], const int d,
                                           const ProbArgs& P, float bound) {
  float z[MAXD];
#pragma unroll
  // an absent shift is all +0.0f, and x - (+0.0f) is x bit for bit
  for (int j = 0; j < MAXD; ++j) z[j] = xin[j] - P.shift[j];
  if constexpr (OBJ == PPS_OBJ_ROSENBROCK) {
    // res += p[2] * (x[i+1] - x[i]^2)^2 + (p[1] - x[i])^2   (test/gpu.jl:14-20)
    const float p1 = P.par[0], p2 = P.par[1];
    float res = 0.0f;
#pragma unroll
    for (int i = 0; i + 1 < MAXD; ++i) {
      if (i + 1 < d) {
        const float dd = z[i + 1] - z[i] * z[i];
        const float e = p1 - z[i];
        res = res + (p2 * (dd * dd) + e * e);
      }
    }
    return res;
  } else if constexpr (OBJ == PPS_OBJ_RASTRIGIN) {
    float res = 10.0f * (float)d;
#pragma unroll
    for (int i = 0; i < MAXD; ++i) {
      if (i < d) res = res + (z[i] * z[i] - 10.0f * cos2pi(z[i]));
    }
    return res;
  } else if constexpr (OBJ == PPS_OBJ_SPHERE) {
    float res = 0.0f;
#pragma unroll
    for (int i = 0; i < MAXD; ++i) {
      if (i < d) res = res + z[i] * z[i];
    }
    return res;
  } else if constexpr (OBJ == PPS_OBJ_BEALE) {
    const float x1 = z[0], x2 = z[MAXD > 1 ? 1 : 0];
    const float a = (1.5f - x1) + x1 * x2;
    const float b = (2.25f - x1) + x1 * (x2 * x2);
    const float c = (2.625f - x1) + x1 * ((x2 * x2) * x2);
    return (a * a + b * b) + c * c;
  } else if constexpr (OBJ == PPS_OBJ_BOOTH) {
    const float x1 = z[0], x2 = z[MAXD > 1 ? 1 : 0];
    const float a = (x1 + 2.0f * x2) - 7.0f;
    const float b = (2.0f * x1 + x2) - 5.0f;
    return a * a + b * b;
  } else if constexpr (OBJ == PPS_OBJ_HIMMELBLAU) {
    const float x1 = z[0], x2 = z[MAXD > 1 ? 1 : 0];
    const float a = (x1 * x1 + x2) - 11.0f;
    const float b = (x1 + x2 * x2) - 7.0f;
    return a * a + b * b;
  } else if constexpr (OBJ == PPS_OBJ_QUADRATIC) {
    const float x1 = z[0], x2 = z[MAXD > 1 ? 1 : 0];
    const float a = x1 - 1.0f, b = x2 + 2.0f;
    return P.par[0] * (a * a) + b * b;
  } else if constexpr (OBJ == PPS_OBJ_BOUNDARY_QUADRATIC) {
    const float a = z[0] + 1.0f;
    return a * a;
  } else if constexpr (OBJ == PPS_OBJ_ROSEN2) {
    const float x1 = z[0], x2 = z[MAXD > 1 ? 1 : 0];
    const float a = P.par[0] - x1;
    const float b = x2 - x1 * x1;
    return a * a + P.par[1] * (b * b);
  } else if constexpr (is_ode_objective(OBJ)) {
    constexpr int M = ode_model(OBJ), S = ode_solver(OBJ);
    if constexpr (S == PPS_ODE_RK4) return ode_sse_rk4<OdeLane<M>>(OdeLane<M>(z), P);
    else if constexpr (S == PPS_ODE_TSIT5) return ode_sse_tsit5<OdeLane<M>, false>(OdeLane<M>(z), P, bound);
    else if constexpr (S == PPS_ODE_ATSIT5) return ode_sse_tsit5<OdeLane<M>, true>(OdeLane<M>(z), P, bound);
    else if constexpr (S == PPS_ODE_TSIT5_SAVEAT)
      return ode_sse_tsit5_saveat<OdeLane<M>, false>(OdeLane<M>(z), P, bound);
    else return ode_sse_tsit5_saveat<OdeLane<M>, true>(OdeLane<M>(z), P, bound);
  } else {
    return __uint_as_float(0x7FC00000u);
  }
}

// ---------------------------------------------------------------------------
// The per-particle PSO step (reference update_particle_state,
// src/kernels.jl:1-26), component j, with r1/r2 drawn from the counter-based
// stream:
//   v' = (w*v + (c1*r1)*(pbest - x)) + (c2*r2)*(gbest - x)
//   x' = clamp(x + v')
// ---------------------------------------------------------------------------
// Constraint penalty of the reference (src/utils.jl:320-372), called from
// handle_constraints (src/kernels.jl:28-36):
//   q_k = max(g_k(x), 0)                      (Julia max: NaN-propagating)
//   theta(q) = 10 if q < 0.001, 20 if q <= 0.1, 100 if q <= 1, else 300
//   gamma(q) = 1 if q < 1, else 2             (the thresholds are Float64
//                                              literals compared after promotion)
//   penalty = h * ((sum_k theta(q_k)) + theta(q_0) q_0^gamma(q_0) + ...)
// with q^1 = q and q^2 = q*q (the correctly rounded power).  g_k sums its
// nonzero linear terms in index order, then its nonzero quadratic terms, then
// adds c_k: for the reference's test constraints (test/constraints.jl:11-14)
// that is their own FP32 evaluation order.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float penalty_theta(float q) {
  const double qd = (double)q;
  return qd < 0.001 ? 10.0f : (qd <= 0.1 ? 20.0f : (qd <= 1.0 ? 100.0f : 300.0f));
}

template <int MAXD>
__device__ float constraint_penalty(const float (&x)[MAXD], const int d, const ProbArgs& P,
                                    float h) {
  float qs[PPS_MAXCONS];
  float sum_theta = 0.0f;
  for (int k = 0; k < P.n_cons; ++k) {
    const float* row = P.cons + (int64_t)k * (2 * d + 1);
    float g = 0.0f;
#pragma unroll
    for (int j = 0; j < MAXD; ++j)
      if (j < d && row[j] != 0.0f) g = g + row[j] * x[j];
#pragma unroll
    for (int j = 0; j < MAXD; ++j)
      if (j < d && row[d + j] != 0.0f) g = g + row[d + j] * (x[j] * x[j]);
    g = g + row[2 * d];
    const float q = jl_max(g, 0.0f);
    qs[k] = q;
    sum_theta = k == 0 ? penalty_theta(q) : sum_theta + penalty_theta(q);
  }
  float pen = 0.0f + sum_theta;
  for (int k = 0; k < P.n_cons; ++k) {
    const float q = qs[k];
    const float pw = q < 1.0f ? q : q * q;
    pen = pen + penalty_theta(q) * pw;
  }
  return h * pen;
}

// Random stream: for particle i at iteration t (1-based) of epoch e, call
// philox(ctr = (i, t, k | TAG_STEP << 16, e), key = seed) for k = 0..ceil(d/2)-1
// and use lanes (x, y, z, w) as (r1[2k], r2[2k], r1[2k+1], r2[2k+1]).
// ---------------------------------------------------------------------------
template <int MAXD, int OBJ>
__device__ __forceinline__ void clamp_position(float (&x)[MAXD], const int d, const ProbArgs& P);

template <int MAXD, int OBJ>
__device__ __forceinline__ void velocity_position(
    float (&x)[MAXD], float (&v)[MAXD], const float (&pb)[MAXD],
    const float* gb, float w, float c1, float c2, uint32_t i, uint32_t t,
    uint32_t epoch, uint32_t k0, uint32_t k1, const int d, const ProbArgs& P) {
  const float c1s = scaled_coef(c1), c2s = scaled_coef(c2);
#pragma unroll
  for (int k = 0; 2 * k < MAXD; ++k) {
    if (2 * k < d) {
      const u4 r = philox4x32_10(u4{i, t, (uint32_t)k | (PPS_TAG_STEP << 16), epoch}, k0, k1);
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int j = 2 * k + h;
        if (j < MAXD && j < d) {
          const float m1 = draw_m(h == 0 ? r.x : r.z);
          const float m2 = draw_m(h == 0 ? r.y : r.w);
          const float a = w * v[j];
          const float b = (c1s * m1) * (pb[j] - x[j]);  // (c1 * r1) * (pb - x)
          const float c = (c2s * m2) * (gb[j] - x[j]);
          v[j] = (a + b) + c;
          x[j] = x[j] + v[j];
        }
      }
    }
  }
  clamp_position<MAXD, OBJ>(x, d, P);
}

// The same step split in two: the draws of particle i at iteration t (they
// depend on neither the state nor the global best, so a kernel can compute
// them while its loads are in flight; kept as draw_m values), then the update
// from those draws.  Together they compute exactly what velocity_position
// computes.
template <int MAXD>
__device__ __forceinline__ void step_draws(float (&r1)[MAXD], float (&r2)[MAXD], uint32_t i,
                                           uint32_t t, uint32_t epoch, uint32_t k0, uint32_t k1,
                                           const int d) {
#pragma unroll
  for (int k = 0; 2 * k < MAXD; ++k) {
    if (2 * k < d) {
      const u4 r = philox4x32_10(u4{i, t, (uint32_t)k | (PPS_TAG_STEP << 16), epoch}, k0, k1);
      r1[2 * k] = draw_m(r.x);
      r2[2 * k] = draw_m(r.y);
      if (2 * k + 1 < MAXD) {
        r1[2 * k + 1] = draw_m(r.z);
        r2[2 * k + 1] = draw_m(r.w);
      }
    }
  }
}
template <int MAXD, int OBJ>
__device__ __forceinline__ void velocity_position_drawn(
    float (&x)[MAXD], float (&v)[MAXD], const float (&pb)[MAXD], const float (&gb)[MAXD],
    float w, float c1, float c2, const float (&r1)[MAXD], const float (&r2)[MAXD], const int d,
    const ProbArgs& P) {
  const float c1s = scaled_coef(c1), c2s = scaled_coef(c2);
#pragma unroll
  for (int j = 0; j < MAXD; ++j) {
    if (j < d) {
      const float a = w * v[j];
      const float b = (c1s * r1[j]) * (pb[j] - x[j]);
      const float c = (c2s * r2[j]) * (gb[j] - x[j]);
      v[j] = (a + b) + c;
      x[j] = x[j] + v[j];
    }
  }
  clamp_position<MAXD, OBJ>(x, d, P);
}

// Bound handling after the position update.
template <int MAXD, int OBJ>
__device__ __forceinline__ void clamp_position(float (&x)[MAXD], const int d, const ProbArgs& P) {
  // The lexicographic clamp is the ODE path's (src/ode_pso.jl:18-19): it is
  // compiled only into the Lotka-Volterra kernels.
  if (!is_ode_objective(OBJ) || P.clamp == PPS_CLAMP_COMPONENT) {
#pragma unroll
    for (int j = 0; j < MAXD; ++j) {
      if (j < d) x[j] = jl_min(jl_max(x[j], P.lb[j]), P.ub[j]);
    }
  } else {
    // Whole-vector max/min under Julia's lexicographic isless
    // (src/ode_pso.jl:18-19): max(x, lb) = isless(lb, x) ? x : lb.
    int c = 0;  // cmp(lb, x): -1, 0, 1
#pragma unroll
    for (int j = 0; j < MAXD; ++j) {
      if (j < d && c == 0) {
        const float a = P.lb[j], b = x[j];
        const bool eq = (__float_as_uint(a) == __float_as_uint(b)) ||
                        (__builtin_isnan(a) & __builtin_isnan(b));
        if (!eq) c = jl_isless(a, b) ? -1 : 1;
      }
    }
    const bool take_lb = !(c < 0);
#pragma unroll
    for (int j = 0; j < MAXD; ++j) if (j < d && take_lb) x[j] = P.lb[j];
    c = 0;  // cmp(ub, x)
#pragma unroll
    for (int j = 0; j < MAXD; ++j) {
      if (j < d && c == 0) {
        const float a = P.ub[j], b = x[j];
        const bool eq = (__float_as_uint(a) == __float_as_uint(b)) ||
                        (__builtin_isnan(a) & __builtin_isnan(b));
        if (!eq) c = jl_isless(a, b) ? -1 : 1;
      }
    }
    const bool take_ub = (c < 0);
#pragma unroll
    for (int j = 0; j < MAXD; ++j) if (j < d && take_ub) x[j] = P.ub[j];
  }
}

// Uniform initial position (reference src/utils.jl:1-5,57-84):
// x_j = r * (ub_j - lb_j) + lb_j with r from philox(ctr = (i, 0,
// m | TAG_INIT << 16, e)) lanes (x, y, z, w) for dimensions 4m..4m+3.
template <int MAXD>
__device__ __forceinline__ void init_position(float (&x)[MAXD], uint32_t i,
                                              uint32_t epoch, uint32_t k0,
                                              uint32_t k1, const int d,
                                              const ProbArgs& P) {
#pragma unroll
  for (int m = 0; 4 * m < MAXD; ++m) {
    if (4 * m < d) {
      const u4 r = philox4x32_10(u4{i, 0u, (uint32_t)m | (PPS_TAG_INIT << 16), epoch}, k0, k1);
      const uint32_t rr[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
      for (int h = 0; h < 4; ++h) {
        const int j = 4 * m + h;
        if (j < MAXD && j < d) x[j] = u01(rr[h]) * (P.ub[j] - P.lb[j]) + P.lb[j];
      }
    }
  }
}

// GPUUnboundedSampler (src/utils.jl:23-55): x_j = u0_j + r * |u0_j| when
// |u0_j| > 0, else r, with r drawn as in init_position.
template <int MAXD>
__device__ __forceinline__ void unbounded_position(float (&x)[MAXD], uint32_t i, uint32_t epoch,
                                                   uint32_t k0, uint32_t k1, const int d,
                                                   const float* u0) {
#pragma unroll
  for (int m = 0; 4 * m < MAXD; ++m) {
    if (4 * m < d) {
      const u4 r = philox4x32_10(u4{i, 0u, (uint32_t)m | (PPS_TAG_INIT << 16), epoch}, k0, k1);
      const uint32_t rr[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
      for (int h = 0; h < 4; ++h) {
        const int j = 4 * m + h;
        if (j < MAXD && j < d) {
          const float c = u0[j], a = __builtin_fabsf(c), u = u01(rr[h]);
          x[j] = a > 0.0f ? c + u * a : u;
        }
      }
    }
  }
}

// Keyed permutation of [0, n) for the Latin hypercube sampler: a balanced
// 4-round Feistel network on b bits (b even, 2^b >= n, 2^b < 4n for n > 4)
// whose round function is Philox(ctr = (right half, round, j | TAG << 16,
// epoch)) lane x masked to b/2 bits, iterated until the value lies below n
// ("cycle walking"; the network is a bijection on [0, 2^b), so the orbit of
// i < n returns below n, on average in fewer than 4 passes).
__device__ __forceinline__ uint32_t lhs_permute(uint32_t i, uint32_t n, uint32_t j,
                                                uint32_t epoch, uint32_t k0, uint32_t k1) {
  uint32_t b = 2;
  while (b < 32 && (1ull << b) < (uint64_t)n) b += 2;
  const uint32_t h = b / 2, mask = (uint32_t)((1ull << h) - 1ull);
  uint32_t x = i;
  do {
    uint32_t L = x >> h, R = x & mask;
#pragma unroll
    for (uint32_t r = 0; r < 4; ++r) {
      const uint32_t f =
          philox4x32_10(u4{R, r, j | (PPS_TAG_LHS_PERM << 16), epoch}, k0, k1).x & mask;
      const uint32_t t = L ^ f;
      L = R;
      R = t;
    }
    x = (L << h) | R;
  } while (x >= n);
  return x;
}

// Latin hypercube position: t = (pi_j(i) + u) / n with u from
// philox(ctr = (i, 0, m | TAG_LHS_JITTER << 16, e)) lanes (x, y, z, w) for
// coordinates 4m..4m+3, then x_j = t * (ub_j - lb_j) + lb_j (FP32).
template <int MAXD>
__device__ __forceinline__ void lhs_position(float (&x)[MAXD], uint32_t i, uint32_t n,
                                             uint32_t epoch, uint32_t k0, uint32_t k1,
                                             const int d, const ProbArgs& P) {
#pragma unroll
  for (int m = 0; 4 * m < MAXD; ++m) {
    if (4 * m < d) {
      const u4 r =
          philox4x32_10(u4{i, 0u, (uint32_t)m | (PPS_TAG_LHS_JITTER << 16), epoch}, k0, k1);
      const uint32_t rr[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
      for (int h = 0; h < 4; ++h) {
        const int j = 4 * m + h;
        if (j < MAXD && j < d) {
          const uint32_t pi = lhs_permute(i, n, (uint32_t)j, epoch, k0, k1);
          const float t = ((float)pi + u01(rr[h])) / (float)n;
          x[j] = t * (P.ub[j] - P.lb[j]) + P.lb[j];
        }
      }
    }
  }
}

}  // namespace pps
