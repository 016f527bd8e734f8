// pps_lbfgs.hip — HybridPSO local phase: projected L-BFGS from every
// particle's best position (reference lbfgs_run!, src/lbfgs.jl:9-21, driven by
// src/hybrid.jl:100-158).
//
// Layout: one particle per group of G = 2^ceil(log2 d) lanes, coordinate l in
// lane l (lanes >= d carry zeros with lb = ub = 0).  A wavefront holds 64/G
// particles.  The history pairs (s_i, y_i), i < m <= 16, live in registers as
// one coordinate per lane: 2m VGPRs per lane instead of 2md per thread.  Dot
// products and objective sums are xor butterflies inside the group
// (v = v + v[l ^ s] for s = G/2 .. 1, the partner's value brought by DPP or a
// swizzle, see partner()), so every lane of a group holds the identical value
// and all branch decisions are group-uniform.
//
// Algorithm (Nocedal & Wright Alg. 7.4 two-loop recursion, projected onto the
// box; the reference's SimpleOptimization._lbfgs and LineSearch.jl are not
// vendored under the reference):
//   x = P(x0); f, g = f(x), grad f(x)
//   repeat maxiters: stop if max|P(x - g) - x| <= abstol
//     p = -H g (two-loop; gamma = s'y / y'y of the newest pair); if g'p >= 0:
//       p = -g and forget the history
//     line search on phi(a) = f(P(x + a p)) (wolfe_search or armijo_search
//       below); stop if it fails
//       (a failed search with history is retried once along -g)
//     s = xn - x, y = grad f(xn) - g; keep the pair if s'y > 1e-10 y'y, s'y > 0,
//       otherwise restart the history
//     stop if the projected step is zero, or (reltol > 0) if
//       |f - f(xn)| <= reltol |f|
//   non-finite final value -> +Inf (src/lbfgs.jl:20)
#include <hip/hip_runtime.h>

#include "pps_device.h"
#include "pps_kernels.h"

namespace pps {

constexpr int LB_MAXH = 16;

// Lane exchanges of the group reductions without an LDS round trip: DPP moves
// inside a row of 16 lanes (quad permutations, row rotations) and, for the
// offsets 4 and 16, a bit-mask swizzle inside 32 lanes.  A group never spans
// a row boundary unless G >= 32, and a row rotation by 8 is exactly the xor-8
// partner; a rotation by 4 equals the xor-4 partner once the values repeat
// with period 8, which the xor-8 step leaves behind.
constexpr int DPP_QUAD_XOR1 = 0xB1;  // quad_perm [1,0,3,2]
constexpr int DPP_QUAD_XOR2 = 0x4E;  // quad_perm [2,3,0,1]
constexpr int DPP_ROW_ROR1 = 0x121;  // row_ror:1
constexpr int DPP_ROW_ROR4 = 0x124;  // row_ror:4
constexpr int DPP_ROW_ROR8 = 0x128;  // row_ror:8
constexpr int DPP_ROW_ROR15 = 0x12F; // row_ror:15
constexpr int SWZ_XOR4 = (4 << 10) | 0x1F;    // ds_swizzle bit mode: lane ^ 4
constexpr int SWZ_XOR16 = (16 << 10) | 0x1F;  // lane ^ 16

template <int CTRL>
__device__ __forceinline__ float dpp(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int PAT>
__device__ __forceinline__ float swz(float v) {
  return __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), PAT));
}

// Value of lane (l ^ off) of the group at butterfly step `off` (steps taken in
// the order G/2, ..., 1, so the periodicity above holds at every step).
template <int G, int OFF>
__device__ __forceinline__ float partner(float v) {
  if constexpr (OFF == 32) return __shfl_xor(v, 32, 64);
  else if constexpr (OFF == 16) return swz<SWZ_XOR16>(v);
  else if constexpr (OFF == 8) return dpp<DPP_ROW_ROR8>(v);
  else if constexpr (OFF == 4) return G == 8 ? swz<SWZ_XOR4>(v) : dpp<DPP_ROW_ROR4>(v);
  else if constexpr (OFF == 2) return dpp<DPP_QUAD_XOR2>(v);
  else return dpp<DPP_QUAD_XOR1>(v);
}

// Butterfly sum: step off adds the partner's value, v = v + v[l ^ off], for
// off = G/2, ..., 1, so every lane of the group ends with the same float.
template <int G>
__device__ __forceinline__ float gsum(float v) {
  if constexpr (G >= 64) v = v + partner<G, 32>(v);
  if constexpr (G >= 32) v = v + partner<G, 16>(v);
  if constexpr (G >= 16) v = v + partner<G, 8>(v);
  if constexpr (G >= 8) v = v + partner<G, 4>(v);
  if constexpr (G >= 4) v = v + partner<G, 2>(v);
  if constexpr (G >= 2) v = v + partner<G, 1>(v);
  return v;
}
template <int G>
__device__ __forceinline__ float gmaxabs(float v) {
  v = __builtin_fabsf(v);
  if constexpr (G >= 64) v = __builtin_fmaxf(v, partner<G, 32>(v));
  if constexpr (G >= 32) v = __builtin_fmaxf(v, partner<G, 16>(v));
  if constexpr (G >= 16) v = __builtin_fmaxf(v, partner<G, 8>(v));
  if constexpr (G >= 8) v = __builtin_fmaxf(v, partner<G, 4>(v));
  if constexpr (G >= 4) v = __builtin_fmaxf(v, partner<G, 2>(v));
  if constexpr (G >= 2) v = __builtin_fmaxf(v, partner<G, 1>(v));
  return v;
}

// z of lane (l + 1) mod G and (l - 1) mod G.
template <int G>
__device__ __forceinline__ float next_lane(float z, int l) {
  if constexpr (G == 16) return dpp<DPP_ROW_ROR15>(z);
  else return __shfl(z, (l + 1) & (G - 1), G);
}
template <int G>
__device__ __forceinline__ float prev_lane(float z, int l) {
  if constexpr (G == 16) return dpp<DPP_ROW_ROR1>(z);
  else return __shfl(z, (l - 1) & (G - 1), G);
}

__device__ __forceinline__ float projc(float x, float lo, float hi) {
  return jl_min(jl_max(x, lo), hi);
}

// Objective value (group-uniform) and gradient component of lane l.
template <int G, int OBJ>
__device__ __forceinline__ void lane_fg(float xl, int l, int d, const ProbArgs& P, float& f,
                                        float& g) {
  const bool act = l < d;
  const float z = act ? (P.has_shift ? xl - P.shift[l] : xl) : 0.0f;
  const float p1 = P.par[0], p2 = P.par[1];
  if constexpr (OBJ == PPS_OBJ_ROSENBROCK) {
    const float zn = next_lane<G>(z, l);  // z[l+1] (used when l+1 < d)
    const float zp = prev_lane<G>(z, l);  // z[l-1] (used when l > 0)
    float term = 0.0f, gl = 0.0f;
    if (act && l + 1 < d) {
      const float dd = zn - z * z;
      const float e = p1 - z;
      term = p2 * (dd * dd) + e * e;
      gl = (-4.0f * p2) * z * dd - 2.0f * e;
    }
    if (act && l > 0) {
      const float dp = z - zp * zp;
      gl = gl + (2.0f * p2) * dp;
    }
    f = gsum<G>(term);
    g = gl;
  } else if constexpr (OBJ == PPS_OBJ_RASTRIGIN) {
    const float term = act ? z * z - 10.0f * cos2pi(z) : 0.0f;
    g = act ? 2.0f * z + 62.83185307179586f * sin2pi(z) : 0.0f;
    f = 10.0f * (float)d + gsum<G>(term);
  } else if constexpr (OBJ == PPS_OBJ_SPHERE) {
    const float term = act ? z * z : 0.0f;
    g = act ? 2.0f * z : 0.0f;
    f = gsum<G>(term);
  } else {
    // fixed-dimension objectives: every lane evaluates the sequential formula
    float xr[2];
    xr[0] = __shfl(xl, 0, G);
    xr[1] = G > 1 ? __shfl(xl, 1 & (G - 1), G) : 0.0f;
    float xa[G > 1 ? 2 : 1];
    xa[0] = xr[0];
    if constexpr (G > 1) xa[1] = xr[1];
    f = objective<(G > 1 ? 2 : 1), OBJ>(xa, d, P, __uint_as_float(0x7FC00000u));
    const float x1 = P.has_shift ? xr[0] - P.shift[0] : xr[0];
    const float x2 = (d > 1) ? (P.has_shift ? xr[1] - P.shift[1] : xr[1]) : 0.0f;
    float g0 = 0.0f, g1 = 0.0f;
    if constexpr (OBJ == PPS_OBJ_BEALE) {
      const float a = (1.5f - x1) + x1 * x2;
      const float b = (2.25f - x1) + x1 * (x2 * x2);
      const float c = (2.625f - x1) + x1 * ((x2 * x2) * x2);
      g0 = 2.0f * a * (x2 - 1.0f) + 2.0f * b * (x2 * x2 - 1.0f) +
           2.0f * c * ((x2 * x2) * x2 - 1.0f);
      g1 = 2.0f * a * x1 + 2.0f * b * (2.0f * x1 * x2) + 2.0f * c * (3.0f * x1 * (x2 * x2));
    } else if constexpr (OBJ == PPS_OBJ_BOOTH) {
      const float a = (x1 + 2.0f * x2) - 7.0f, b = (2.0f * x1 + x2) - 5.0f;
      g0 = 2.0f * a + 4.0f * b;
      g1 = 4.0f * a + 2.0f * b;
    } else if constexpr (OBJ == PPS_OBJ_HIMMELBLAU) {
      const float a = (x1 * x1 + x2) - 11.0f, b = (x1 + x2 * x2) - 7.0f;
      g0 = 4.0f * x1 * a + 2.0f * b;
      g1 = 2.0f * a + 4.0f * x2 * b;
    } else if constexpr (OBJ == PPS_OBJ_QUADRATIC) {
      g0 = 2.0f * p1 * (x1 - 1.0f);
      g1 = 2.0f * (x2 + 2.0f);
    } else if constexpr (OBJ == PPS_OBJ_BOUNDARY_QUADRATIC) {
      g0 = 2.0f * (x1 + 1.0f);
    } else if constexpr (OBJ == PPS_OBJ_ROSEN2) {
      const float a = p1 - x1, b = x2 - x1 * x1;
      g0 = -2.0f * a - 4.0f * p2 * x1 * b;
      g1 = 2.0f * p2 * b;
    }
    g = l == 0 ? g0 : (l == 1 ? g1 : 0.0f);
  }
}

// Armijo backtracking: the first a = rho^k (k <= max_backtracks) with
// f(P(x + a p)) <= f + c g'(P(x + a p) - x).  Group-uniform result.
template <int G, int OBJ>
__device__ __forceinline__ bool armijo_search(float x, float f, float g, float p, float lo,
                                              float hi, int l, int d, const ProbArgs& P,
                                              const LbfgsArgs& A, float& xn, float& fn,
                                              float& gn) {
  float alpha = 1.0f;
  for (int bt = 0; bt <= A.max_backtracks; ++bt) {
    xn = projc(x + alpha * p, lo, hi);
    const float dx = xn - x;
    lane_fg<G, OBJ>(xn, l, d, P, fn, gn);
    const float decr = gsum<G>(g * dx);
    if (fn <= f + A.armijo_c * decr) return true;
    alpha = alpha * A.armijo_rho;
  }
  return false;
}

// Strong-Wolfe search (Nocedal & Wright Alg. 3.5 with zoom, Alg. 3.6) on
// phi(a) = f(P(x + a p)): accept a with phi(a) <= phi(0) + c1 a phi'(0) and
// |phi'(a)| <= c2 |phi'(0)|.  phi'(0) counts the lanes the projection lets
// move along p (not on a bound p points out of), phi'(a) the lanes whose
// unprojected x + a p lies strictly inside the box.  Trial steps alpha_init,
// doubled up to alpha_max (ls_maxiters of them) until one brackets an
// acceptable step; zoom's trial is the minimiser of the quadratic through
// phi(lo), phi'(lo), phi(hi) inside the middle 80 % of the bracket, else the
// midpoint; without a step meeting both conditions the last step of
// sufficient decrease is taken (none: false).  Every branch depends on group
// sums only, so the group stays together.
template <int G, int OBJ>
__device__ __forceinline__ bool wolfe_search(float x, float f, float g, float p, float lo,
                                             float hi, int l, int d, const ProbArgs& P,
                                             const LbfgsArgs& A, float& xn, float& fn,
                                             float& gn) {
  const float d0 = gsum<G>(((x <= lo && p < 0.0f) || (x >= hi && p > 0.0f)) ? 0.0f : g * p);
  if (!(d0 < 0.0f)) return false;
  const float c1 = A.armijo_c, curv = -(A.wolfe_c2 * d0);
  float xl = x, gl = g;  // point at a_lo
  float a_lo = 0.0f, f_lo = f, d_lo = d0, a_hi = 0.0f, f_hi = 0.0f;
  float a = A.alpha_init;
  float xt = x, ft = f, gt = g;
  bool bracket = false;
  for (int i = 0; i < A.ls_maxiters; ++i) {
    const float u = x + a * p;
    xt = projc(u, lo, hi);
    lane_fg<G, OBJ>(xt, l, d, P, ft, gt);
    if (ft > f + (c1 * a) * d0 || (i > 0 && ft >= f_lo)) {
      a_hi = a;
      f_hi = ft;
      bracket = true;
      break;
    }
    const float da = gsum<G>((lo < u && u < hi) ? gt * p : 0.0f);
    if (__builtin_fabsf(da) <= curv) {
      xn = xt;
      fn = ft;
      gn = gt;
      return true;
    }
    if (da >= 0.0f) {
      a_hi = a_lo;
      f_hi = f_lo;
      a_lo = a;
      f_lo = ft;
      d_lo = da;
      xl = xt;
      gl = gt;
      bracket = true;
      break;
    }
    a_lo = a;
    f_lo = ft;
    d_lo = da;
    xl = xt;
    gl = gt;
    if (!(a < A.alpha_max)) break;
    a = a + a;
    if (a > A.alpha_max) a = A.alpha_max;
  }
  if (bracket) {
    for (int j = 0; j < A.zoom_maxiters; ++j) {
      const float delta = a_hi - a_lo;
      const float den = 2.0f * ((f_hi - f_lo) - d_lo * delta);
      a = a_lo - ((d_lo * delta) * delta) / den;
      const float b0 = __builtin_fminf(a_lo, a_hi), b1 = __builtin_fmaxf(a_lo, a_hi);
      const float w = b1 - b0;
      if (!(a >= b0 + 0.1f * w && a <= b1 - 0.1f * w)) a = a_lo + 0.5f * delta;
      const float u = x + a * p;
      xt = projc(u, lo, hi);
      lane_fg<G, OBJ>(xt, l, d, P, ft, gt);
      if (ft > f + (c1 * a) * d0 || ft >= f_lo) {
        a_hi = a;
        f_hi = ft;
      } else {
        const float da = gsum<G>((lo < u && u < hi) ? gt * p : 0.0f);
        if (__builtin_fabsf(da) <= curv) {
          xn = xt;
          fn = ft;
          gn = gt;
          return true;
        }
        if (da * (a_hi - a_lo) >= 0.0f) {
          a_hi = a_lo;
          f_hi = f_lo;
        }
        a_lo = a;
        f_lo = ft;
        d_lo = da;
        xl = xt;
        gl = gt;
      }
    }
  }
  if (!(a_lo > 0.0f)) return false;
  xn = xl;
  fn = f_lo;
  gn = gl;
  return true;
}

// H = history capacity (m <= H): the pairs live in registers, so a kernel
// sized for the usual m <= 10 holds 18 fewer per lane than one sized for 16
// (71 against 90 VGPRs for 16-D Rastrigin: 7 waves per SIMD instead of 5).
template <int G, int H, int OBJ>
__global__ __launch_bounds__(256) void k_lbfgs(SwarmDev S, ProbArgs P, LbfgsArgs A, float* out_x,
                                               float* out_f) {
  const int64_t tid = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t i = tid / G;
  const int l = (int)(tid & (G - 1));
  if (i >= S.n) return;  // whole groups leave together
  const int d = P.d;
  const bool act = l < d;
  const float lo = act ? P.lb[l] : 0.0f, hi = act ? P.ub[l] : 0.0f;
  float x = projc(act ? S.pb[(int64_t)l * S.npad + i] : 0.0f, lo, hi);
  float f, g;
  lane_fg<G, OBJ>(x, l, d, P, f, g);
  float Sh[H], Yh[H], RHO[H];
#pragma unroll
  for (int h = 0; h < H; ++h) Sh[h] = Yh[h] = RHO[h] = 0.0f;
  int k = 0;
  const int m = A.history;
  for (int it = 0; it < A.maxiters; ++it) {
    const float pg = projc(x - g, lo, hi) - x;
    const float pgn = gmaxabs<G>(pg);
    if (!(pgn > A.abstol)) break;
    float q = g, a[H];
#pragma unroll
    for (int h = 0; h < H; ++h) {
      a[h] = 0.0f;
      if (h < k) {
        a[h] = RHO[h] * gsum<G>(Sh[h] * q);
        q = q - a[h] * Yh[h];
      }
    }
    float gamma = 1.0f;
    if (k > 0) gamma = gsum<G>(Sh[0] * Yh[0]) / gsum<G>(Yh[0] * Yh[0]);
    float r = gamma * q;
#pragma unroll
    for (int h = H - 1; h >= 0; --h) {
      if (h < k) {
        const float b = RHO[h] * gsum<G>(Yh[h] * r);
        r = r + (a[h] - b) * Sh[h];
      }
    }
    float p = -r;
    const float gp = gsum<G>(g * p);
    if (!(gp < 0.0f)) {
      p = -g;
      k = 0;
    }
    // Line search; if the quasi-Newton direction fails, forget the history
    // and search once more along -g.
    float fn = 0.0f, gn = 0.0f, xn = x;
    bool accepted = false;
    for (int attempt = 0; attempt < 2 && !accepted; ++attempt) {
      if (attempt == 1) {
        if (k == 0) break;
        p = -g;
        k = 0;
      }
      accepted = A.linesearch == PPS_LINESEARCH_STRONG_WOLFE
                     ? wolfe_search<G, OBJ>(x, f, g, p, lo, hi, l, d, P, A, xn, fn, gn)
                     : armijo_search<G, OBJ>(x, f, g, p, lo, hi, l, d, P, A, xn, fn, gn);
    }
    if (!accepted) break;
    const float s = xn - x, y = gn - g;
    const float ss = gmaxabs<G>(s);
    const float sy = gsum<G>(s * y), yy = gsum<G>(y * y);
    if (sy > 1e-10f * yy && sy > 0.0f) {
      const int top = k < m ? k : m - 1;
#pragma unroll
      for (int h = H - 1; h >= 1; --h) {
        if (h <= top) {
          Sh[h] = Sh[h - 1];
          Yh[h] = Yh[h - 1];
          RHO[h] = RHO[h - 1];
        }
      }
      Sh[0] = s;
      Yh[0] = y;
      RHO[0] = 1.0f / sy;
      if (k < m) ++k;
    } else {
      k = 0;  // no positive curvature along s: restart from a scaled gradient
    }
    const bool stall = A.reltol > 0.0f && __builtin_fabsf(f - fn) <= A.reltol * __builtin_fabsf(f);
    x = xn;
    g = gn;
    f = fn;
    if (!(ss > 0.0f) || stall) break;
  }
  if (act) out_x[(int64_t)l * S.npad + i] = x;
  if (l == 0) out_f[i] = __builtin_isfinite(f) ? f : __builtin_inff();
}

template <int G, int OBJ>
static hipError_t launch_lbfgs_g(const SwarmDev& S, const ProbArgs& P, const LbfgsArgs& A,
                                 float* ox, float* of, hipStream_t st) {
  const int64_t threads = S.n * G;
  const int64_t blocks = (threads + 255) / 256;
  if (A.history <= 10)
    hipLaunchKernelGGL((k_lbfgs<G, 10, OBJ>), dim3((unsigned)blocks), dim3(256), 0, st, S, P, A,
                       ox, of);
  else
    hipLaunchKernelGGL((k_lbfgs<G, LB_MAXH, OBJ>), dim3((unsigned)blocks), dim3(256), 0, st, S, P,
                       A, ox, of);
  return hipGetLastError();
}

template <int OBJ>
static hipError_t launch_lbfgs_obj(const SwarmDev& S, const ProbArgs& P, const LbfgsArgs& A,
                                   float* ox, float* of, hipStream_t st) {
  const int d = P.d;
  if (d <= 1) return launch_lbfgs_g<1, OBJ>(S, P, A, ox, of, st);
  if (d <= 2) return launch_lbfgs_g<2, OBJ>(S, P, A, ox, of, st);
  if (d <= 4) return launch_lbfgs_g<4, OBJ>(S, P, A, ox, of, st);
  if (d <= 8) return launch_lbfgs_g<8, OBJ>(S, P, A, ox, of, st);
  if (d <= 16) return launch_lbfgs_g<16, OBJ>(S, P, A, ox, of, st);
  if (d <= 32) return launch_lbfgs_g<32, OBJ>(S, P, A, ox, of, st);
  return launch_lbfgs_g<64, OBJ>(S, P, A, ox, of, st);
}

bool lbfgs_supported(int obj, int d) {
  switch (obj) {
    case PPS_OBJ_ROSENBROCK:
    case PPS_OBJ_RASTRIGIN:
    case PPS_OBJ_SPHERE: return d >= 1 && d <= PPS_MAXD;
    case PPS_OBJ_BEALE:
    case PPS_OBJ_BOOTH:
    case PPS_OBJ_HIMMELBLAU:
    case PPS_OBJ_QUADRATIC:
    case PPS_OBJ_ROSEN2: return d == 2;
    case PPS_OBJ_BOUNDARY_QUADRATIC: return d == 1;
    default: return false;
  }
}

hipError_t launch_lbfgs(const SwarmDev& S, const ProbArgs& P, const LbfgsArgs& A, float* out_x,
                        float* out_f, hipStream_t st) {
  switch (P.obj) {
    case PPS_OBJ_ROSENBROCK: return launch_lbfgs_obj<PPS_OBJ_ROSENBROCK>(S, P, A, out_x, out_f, st);
    case PPS_OBJ_RASTRIGIN: return launch_lbfgs_obj<PPS_OBJ_RASTRIGIN>(S, P, A, out_x, out_f, st);
    case PPS_OBJ_SPHERE: return launch_lbfgs_obj<PPS_OBJ_SPHERE>(S, P, A, out_x, out_f, st);
    case PPS_OBJ_BEALE: return launch_lbfgs_g<2, PPS_OBJ_BEALE>(S, P, A, out_x, out_f, st);
    case PPS_OBJ_BOOTH: return launch_lbfgs_g<2, PPS_OBJ_BOOTH>(S, P, A, out_x, out_f, st);
    case PPS_OBJ_HIMMELBLAU: return launch_lbfgs_g<2, PPS_OBJ_HIMMELBLAU>(S, P, A, out_x, out_f, st);
    case PPS_OBJ_QUADRATIC: return launch_lbfgs_g<2, PPS_OBJ_QUADRATIC>(S, P, A, out_x, out_f, st);
    case PPS_OBJ_ROSEN2: return launch_lbfgs_g<2, PPS_OBJ_ROSEN2>(S, P, A, out_x, out_f, st);
    case PPS_OBJ_BOUNDARY_QUADRATIC:
      return launch_lbfgs_g<1, PPS_OBJ_BOUNDARY_QUADRATIC>(S, P, A, out_x, out_f, st);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace pps
