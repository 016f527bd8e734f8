// pps_shared.h — constants shared by libpps device code and host runtime.
#pragma once
#include "pps.h"

// ODE estimation objectives (two-state model, four rates, observations): a
// model (Lotka-Volterra 16..20, FitzHugh-Nagumo 21..25) times a solver, in
// the same order for every model.
#define PPS_ODE_LOTKA_VOLTERRA 0
#define PPS_ODE_FITZHUGH_NAGUMO 1
#define PPS_ODE_RK4 0
#define PPS_ODE_TSIT5 1
#define PPS_ODE_ATSIT5 2
#define PPS_ODE_TSIT5_SAVEAT 3
#define PPS_ODE_ATSIT5_SAVEAT 4
constexpr bool is_ode_objective(int obj) {
  return obj >= PPS_OBJ_LOTKA_VOLTERRA && obj <= PPS_OBJ_FITZHUGH_NAGUMO_ATSIT5_SAVEAT;
}
constexpr int ode_model(int obj) {
  return obj >= PPS_OBJ_FITZHUGH_NAGUMO ? PPS_ODE_FITZHUGH_NAGUMO : PPS_ODE_LOTKA_VOLTERRA;
}
constexpr int ode_solver(int obj) {
  return obj - (obj >= PPS_OBJ_FITZHUGH_NAGUMO ? PPS_OBJ_FITZHUGH_NAGUMO : PPS_OBJ_LOTKA_VOLTERRA);
}
constexpr bool ode_adaptive(int obj) {
  return ode_solver(obj) == PPS_ODE_ATSIT5 || ode_solver(obj) == PPS_ODE_ATSIT5_SAVEAT;
}
constexpr bool ode_saveat(int obj) {
  return ode_solver(obj) == PPS_ODE_TSIT5_SAVEAT || ode_solver(obj) == PPS_ODE_ATSIT5_SAVEAT;
}

#define PPS_MAXD 64    // largest supported dimension
#define PPS_MAXPAR 8   // objective parameters carried in kernel arguments

// Philox4x32 round multipliers and Weyl key increments (Salmon et al., SC'11).
#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

// Workgroup size (= particle tile) of the multi-workgroup kernels; results do
// not depend on it (256 measured fastest against 512 and 1024).
#ifndef PPS_BLK
#define PPS_BLK 256
#endif
// The ODE objectives run 64-particle workgroups instead: each particle's loss
// is one long dependent chain, so the few waves of an estimation swarm do best
// spread one per CU (2^14 particles, RK4: 39.7 -> 37.7 ms per 300 iterations;
// the streaming objectives are much slower with 64).
constexpr int block_of(int obj) { return is_ode_objective(obj) ? 64 : PPS_BLK; }

// Third counter word = block index | tag << 16.
#define PPS_TAG_STEP 0u
#define PPS_TAG_INIT 1u
#define PPS_TAG_LHS_PERM 2u    // Latin hypercube: Feistel round function
#define PPS_TAG_LHS_JITTER 3u  // Latin hypercube: offset inside the stratum
// Internal sampler value: the x rows already hold the initial positions.
#define PPS_SAMPLER_POSITIONS 3

// Taylor coefficients (-1)^k (2 pi)^(2k) / (2k)! rounded to FP32, k = 1..7.
#define COS2PI_C1 (-0x1.3bd3ccp+4f)
#define COS2PI_C2 (0x1.03c1f0p+6f)
#define COS2PI_C3 (-0x1.55d3c8p+6f)
#define COS2PI_C4 (0x1.e1f506p+5f)
#define COS2PI_C5 (-0x1.a6d1f2p+4f)
#define COS2PI_C6 (0x1.f9d38ap+2f)
#define COS2PI_C7 (-0x1.b6e250p+0f)

#include <stdint.h>

namespace pps {

// Problem constants passed by value as a kernel argument (a few hundred bytes
// of kernarg, read through scalar loads).
struct ProbArgs {
  int32_t obj;
  int32_t d;
  int32_t has_shift;
  int32_t clamp;  // PPS_CLAMP_*
  float lb[PPS_MAXD];
  float ub[PPS_MAXD];
  float shift[PPS_MAXD];
  float par[PPS_MAXPAR];
  const float* obs;  // Lotka-Volterra observations (device), 2*n_obs floats
  int32_t n_obs;
  int32_t steps_per_obs;
  // Inequality constraints (device): n_cons rows of 2d+1 floats
  // [a_0..a_{d-1}, q_0..q_{d-1}, c], g(x) = sum a_j x_j + sum q_j x_j^2 + c.
  const float* cons;
  int32_t n_cons;
  // Adaptive-step ODE losses stop summing once they reach the particle's best
  // cost (1), or are always summed in full (0).  The swarm state is the same
  // either way.
  int32_t ode_stop;
};

// Global-best record in device memory (async and queue-lock modes keep their
// live global best here; every mode leaves its final result here).
struct GbRec {
  float cost;
  uint32_t idx;
  uint32_t lock;     // queue-lock mutex word
  uint32_t err;      // set when a bounded lock wait gave up
  float pos[PPS_MAXD];
};

// Device view of one swarm.  SoA: coordinate j of particle i is at
// [j * npad + i]; npad is n rounded up to 64.
//
// Hand-over of the global best between sync iterations: each of the G
// workgroups owns the particle tiles b, b+G, b+2G, ... (tile = `block`
// consecutive particles) and, after stepping them, writes its block argmin key
// bkey[par][b] and the winner's position cand[par][b][0..d-1]; iteration t
// writes parity t & 1 and reads parity (t-1) & 1, so no workgroup reads what
// another is writing within one launch.
struct SwarmDev {
  float* x;
  float* v;
  float* pb;
  float* pc;
  int64_t n;
  int64_t npad;
  uint64_t* bkey;  // [2][G]
  float* cand;     // [2][G][d]
  GbRec* gbrec;
  // Position in the random stream, in device memory: launch k of a run
  // executes iteration t = tbase[0] + k of epoch tbase[1], so a captured
  // iteration loop can be replayed from any starting iteration and after any
  // number of reinitialisations.  The host writes both words before each run.
  // tbase[2] is the offset into wtab of the current chunk of a chunk graph.
  const uint32_t* tbase;
  // Inertia weights of the current run's iterations (device; w_k as the
  // host's iterated FP32 product), read by launches with StepArgs::w_table.
  const float* wtab;
  uint32_t key0, key1;  // Philox key (seed)
  uint32_t epoch;       // epoch of the kernels launched outside a graph
  int32_t G;
  int32_t block;
  int32_t sampler;   // PPS_SAMPLER_* of the next initialisation
  const float* u0;   // d floats (device): the unbounded sampler's centre
};

// Local-search (projected L-BFGS) settings passed to the kernel.
struct LbfgsArgs {
  int32_t maxiters;
  int32_t history;
  float abstol;
  float armijo_c;
  float armijo_rho;
  int32_t max_backtracks;
  int32_t linesearch;  // PPS_LINESEARCH_*
  float wolfe_c2;
  float alpha_init;
  float alpha_max;
  int32_t ls_maxiters;
  int32_t zoom_maxiters;
  float reltol;
};

}  // namespace pps
