// Internal declarations shared by the HIP translation units of libodehip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "odehip.h"
#include "tableaus.hpp"

namespace odehip {

// Device-resident state of one solve: the attempt loop's scalars, read by
// every kernel of an attempt and written only by the controller kernel.
struct SolverState {
  double t, dt, t0, tf;
  double err;         // last scaled error norm
  double dt_initial;
  double safety;      // 0.38^(1/(1+q))
  double inv_q1;      // 1/(1+q)
  double dt0, d1;     // initial-step scratch
  double err_prev;    // error of the last accepted step (PI controller)
  double hist_err[3], hist_dt[3], hist_acc[3];  // last attempts, newest first
  long accepted, rejected, attempts, rhs_evals;
  long max_accepted, max_attempts;
  long times_cap, trace_cap;
  long learned_steps;  // steps sized by the learned controller
  double dt_acc;       // length of the last accepted step
  long dense_next;     // first dense-output time not yet written
  long dense_lo, dense_hi;  // rows the current dense pass writes
  long dense_seen;     // accepted count the dense pass last examined
  long dense_steps, dense_rhs_evals;
  double mass_uniform;  // ensembles: the system's common mass (all equal), else 0
  int dense_go;        // the dense pass of this attempt runs
  int pad2;
  int cur;       // which ping-pong buffer holds x0
  int done;      // nonzero: every kernel returns immediately
  int status;    // ODEHIP_STATUS_*
  int k1_valid;  // 0: compute k1; 1: copy it from the FSAL buffer; 2: keep it
  int fsal;      // FSAL / k1 reuse enabled
  int tableau;
  int n_groups;  // reduction groups of the error norm
  int controller;      // ODEHIP_CONTROLLER_*
  int rejects_in_row;  // consecutive rejections up to now
  int just_accepted;   // the last judged attempt was accepted
  int pad;
};

// Arguments of the N-body RHS kernels.  Exactly one of y / (xb0, xb1, st)
// selects the input state: y if non-null, else the x0 buffer st->cur names.
struct NBodyArgs {
  const double* y;
  const double* xb0;
  const double* xb1;
  const SolverState* st;
  double* k;            // output dxdt
  const double* mass;   // nullable
  const double* kfsal;  // FSAL source for stage-0 reuse (solve only)
  int n;                // particles
  int stage0;           // 1: this launch computes k1 of an attempt;
                        // 2: a dense-output evaluation (runs iff st->dense_go)
  double eps2;
  float near2;          // mixed variant: pairs with FP32 r^2 below this are
                        // computed in FP64 (0: none)
};

// Near-field threshold of the mixed RHS for a system of n bodies: pairs with
// r^2 < coefficient / n take the FP64 term.  With masses ~1/n this selects the
// pairs whose single term is a sizeable part of a body's acceleration.
__host__ __device__ inline float mixed_near2(double coefficient, int n) {
  return n > 0 ? static_cast<float>(coefficient / n) : 0.0f;
}

// ---- launchers (nbody_kernels.hip) ----
hipError_t launch_nbody_rhs(const NBodyArgs& a, int variant, hipStream_t s);
hipError_t launch_probe_fp64(double* out, long long* clk, int blocks, int iters,
                             hipStream_t s);
hipError_t launch_momentum_terms(const double* x, int n, const double* mass,
                                 double* out, hipStream_t s);
hipError_t launch_nbody_energy(const double* x, int n, double eps2,
                               const double* mass, double* per_particle,
                               hipStream_t s);

// ---- tableaux (built in, or registered at run time; odehip_abi.hip) ----
bool valid_tableau(int id);
// The host copy of a valid tableau id.
const Tableau& host_tableau(int id);
// Constant-table slot uploads of a registered tableau, on the current device
// (rk_kernels.hip, ensemble_kernels.hip).
hipError_t upload_rk_tableau(int id, const Tableau& T);
hipError_t upload_ens_tableau(int id, const Tableau& T);

// ---- launchers (rk_kernels.hip) ----
hipError_t launch_simple_rhs(int kind, const double* y, const double* xb0,
                             const double* xb1, const SolverState* st,
                             double* k, const double* kfsal, int stage0,
                             size_t n_var, hipStream_t s);
hipError_t launch_rk_stage_host_dt(int tab, int stage, double dt, double* temp,
                                   const double* ks, const double* x0,
                                   size_t n_var, hipStream_t s);
hipError_t launch_rk_update_host_dt(int tab, double dt, double* x, double* err,
                                    const double* ks, const double* x0,
                                    size_t n_var, hipStream_t s);
hipError_t launch_rk_target(double* sc, const double* atol, const double* rtol,
                            const double* x, const double* x0, size_t n_var,
                            hipStream_t s);
hipError_t launch_init_target(double* sc, const double* x0, const double* atol,
                              const double* rtol, size_t n_var, hipStream_t s);
hipError_t launch_euler_host_dt(double* x1, const double* x0, const double* f0,
                                double dt0, size_t n_var, hipStream_t s);
hipError_t launch_difference(double* df, const double* f0, size_t n_var,
                             hipStream_t s);
hipError_t launch_norm_partials(const double* v, const double* sub,
                                const double* sc, size_t n_var,
                                double* partials, hipStream_t s);
hipError_t launch_fold_norm(const double* partials, int n_groups, size_t n_var,
                            double* out, hipStream_t s);
hipError_t launch_fill(double* dst, double v, size_t n, hipStream_t s);
hipError_t launch_sum_tree(const double* v, size_t n, double* out,
                           hipStream_t s);

// Device buffers of one single-system solve.
struct SolveBufs {
  SolverState* st;
  double* xb0;
  double* xb1;
  double* ks;
  double* ytmp0;
  double* ytmp1;
  double* kfsal;
  double* sc0;
  double* partA;
  double* partB;
  double* times;
  double* trace;
  double* states;  // accepted states (states_cap x n_var), or null
  long times_cap;
  long trace_cap;
  long states_cap;
  const double* atol;
  const double* rtol;
  size_t n_var;
  int tableau;
  int rhs_kind;
  int nbody_variant;
  int n_particles;
  double eps2;
  float near2;            // mixed RHS near-field threshold (NBodyArgs::near2)
  const double* mass;
  int controller;
  const struct MlpInfer* mlp;  // learned controller's network, or null
  double* dense_t;        // device, n_dense output times (DOP853 only)
  double* dense_out;      // device, n_dense x n_var
  long n_dense;
};

hipError_t enqueue_solve_init(const SolveBufs& b, const double* x_in,
                              double t0, double tf, long max_acc,
                              long max_att, int fsal, hipStream_t s);
hipError_t enqueue_initial_step(const SolveBufs& b, hipStream_t s);
hipError_t enqueue_attempt(const SolveBufs& b, hipStream_t s);
hipError_t enqueue_solve_finish(const SolveBufs& b, double* x_out,
                                hipStream_t s);
hipError_t enqueue_rhs(const SolveBufs& b, const double* y_fixed, double* k,
                       int stage0, hipStream_t s);

inline int n_groups_for(size_t n) {
  return static_cast<int>((n + ODEHIP_REDUCE_GROUP - 1) / ODEHIP_REDUCE_GROUP);
}

}  // namespace odehip
