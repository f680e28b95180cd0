// Embedded Runge-Kutta kernels on gfx950: stage combination, the fused
// final-state / error / error-target / norm pass, the device-resident
// accept/reject + step-size controller, and the device initial-step
// estimate.  Semantics follow src/integrator/RKEmbeddedParallel/
// RKEmbeddedParallel.hpp; each kernel cites the lines it replaces.  Compiled
// with -ffp-contract=off so every expression rounds exactly as the
// reference's.
#include <cmath>

#include "mlp.hpp"
#include "rk_device.hpp"

namespace odehip {
namespace {

// built-in tableaux, then the slots of registered ones
__constant__ Tableau c_tab[ODEHIP_MAX_TABLEAUS] = {
    kTableaus[0], kTableaus[1], kTableaus[2],
    kTableaus[3], kTableaus[4], kTableaus[5]};

__constant__ Dop853Dense c_dense = kDop853Dense;
__constant__ Dopri5Dense c_dense5 = kDopri5Dense;

constexpr int kBlock = 256;

__device__ inline size_t gid() {
  return static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
}

// ----------------------------------------------------------- simple RHS ----
__global__ void k_exp_rhs(const double* y, const double* xb0, const double* xb1,
                          const SolverState* st, double* k, const double* kfsal,
                          int stage0, size_t n_var) {
  if (st && (stage0 == 2 ? st->dense_go == 0 : st->done)) return;
  const size_t i = gid();
  if (st && stage0 == 1 && st->k1_valid) {
    if (st->k1_valid == 1 && i < n_var) k[i] = kfsal[i];
    return;
  }
  const double* x = y ? y : (st->cur ? xb1 : xb0);
  if (i < n_var) k[i] = x[i];  // ExponentialODE.hpp:16-18
}

__global__ void k_vdp_rhs(const double* y, const double* xb0, const double* xb1,
                          const SolverState* st, double* k, const double* kfsal,
                          int stage0) {
  if (st && (stage0 == 2 ? st->dense_go == 0 : st->done)) return;
  const int i = threadIdx.x;
  if (st && stage0 == 1 && st->k1_valid) {
    if (st->k1_valid == 1 && i < 2) k[i] = kfsal[i];
    return;
  }
  const double* x = y ? y : (st->cur ? xb1 : xb0);
  // VanDerPolODE.hpp:16-25 (eps = 1)
  if (i == 0) {
    k[0] = x[1];
  } else if (i == 1) {
    const double eps = 1.0;
    k[1] = eps * (1.0 - x[0] * x[0]) * x[1] - x[0];
  }
}

// ------------------------------------------------------- stage kernels -----
__global__ void k_stage_host_dt(int tab, int stage, double dt, double* temp,
                                const double* ks, const double* x0,
                                size_t n_var) {
  const size_t i = gid();
  if (i < n_var) temp[i] = stage_value(c_tab[tab], stage, dt, ks, x0, i, n_var);
}

__global__ void k_stage_dev(const SolverState* st, int tab, int stage,
                            double* temp, const double* ks, const double* xb0,
                            const double* xb1, size_t n_var) {
  if (st->done) return;
  const size_t i = gid();
  const double* x0 = st->cur ? xb1 : xb0;
  if (i < n_var) temp[i] = stage_value(c_tab[tab], stage, st->dt, ks, x0, i, n_var);
}

__global__ void k_update_host_dt(int tab, double dt, double* x, double* err,
                                 const double* ks, const double* x0,
                                 size_t n_var) {
  const size_t i = gid();
  if (i < n_var) {
    double xv, ev;
    update_values(c_tab[tab], dt, ks, x0, i, n_var, xv, ev);
    x[i] = xv;
    err[i] = ev;
  }
}

// detail::update_error_target_kernel (RKEmbeddedParallel.hpp:125-130).
__global__ void k_target(double* sc, const double* atol, const double* rtol,
                         const double* x, const double* x0, size_t n_var) {
  const size_t i = gid();
  if (i < n_var) sc[i] = atol[i] + rtol[i] * max_abs(x[i], x0[i]);
}

// Fused final stage of an attempt: new state, embedded error, error target
// and the scaled squared error folded into this group's partial -- one
// coalesced pass over ks (RKEmbeddedParallel.hpp:47-55, 103-137).
__global__ __launch_bounds__(kBlock) void k_final_dev(
    const SolverState* st, int tab, const double* ks, double* xb0, double* xb1,
    const double* atol, const double* rtol, double* partials, double* kfsal,
    size_t n_var) {
  if (st->done) return;
  const Tableau& T = c_tab[tab];
  const int cur = st->cur;
  const double* x0 = cur ? xb1 : xb0;
  double* x = cur ? xb0 : xb1;
  const double dt = st->dt;
  const size_t i = gid();
  double term = 0.0;
  if (i < n_var) {
    double xv, ev;
    update_values(T, dt, ks, x0, i, n_var, xv, ev);
    x[i] = xv;
    const double sc = atol[i] + rtol[i] * max_abs(xv, x0[i]);
    const double scaled = ev / sc;  // scaled_value_squared_kernel
    term = scaled * scaled;
    if (st->fsal && T.fsal) kfsal[i] = ks[(T.S - 1) * n_var + i];
  }
  const double g = group_reduce(term);
  if (threadIdx.x == 0) partials[blockIdx.x] = g;
}

// Accept/reject and step-size control (RKEmbeddedParallel.hpp:57-79) on the
// device, plus the termination tests and counters.
__global__ void k_controller(SolverState* st, const double* partials,
                             size_t n_var, double* times, double* trace) {
  if (st->done) return;
  const double sum = fold_partials_block(partials, st->n_groups);
  if (threadIdx.x != 0) return;
  ctrl_attempt(st, c_tab[st->tableau], sum, n_var, times, trace);
}

// The same with the learned controller (a tile of one system).
__global__ __launch_bounds__(64) void k_controller_learned(
    SolverState* st, const double* partials, size_t n_var, double* times,
    double* trace, const MlpInfer* M, int n_particles) {
  __shared__ MlpTileLds tile;
  if (st->done) return;  // read by every thread before thread 0 writes it
  const double sum = fold_partials_block(partials, st->n_groups);
  bool accept = false, judged = false;
  float raw[kMlpIn];
  for (int i = 0; i < kMlpIn; ++i) raw[i] = 0.0f;
  __syncthreads();
  if (threadIdx.x == 0) {
    judged = ctrl_judge(st, c_tab[st->tableau], sum, n_var, times, trace, &accept);
    if (judged) mlp_features(st, n_particles, raw);
  }
  if (threadIdx.x < 16) mlp_tile_set_row(M, tile, threadIdx.x, raw);
  __syncthreads();
  mlp_tile_forward(M, tile);
  if (threadIdx.x == 0 && judged) {
    double y = tile.y[0];
    y = y < kMlpLogRatioMin ? kMlpLogRatioMin : (y > kMlpLogRatioMax ? kMlpLogRatioMax : y);
    ctrl_step(st, accept, std::isfinite(y) ? std::exp(y) : NAN);
  }
}

// ------------------------------------------------------ initial step -------
__global__ void k_init_state(SolverState* st, double t0, double tf,
                             long max_acc, long max_att, int fsal, int tab,
                             int n_groups, double* times, long times_cap,
                             long trace_cap, int controller) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  ctrl_init(st, c_tab[tab], t0, tf, max_acc, max_att, fsal, tab, n_groups,
            times_cap, trace_cap, controller);
  if (times && times_cap > 0) times[0] = t0;
}

// compute_error_target_kernel (RKEmbeddedParallel.hpp:151-157).
__global__ void k_init_target(double* sc, const double* x0, const double* atol,
                              const double* rtol, size_t n_var) {
  const size_t i = gid();
  if (i < n_var) sc[i] = std::fabs(x0[i]) * rtol[i] + atol[i];
}

// Group partials of sum(((v - sub) / sc)^2) in the canonical order; sub may
// be null.  difference_kernel (:165-168) is folded in as v - sub.
__global__ __launch_bounds__(kBlock) void k_norm_partials(
    const double* v, const double* sub, const double* sc, size_t n_var,
    double* partials) {
  const size_t i = gid();
  double term = 0.0;
  if (i < n_var) {
    const double d = sub ? (v[i] - sub[i]) : v[i];
    const double s = d / sc[i];
    term = s * s;
  }
  const double g = group_reduce(term);
  if (threadIdx.x == 0) partials[blockIdx.x] = g;
}

__global__ void k_fold_sum(const double* partials, int n_groups, double* out) {
  const double s = fold_partials_block(partials, n_groups);
  if (threadIdx.x == 0) *out = s;
}

__global__ void k_fold_norm(const double* partials, int n_groups, size_t n_var,
                            double* out) {
  const double s = fold_partials_block(partials, n_groups);
  if (threadIdx.x != 0) return;
  *out = std::sqrt(s / static_cast<double>(static_cast<long>(n_var)));
}

// estimate_initial_step (RKEmbeddedParallel.hpp:183-185): d0, d1, dt0.
__global__ void k_init_step_a(SolverState* st, const double* pa,
                              const double* pb, size_t n_var) {
  const double sa = fold_partials_block(pa, st->n_groups);
  __syncthreads();
  const double sb = fold_partials_block(pb, st->n_groups);
  if (threadIdx.x != 0) return;
  ctrl_init_step_a(st, sa, sb, n_var);
}

// Host-dt forms of the initial-step kernels (:151-168).
__global__ void k_euler_host_dt(double* x1, const double* x0, const double* f0,
                                double dt0, size_t n_var) {
  const size_t i = gid();
  if (i < n_var) x1[i] = x0[i] + f0[i] * dt0;
}

__global__ void k_difference(double* df, const double* f0, size_t n_var) {
  const size_t i = gid();
  if (i < n_var) df[i] -= f0[i];
}

// euler_step_kernel (:159-163).
__global__ void k_euler(const SolverState* st, double* x1, const double* x0,
                        const double* f0, size_t n_var) {
  const size_t i = gid();
  if (i < n_var) x1[i] = x0[i] + f0[i] * st->dt0;
}

// (:193-199): d2, dt1, the initial step.
__global__ void k_init_step_b(SolverState* st, const double* pa, size_t n_var) {
  const double sa = fold_partials_block(pa, st->n_groups);
  if (threadIdx.x != 0) return;
  ctrl_init_step_b(st, c_tab[st->tableau], sa, n_var);
}

// RawOutput::save_state of an accepted step (RawOutput.hpp:15-19): the new
// x0 into slot `accepted` of the device state list.  Runs after every
// controller launch; once a solve is done it rewrites the same slot with the
// same state (harmless).
__global__ void k_save_state(const SolverState* st, const double* xb0,
                             const double* xb1, double* states, long cap,
                             size_t n_var) {
  if (!st->just_accepted) return;
  const long slot = st->accepted;
  if (slot >= cap) return;
  const size_t i = gid();
  const double* x0 = st->cur ? xb1 : xb0;
  if (i < n_var) states[static_cast<size_t>(slot) * n_var + i] = x0[i];
}

// ------------------------------------------------------- dense output ----
// DOP853's continuous extension after an accepted step (the reference's
// StepperDopr853::prepare_dense / dense_out, StepperDopr853.hpp:200-289,
// driven by DenseOutput::save, Output.hpp:99-114).  When the pass runs, the
// controller has already swapped the ping-pong buffers: the step went from
// the buffer st->cur does not name to the one it names, over st->dt_acc,
// ending at st->t.

// The gate of the dense pass (one thread).
__global__ void k_dense_begin(SolverState* st, const double* times, long n_dense,
                              int extra_rhs) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  dense_gate(st, times, n_dense, extra_rhs);
}

// The input of extra stage 13 + x (StepperDopr853.hpp:239-262).
__global__ void k_dense_stage(const SolverState* st, int x, double* ytmp,
                              const double* ks, const double* k13,
                              const double* xb0, const double* xb1, size_t n) {
  if (!st->dense_go) return;
  const size_t i = gid();
  if (i >= n) return;
  ytmp[i] = dop853_dense_stage(c_dense, x, ks, k13, n, i, st->cur ? xb0 : xb1, st->dt_acc);
}

// The eight polynomial coefficients of element i and the rows of this pass.
__global__ __launch_bounds__(kBlock) void k_dense_out(
    const SolverState* st, const double* ks, const double* k13,
    const double* xb0, const double* xb1, const double* times, double* out,
    size_t n) {
  if (!st->dense_go) return;
  const size_t i = gid();
  if (i >= n) return;
  dop853_dense_rows(c_dense, st, ks, k13, n, i, st->cur ? xb0 : xb1, st->cur ? xb1 : xb0,
                    times, out, n);
}

// DOPRI5: the five coefficients of element i from stages 1, 3..7 -- stage 7
// is f at the new state, its input being the new state bit for bit -- and the
// rows of this pass.
__global__ __launch_bounds__(kBlock) void k_dense_out5(
    const SolverState* st, const double* ks, const double* xb0, const double* xb1,
    const double* times, double* out, size_t n) {
  if (!st->dense_go) return;
  const size_t i = gid();
  if (i >= n) return;
  dopri5_dense_rows(c_dense5, st, ks, n, i, st->cur ? xb0 : xb1, st->cur ? xb1 : xb0,
                    times, out, n);
}

// Final state out: x_out = the x0 buffer.
__global__ void k_copy_out(const SolverState* st, const double* xb0,
                           const double* xb1, double* out, size_t n_var) {
  const size_t i = gid();
  const double* x0 = st->cur ? xb1 : xb0;
  if (i < n_var) out[i] = x0[i];
}

__global__ void k_fill(double* dst, double v, size_t n) {
  const size_t i = gid();
  if (i < n) dst[i] = v;
}

// Fixed-shape pairwise tree sum (used for energy / momentum totals).
__global__ __launch_bounds__(kBlock) void k_sum_partials(const double* v,
                                                         size_t n,
                                                         double* partials) {
  const size_t i = gid();
  const double g = group_reduce(i < n ? v[i] : 0.0);
  if (threadIdx.x == 0) partials[blockIdx.x] = g;
}

inline unsigned blocks_for(size_t n) {
  return static_cast<unsigned>((n + kBlock - 1) / kBlock);
}

}  // namespace

// ---------------------------------------------------------------- launchers --

hipError_t launch_simple_rhs(int kind, const double* y, const double* xb0,
                             const double* xb1, const SolverState* st,
                             double* k, const double* kfsal, int stage0,
                             size_t n_var, hipStream_t s) {
  if (kind == ODEHIP_RHS_EXPONENTIAL) {
    hipLaunchKernelGGL(k_exp_rhs, dim3(blocks_for(n_var)), dim3(kBlock), 0, s,
                       y, xb0, xb1, st, k, kfsal, stage0, n_var);
  } else if (kind == ODEHIP_RHS_VANDERPOL) {
    if (n_var != 2) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_vdp_rhs, dim3(1), dim3(64), 0, s, y, xb0, xb1, st, k,
                       kfsal, stage0);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_rk_stage_host_dt(int tab, int stage, double dt, double* temp,
                                   const double* ks, const double* x0,
                                   size_t n_var, hipStream_t s) {
  hipLaunchKernelGGL(k_stage_host_dt, dim3(blocks_for(n_var)), dim3(kBlock), 0,
                     s, tab, stage, dt, temp, ks, x0, n_var);
  return hipGetLastError();
}

hipError_t launch_rk_update_host_dt(int tab, double dt, double* x, double* err,
                                    const double* ks, const double* x0,
                                    size_t n_var, hipStream_t s) {
  hipLaunchKernelGGL(k_update_host_dt, dim3(blocks_for(n_var)), dim3(kBlock), 0,
                     s, tab, dt, x, err, ks, x0, n_var);
  return hipGetLastError();
}

hipError_t launch_rk_target(double* sc, const double* atol, const double* rtol,
                            const double* x, const double* x0, size_t n_var,
                            hipStream_t s) {
  hipLaunchKernelGGL(k_target, dim3(blocks_for(n_var)), dim3(kBlock), 0, s, sc,
                     atol, rtol, x, x0, n_var);
  return hipGetLastError();
}

hipError_t launch_init_target(double* sc, const double* x0, const double* atol,
                              const double* rtol, size_t n_var, hipStream_t s) {
  hipLaunchKernelGGL(k_init_target, dim3(blocks_for(n_var)), dim3(kBlock), 0, s, sc,
                     x0, atol, rtol, n_var);
  return hipGetLastError();
}

hipError_t launch_euler_host_dt(double* x1, const double* x0, const double* f0,
                                double dt0, size_t n_var, hipStream_t s) {
  hipLaunchKernelGGL(k_euler_host_dt, dim3(blocks_for(n_var)), dim3(kBlock), 0, s,
                     x1, x0, f0, dt0, n_var);
  return hipGetLastError();
}

hipError_t launch_difference(double* df, const double* f0, size_t n_var,
                             hipStream_t s) {
  hipLaunchKernelGGL(k_difference, dim3(blocks_for(n_var)), dim3(kBlock), 0, s, df,
                     f0, n_var);
  return hipGetLastError();
}

hipError_t launch_norm_partials(const double* v, const double* sub,
                                const double* sc, size_t n_var,
                                double* partials, hipStream_t s) {
  hipLaunchKernelGGL(k_norm_partials, dim3(blocks_for(n_var)), dim3(kBlock), 0,
                     s, v, sub, sc, n_var, partials);
  return hipGetLastError();
}

hipError_t launch_fold_norm(const double* partials, int n_groups, size_t n_var,
                            double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_fold_norm, dim3(1), dim3(64), 0, s, partials, n_groups,
                     n_var, out);
  return hipGetLastError();
}

hipError_t launch_fill(double* dst, double v, size_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill, dim3(blocks_for(n)), dim3(kBlock), 0, s, dst, v, n);
  return hipGetLastError();
}

// out[0] = sum of v in the canonical order (out must have room for the group
// partials after out[0]: n_groups_for(n) entries at out + 1).
hipError_t launch_sum_tree(const double* v, size_t n, double* out,
                           hipStream_t s) {
  const int ng = n_groups_for(n);
  hipLaunchKernelGGL(k_sum_partials, dim3(ng), dim3(kBlock), 0, s, v, n, out + 1);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_fold_sum, dim3(1), dim3(64), 0, s, out + 1, ng, out);
  return hipGetLastError();
}

hipError_t enqueue_rhs(const SolveBufs& b, const double* y_fixed, double* k,
                       int stage0, hipStream_t s) {
  if (b.rhs_kind == ODEHIP_RHS_NBODY) {
    NBodyArgs a{};
    a.y = y_fixed;
    a.xb0 = b.xb0;
    a.xb1 = b.xb1;
    a.st = b.st;
    a.k = k;
    a.mass = b.mass;
    a.kfsal = b.kfsal;
    a.n = b.n_particles;
    a.stage0 = stage0;
    a.eps2 = b.eps2;
    a.near2 = b.near2;
    return launch_nbody_rhs(a, b.nbody_variant, s);
  }
  return launch_simple_rhs(b.rhs_kind, y_fixed, b.xb0, b.xb1, b.st, k, b.kfsal,
                           stage0, b.n_var, s);
}

hipError_t enqueue_solve_init(const SolveBufs& b, const double* x_in,
                              double t0, double tf, long max_acc, long max_att,
                              int fsal, hipStream_t s) {
  hipError_t e = hipMemcpyAsync(b.xb0, x_in, b.n_var * sizeof(double),
                                hipMemcpyDeviceToDevice, s);
  if (e != hipSuccess) return e;
  if (b.states && b.states_cap > 0) {
    e = hipMemcpyAsync(b.states, x_in, b.n_var * sizeof(double), hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_init_state, dim3(1), dim3(64), 0, s, b.st, t0, tf,
                     max_acc, max_att, fsal, b.tableau, n_groups_for(b.n_var),
                     b.times, b.times_cap, b.trace_cap, b.controller);
  return hipGetLastError();
}

hipError_t enqueue_initial_step(const SolveBufs& b, hipStream_t s) {
  const size_t n = b.n_var;
  double* f0 = b.ks;
  double* f1 = b.ks + n;
  hipLaunchKernelGGL(k_init_target, dim3(blocks_for(n)), dim3(kBlock), 0, s,
                     b.sc0, b.xb0, b.atol, b.rtol, n);
  hipError_t e = enqueue_rhs(b, b.xb0, f0, 0, s);
  if (e != hipSuccess) return e;
  if ((e = launch_norm_partials(b.xb0, nullptr, b.sc0, n, b.partA, s)) != hipSuccess) return e;
  if ((e = launch_norm_partials(f0, nullptr, b.sc0, n, b.partB, s)) != hipSuccess) return e;
  hipLaunchKernelGGL(k_init_step_a, dim3(1), dim3(64), 0, s, b.st, b.partA,
                     b.partB, n);
  hipLaunchKernelGGL(k_euler, dim3(blocks_for(n)), dim3(kBlock), 0, s, b.st,
                     b.ytmp0, b.xb0, f0, n);
  if ((e = enqueue_rhs(b, b.ytmp0, f1, 0, s)) != hipSuccess) return e;
  if ((e = launch_norm_partials(f1, f0, b.sc0, n, b.partA, s)) != hipSuccess) return e;
  hipLaunchKernelGGL(k_init_step_b, dim3(1), dim3(64), 0, s, b.st, b.partA, n);
  return hipGetLastError();
}

hipError_t enqueue_attempt(const SolveBufs& b, hipStream_t s) {
  const Tableau& T = host_tableau(b.tableau);
  const size_t n = b.n_var;
  hipError_t e = enqueue_rhs(b, nullptr, b.ks, 1, s);
  if (e != hipSuccess) return e;
  for (int stage = 1; stage < T.S; ++stage) {
    double* temp = (stage & 1) ? b.ytmp1 : b.ytmp0;
    hipLaunchKernelGGL(k_stage_dev, dim3(blocks_for(n)), dim3(kBlock), 0, s,
                       b.st, b.tableau, stage, temp, b.ks, b.xb0, b.xb1, n);
    if ((e = enqueue_rhs(b, temp, b.ks + stage * n, 0, s)) != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_final_dev, dim3(blocks_for(n)), dim3(kBlock), 0, s, b.st,
                     b.tableau, b.ks, b.xb0, b.xb1, b.atol, b.rtol, b.partA,
                     b.kfsal, n);
  if (b.controller == ODEHIP_CONTROLLER_LEARNED) {
    hipLaunchKernelGGL(k_controller_learned, dim3(1), dim3(64), 0, s, b.st, b.partA, n,
                       b.times, b.trace, b.mlp, b.n_particles);
  } else {
    hipLaunchKernelGGL(k_controller, dim3(1), dim3(64), 0, s, b.st, b.partA, n, b.times,
                       b.trace);
  }
  if (b.states && b.states_cap > 0)
    hipLaunchKernelGGL(k_save_state, dim3(blocks_for(n)), dim3(kBlock), 0, s, b.st, b.xb0,
                       b.xb1, b.states, b.states_cap, n);
  if (b.n_dense > 0 && b.tableau == ODEHIP_TABLEAU_DOPRI5) {
    hipLaunchKernelGGL(k_dense_begin, dim3(1), dim3(64), 0, s, b.st, b.dense_t, b.n_dense, 0);
    hipLaunchKernelGGL(k_dense_out5, dim3(blocks_for(n)), dim3(kBlock), 0, s, b.st, b.ks,
                       b.xb0, b.xb1, b.dense_t, b.dense_out, n);
  } else if (b.n_dense > 0) {
    hipLaunchKernelGGL(k_dense_begin, dim3(1), dim3(64), 0, s, b.st, b.dense_t, b.n_dense, 4);
    if ((e = enqueue_rhs(b, nullptr, b.kfsal, 2, s)) != hipSuccess) return e;
    for (int x = 0; x < 3; ++x) {
      double* temp = (x & 1) ? b.ytmp1 : b.ytmp0;
      hipLaunchKernelGGL(k_dense_stage, dim3(blocks_for(n)), dim3(kBlock), 0, s, b.st, x,
                         temp, b.ks, b.kfsal, b.xb0, b.xb1, n);
      if ((e = enqueue_rhs(b, temp, b.ks + (x + 1) * n, 2, s)) != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_dense_out, dim3(blocks_for(n)), dim3(kBlock), 0, s, b.st, b.ks,
                       b.kfsal, b.xb0, b.xb1, b.dense_t, b.dense_out, n);
  }
  return hipGetLastError();
}

hipError_t enqueue_solve_finish(const SolveBufs& b, double* x_out,
                                hipStream_t s) {
  hipLaunchKernelGGL(k_copy_out, dim3(blocks_for(b.n_var)), dim3(kBlock), 0, s,
                     b.st, b.xb0, b.xb1, x_out, b.n_var);
  return hipGetLastError();
}

hipError_t upload_rk_tableau(int id, const Tableau& T) {
  if (id < ODEHIP_NUM_TABLEAUS || id >= ODEHIP_MAX_TABLEAUS) return hipErrorInvalidValue;
  return hipMemcpyToSymbol(HIP_SYMBOL(c_tab), &T, sizeof(Tableau), id * sizeof(Tableau),
                           hipMemcpyHostToDevice);
}

}  // namespace odehip
