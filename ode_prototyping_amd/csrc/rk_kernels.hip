// Embedded Runge-Kutta kernels on gfx950: stage combination, the fused
// final-state / error / error-target / norm pass, the device-resident
// accept/reject + step-size controller, and the device initial-step
// estimate.  Semantics follow src/integrator/RKEmbeddedParallel/
// RKEmbeddedParallel.hpp; each kernel cites the lines it replaces.  Compiled
// with -ffp-contract=off so every expression rounds exactly as the
// reference's.
#include <cmath>

#include "crpow.hpp"
#include "internal.hpp"

namespace odehip {
namespace {

__constant__ Tableau c_tab[ODEHIP_NUM_TABLEAUS] = {
    kTableaus[0], kTableaus[1], kTableaus[2],
    kTableaus[3], kTableaus[4], kTableaus[5]};

constexpr int kBlock = 256;

__device__ inline size_t gid() {
  return static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
}

__device__ inline double max_abs(double a, double b) {
  // std::max(std::abs(x), std::abs(x0)): the first argument unless it is
  // less than the second.
  const double fa = std::fabs(a), fb = std::fabs(b);
  return (fa < fb) ? fb : fa;
}

// Canonical group reduction (odehip.h, ODEHIP_REDUCE_GROUP): 256 threads; each
// 64-lane quad reduced with halving xor strides (lane 0 holds
// v0 + v32, + (v16 + v48), ...), then (q0 + q1) + (q2 + q3).
__device__ inline double group_reduce(double v) {
  __shared__ double q[4];
#pragma unroll
  for (int s = 32; s >= 1; s >>= 1) v = v + __shfl_xor(v, s, 64);
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  if (lane == 0) q[wave] = v;
  __syncthreads();
  return (q[0] + q[1]) + (q[2] + q[3]);
}

// Sequential fold of group partials from 0.0 (the last level of the
// canonical order).
__device__ inline double fold_partials(const double* p, int n_groups) {
  double r = 0.0;
  for (int g = 0; g < n_groups; ++g) r = r + p[g];
  return r;
}

// The same fold, called by every thread of a 64-thread block: the partials
// are first staged through LDS by all lanes (coalesced), then lane 0 adds
// them in order from LDS.  Returns the sum in lane 0.
constexpr int kFoldLds = 4096;
__device__ inline double fold_partials_block(const double* p, int n_groups) {
  __shared__ double buf[kFoldLds];
  if (n_groups > kFoldLds) return threadIdx.x == 0 ? fold_partials(p, n_groups) : 0.0;
  for (int g = threadIdx.x; g < n_groups; g += blockDim.x) buf[g] = p[g];
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll 8
    for (int g = 0; g < n_groups; ++g) r = r + buf[g];
  }
  return r;
}

// ----------------------------------------------------------- simple RHS ----
__global__ void k_exp_rhs(const double* y, const double* xb0, const double* xb1,
                          const SolverState* st, double* k, const double* kfsal,
                          int stage0, size_t n_var) {
  if (st && st->done) return;
  const size_t i = gid();
  if (st && stage0 && st->k1_valid) {
    if (st->k1_valid == 1 && i < n_var) k[i] = kfsal[i];
    return;
  }
  const double* x = y ? y : (st->cur ? xb1 : xb0);
  if (i < n_var) k[i] = x[i];  // ExponentialODE.hpp:16-18
}

__global__ void k_vdp_rhs(const double* y, const double* xb0, const double* xb1,
                          const SolverState* st, double* k, const double* kfsal,
                          int stage0) {
  if (st && st->done) return;
  const int i = threadIdx.x;
  if (st && stage0 && st->k1_valid) {
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
// detail::rk_stage_kernel (RKEmbeddedParallel.hpp:91-101).
__device__ inline double stage_value(const Tableau& T, int stage, double dt,
                                     const double* ks, const double* x0,
                                     size_t i, size_t n_var) {
  double tmp = 0.0;
  for (int j = 0; j < stage; ++j) tmp += T.a[stage - 1][j] * ks[j * n_var + i];
  return x0[i] + tmp * dt;
}

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

// detail::update_state_and_error_kernel (RKEmbeddedParallel.hpp:103-123).
__device__ inline void update_values(const Tableau& T, double dt,
                                     const double* ks, const double* x0,
                                     size_t i, size_t n_var, double& x,
                                     double& err) {
  double xs = 0.0, es = 0.0;
  for (int j = 0; j < T.S; ++j) {
    const double kj = ks[j * n_var + i];
    xs += T.b[j] * kj;
    es += T.db[j] * kj;
  }
  x = x0[i] + xs * dt;
  err = es * dt;
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
  const Tableau& T = c_tab[st->tableau];
  const double scaled_error =
      std::sqrt(sum / static_cast<double>(static_cast<long>(n_var)));
  const long attempt = st->attempts;
  st->attempts = attempt + 1;
  st->rhs_evals += st->k1_valid ? (T.S - 1) : T.S;
  st->err = scaled_error;
  double t = st->t;
  double dt = st->dt;
  const bool finite = std::isfinite(scaled_error);
  const bool accept = finite && scaled_error <= 1.0;
  if (trace && attempt < st->trace_cap) {
    trace[3 * attempt] = scaled_error;
    trace[3 * attempt + 1] = dt;
    trace[3 * attempt + 2] = accept ? 1.0 : 0.0;
  }
  if (!finite) {
    st->rejected += 1;
    st->status = ODEHIP_STATUS_NONFINITE;
    st->done = 1;
    return;
  }
  if (accept) {
    t += dt;
    st->cur ^= 1;  // x0 = x (the ping-pong buffers swap roles)
    const long acc = st->accepted + 1;
    st->accepted = acc;
    if (times && acc < st->times_cap) times[acc] = t;
    st->k1_valid = (st->fsal && T.fsal) ? 1 : 0;
  } else {
    st->rejected += 1;
    st->k1_valid = st->fsal ? 2 : 0;
  }
  const double max_step_scale = 6.0;
  const double min_step_scale = 0.33;
  const double tf = st->tf, t0 = st->t0;
  const double dtnew = dt * st->safety / cr_pow(scaled_error, st->inv_q1);
  if (std::fabs(dtnew) > max_step_scale * std::fabs(dt)) {
    dt *= max_step_scale;
  } else if (std::fabs(dtnew) < min_step_scale * std::fabs(dt)) {
    dt *= min_step_scale;
  } else if (dtnew / (tf - t0) < 1.0e-12) {
    dt = (tf - t0) * 1.0e-12;
  } else {
    dt = dtnew;
  }
  if (t + dt > tf) dt = tf - t;
  st->t = t;
  st->dt = dt;
  if (!(t < tf)) {
    st->status = ODEHIP_STATUS_REACHED_END;
    st->done = 1;
  } else if (accept && st->max_accepted > 0 && st->accepted >= st->max_accepted) {
    st->status = ODEHIP_STATUS_MAX_ACCEPTED;
    st->done = 1;
  } else if (st->max_attempts > 0 && st->attempts >= st->max_attempts) {
    st->status = ODEHIP_STATUS_MAX_ATTEMPTS;
    st->done = 1;
  }
}

// ------------------------------------------------------ initial step -------
__global__ void k_init_state(SolverState* st, double t0, double tf,
                             long max_acc, long max_att, int fsal, int tab,
                             int n_groups, double* times, long times_cap,
                             long trace_cap) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const Tableau& T = c_tab[tab];
  const int q = T.p < T.pt ? T.p : T.pt;
  st->t = t0;
  st->t0 = t0;
  st->tf = tf;
  st->dt = 0.0;
  st->err = 0.0;
  st->dt_initial = 0.0;
  st->inv_q1 = 1.0 / (1.0 + q);
  st->safety = cr_pow(0.38, (1.0 / (1.0 + q)));
  st->dt0 = 0.0;
  st->d1 = 0.0;
  st->accepted = 0;
  st->rejected = 0;
  st->attempts = 0;
  st->rhs_evals = 0;
  st->max_accepted = max_acc;
  st->max_attempts = max_att;
  st->times_cap = times_cap;
  st->trace_cap = trace_cap;
  st->cur = 0;
  st->fsal = fsal;
  st->tableau = tab;
  st->n_groups = n_groups;
  st->k1_valid = 0;
  st->status = ODEHIP_STATUS_RUNNING;
  st->done = (t0 < tf) ? 0 : 1;
  if (st->done) st->status = ODEHIP_STATUS_REACHED_END;
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
  const double nv = static_cast<double>(static_cast<long>(n_var));
  const double d0 = std::sqrt(sa / nv);
  const double d1 = std::sqrt(sb / nv);
  const double dt0 = (d0 < 1.0e-5 or d1 < 1.0e-5) ? 1.0e-6 : 0.01 * (d0 / d1);
  st->dt0 = dt0;
  st->d1 = d1;
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
  const Tableau& T = c_tab[st->tableau];
  const double nv = static_cast<double>(static_cast<long>(n_var));
  const double dt0 = st->dt0;
  const double d1 = st->d1;
  const double d2 = std::sqrt(sa / nv) / dt0;
  const int p = T.p;
  const double m = (d1 < d2) ? d2 : d1;
  const double hi = 1.0e-6, lo = dt0 * 1.0e-3;
  const double dt1 = (m <= 1.0e-15) ? ((hi < lo) ? lo : hi)
                                    : cr_pow(0.01 / m, (1.0 / (1.0 + p)));
  const double cap = 100.0 * dt0;
  const double dt = (dt1 < cap) ? dt1 : cap;
  st->dt = dt;
  st->dt_initial = dt;
  st->rhs_evals += 2;
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
  hipLaunchKernelGGL(k_init_state, dim3(1), dim3(64), 0, s, b.st, t0, tf,
                     max_acc, max_att, fsal, b.tableau, n_groups_for(b.n_var),
                     b.times, b.times_cap, b.trace_cap);
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
  const Tableau& T = kTableaus[b.tableau];
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
  hipLaunchKernelGGL(k_controller, dim3(1), dim3(64), 0, s, b.st, b.partA, n,
                     b.times, b.trace);
  return hipGetLastError();
}

hipError_t enqueue_solve_finish(const SolveBufs& b, double* x_out,
                                hipStream_t s) {
  hipLaunchKernelGGL(k_copy_out, dim3(blocks_for(b.n_var)), dim3(kBlock), 0, s,
                     b.st, b.xb0, b.xb1, x_out, b.n_var);
  return hipGetLastError();
}

}  // namespace odehip
