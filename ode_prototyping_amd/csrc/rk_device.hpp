// Device-side arithmetic of the embedded Runge-Kutta method shared by the
// single-system kernels (rk_kernels.hip) and the ensemble kernels
// (ensemble_kernels.hip): per-element stage / update formulas, the canonical
// group reduction, and the accept/reject + step-size controller.  Both solve
// paths call these same functions, so an ensemble member is computed with
// exactly the arithmetic of a single-system solve.
#pragma once

#include <cmath>

#include "crpow.hpp"
#include "internal.hpp"

namespace odehip {

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

// detail::rk_stage_kernel (RKEmbeddedParallel.hpp:91-101): element i of
// x0 + dt * sum_{j<stage} a[stage-1][j] k_j, stage vectors n_var apart.
__device__ inline double stage_value(const Tableau& T, int stage, double dt,
                                     const double* ks, const double* x0,
                                     size_t i, size_t n_var) {
  double tmp = 0.0;
  for (int j = 0; j < stage; ++j) tmp += T.a[stage - 1][j] * ks[j * n_var + i];
  return x0[i] + tmp * dt;
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

// ---------------------------------------------------------- controller ----

// Solver scalars before the first attempt (RKEmbeddedParallel.hpp:13-33).
__device__ inline void ctrl_init(SolverState* st, const Tableau& T, double t0,
                                 double tf, long max_acc, long max_att,
                                 int fsal, int tab, int n_groups,
                                 long times_cap, long trace_cap,
                                 int controller) {
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
  st->err_prev = 1.0e-4;
  for (int h = 0; h < 3; ++h) {
    st->hist_err[h] = 0.0;
    st->hist_dt[h] = 0.0;
    st->hist_acc[h] = 0.0;
  }
  st->accepted = 0;
  st->rejected = 0;
  st->attempts = 0;
  st->rhs_evals = 0;
  st->max_accepted = max_acc;
  st->max_attempts = max_att;
  st->times_cap = times_cap;
  st->trace_cap = trace_cap;
  st->learned_steps = 0;
  st->dt_acc = 0.0;
  st->dense_next = 0;
  st->dense_lo = 0;
  st->dense_hi = 0;
  st->dense_seen = 0;
  st->dense_steps = 0;
  st->dense_rhs_evals = 0;
  st->dense_go = 0;
  st->pad2 = 0;
  st->cur = 0;
  st->fsal = fsal;
  st->tableau = tab;
  st->n_groups = n_groups;
  st->k1_valid = 0;
  st->controller = controller;
  st->rejects_in_row = 0;
  st->just_accepted = 0;
  st->status = ODEHIP_STATUS_RUNNING;
  st->done = (t0 < tf) ? 0 : 1;
  if (st->done) st->status = ODEHIP_STATUS_REACHED_END;
}

// First half of an attempt's end (RKEmbeddedParallel.hpp:56-63): the scaled
// error norm from the sum of squares, counters, trace, accept/reject, the
// time update and the x0 = x swap.  Returns false when the solve stops on a
// non-finite error norm (nothing else to do for this attempt).
__device__ inline bool ctrl_judge(SolverState* st, const Tableau& T,
                                  double sum, size_t n_var, double* times,
                                  double* trace, bool* accepted_out) {
  const double scaled_error =
      std::sqrt(sum / static_cast<double>(static_cast<long>(n_var)));
  const long attempt = st->attempts;
  st->attempts = attempt + 1;
  st->rhs_evals += st->k1_valid ? (T.S - 1) : T.S;
  st->err = scaled_error;
  const double dt = st->dt;
  const bool finite = std::isfinite(scaled_error);
  const bool accept = finite && scaled_error <= 1.0;
  *accepted_out = accept;
  st->just_accepted = accept ? 1 : 0;
  if (trace && attempt < st->trace_cap) {
    trace[3 * attempt] = scaled_error;
    trace[3 * attempt + 1] = dt;
    trace[3 * attempt + 2] = accept ? 1.0 : 0.0;
  }
  // attempt history, most recent first (controller features)
  for (int h = 2; h > 0; --h) {
    st->hist_err[h] = st->hist_err[h - 1];
    st->hist_dt[h] = st->hist_dt[h - 1];
    st->hist_acc[h] = st->hist_acc[h - 1];
  }
  st->hist_err[0] = scaled_error;
  st->hist_dt[0] = dt;
  st->hist_acc[0] = accept ? 1.0 : 0.0;
  if (!finite) {
    st->rejected += 1;
    st->status = ODEHIP_STATUS_NONFINITE;
    st->done = 1;
    return false;
  }
  if (accept) {
    st->t = st->t + dt;
    st->dt_acc = dt;
    st->cur ^= 1;  // x0 = x (the ping-pong buffers swap roles)
    const long acc = st->accepted + 1;
    st->accepted = acc;
    if (times && acc < st->times_cap) times[acc] = st->t;
    st->k1_valid = (st->fsal && T.fsal) ? 1 : 0;
    st->rejects_in_row = 0;
  } else {
    st->rejected += 1;
    st->k1_valid = st->fsal ? 2 : 0;
    st->rejects_in_row += 1;
  }
  return true;
}

// Step-size proposal of the PI controller (Gustafsson's PI.3.4; Hairer &
// Wanner, Solving ODEs II, IV.2): dt * (0.38/err)^(0.3/k) *
// (err_prev/err)^(0.4/k), k = 1 + q -- the reference controller's steady
// error 0.38 --, err_prev the error of the last accepted step (1e-4 before
// the first one, floored at 1e-4), no growth right after a rejection.
__device__ inline double pi_proposal(const SolverState* st, double err,
                                     bool accept) {
  const double k_inv = st->inv_q1;
  double fac = cr_pow(0.38 / err, 0.3 * k_inv) * cr_pow(st->err_prev / err, 0.4 * k_inv);
  if (!accept && fac > 1.0) fac = 1.0;
  return st->dt * fac;
}

// Second half (RKEmbeddedParallel.hpp:64-79): the new step size, its clamps
// and the clip to tf, then the termination tests.  `learned_ratio` is the
// learned controller's proposed dt_new / dt (ODEHIP_CONTROLLER_LEARNED only,
// capped at 1 right after a rejection); the PI proposal replaces it when it
// is not finite, before three attempts of history exist, or after two
// rejections in a row.
__device__ inline void ctrl_step(SolverState* st, bool accept,
                                 double learned_ratio) {
  const double scaled_error = st->err;
  double dt = st->dt;
  const double t = st->t;
  double dtnew;
  if (st->controller == ODEHIP_CONTROLLER_PI) {
    dtnew = pi_proposal(st, scaled_error, accept);
  } else if (st->controller == ODEHIP_CONTROLLER_LEARNED) {
    const bool usable = std::isfinite(learned_ratio) && st->attempts >= 3 &&
                        st->rejects_in_row < 2;
    if (usable) {
      // no growth right after a rejection
      dtnew = dt * ((!accept && learned_ratio > 1.0) ? 1.0 : learned_ratio);
      st->learned_steps += 1;
    } else {
      dtnew = pi_proposal(st, scaled_error, accept);
    }
  } else {
    dtnew = dt * st->safety / cr_pow(scaled_error, st->inv_q1);
  }
  if (accept) st->err_prev = scaled_error < 1.0e-4 ? 1.0e-4 : scaled_error;
  const double max_step_scale = 6.0;
  const double min_step_scale = 0.33;
  const double tf = st->tf, t0 = st->t0;
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

// A whole attempt end with the reference or PI controller.
__device__ inline void ctrl_attempt(SolverState* st, const Tableau& T,
                                    double sum, size_t n_var, double* times,
                                    double* trace) {
  bool accept = false;
  if (ctrl_judge(st, T, sum, n_var, times, trace, &accept))
    ctrl_step(st, accept, NAN);
}

// ------------------------------------------------------- dense output ----

// DenseOutput::save's test (NumericalRecipes/Output.hpp:109) after an
// attempt: if it was accepted and ended strictly after the next output
// time(s), rows [dense_lo, dense_hi) are this step's and dense_go is set.
// extra_rhs: RHS evaluations the pass costs (counted; with DOP853 f at the
// step's end also becomes the next k1 when k1 reuse is on).
__device__ inline void dense_gate(SolverState* st, const double* times, long n_dense,
                                  int extra_rhs) {
  int go = 0;
  if (st->just_accepted && st->accepted != st->dense_seen) {
    st->dense_seen = st->accepted;
    const long lo = st->dense_next;
    long hi = lo;
    const double t = st->t;
    while (hi < n_dense && t > times[hi]) ++hi;
    if (hi > lo) {
      go = 1;
      st->dense_lo = lo;
      st->dense_hi = hi;
      st->dense_next = hi;
      st->dense_steps += 1;
      st->dense_rhs_evals += extra_rhs;
      if (extra_rhs && st->fsal) st->k1_valid = 1;
    }
  }
  st->dense_go = go;
}

// DOP853's stage vector j of a dense pass: 0..11 the accepted attempt's
// stages (nv apart in ks), 12 = f at the step's end (k13), 13..15 the extra
// stages, kept in the slots of stages 1..3, which the extension never reads.
__device__ inline const double* dop853_dense_vec(const double* ks, const double* k13, int j,
                                                 size_t nv) {
  if (j < 12) return ks + static_cast<size_t>(j) * nv;
  if (j == 12) return k13;
  return ks + static_cast<size_t>(j - 12) * nv;
}

// Element e of the input of extra stage 13 + x (StepperDopr853.hpp:239-262):
// y + h (w_0 k_{i0} + ... + w_7 k_{i7}), the sum left to right.
__device__ inline double dop853_dense_stage(const Dop853Dense& D, int x, const double* ks,
                                            const double* k13, size_t nv, size_t e,
                                            const double* y, double h) {
  double acc = D.extra_w[x][0] * dop853_dense_vec(ks, k13, D.extra_idx[x][0], nv)[e];
  for (int m = 1; m < 8; ++m)
    acc = acc + D.extra_w[x][m] * dop853_dense_vec(ks, k13, D.extra_idx[x][m], nv)[e];
  return y[e] + h * acc;
}

// StepperDopr853::prepare_dense / dense_out (StepperDopr853.hpp:214-237,
// 263-289) for element e: the eight polynomial coefficients from 14 stage
// vectors, then out[row * row_stride + e] for the pass's rows.
__device__ inline void dop853_dense_rows(const Dop853Dense& D, const SolverState* st,
                                         const double* ks, const double* k13, size_t nv,
                                         size_t e, const double* y, const double* yout,
                                         const double* times, double* out,
                                         size_t row_stride) {
  const double h = st->dt_acc;
  const double yi = y[e];
  const double ydiff = yout[e] - yi;
  const double bspl = h * ks[e] - ydiff;
  const double rc3 = ydiff - h * k13[e] - bspl;
  double rc[4];
  for (int r = 0; r < 4; ++r) {
    double acc = D.d_w[r][0] * dop853_dense_vec(ks, k13, D.d_idx[0], nv)[e];
    for (int m = 1; m < 12; ++m)
      acc = acc + D.d_w[r][m] * dop853_dense_vec(ks, k13, D.d_idx[m], nv)[e];
    rc[r] = h * acc;
  }
  const double t_end = st->t;
  const long hi = st->dense_hi;
  for (long row = st->dense_lo; row < hi; ++row) {
    const double s = (times[row] - (t_end - h)) / h;
    const double s1 = 1.0 - s;
    out[static_cast<size_t>(row) * row_stride + e] =
        yi + s * (ydiff + s1 * (bspl + s * (rc3 + s1 * (rc[0] + s * (rc[1] + s1 * (rc[2] + s * rc[3]))))));
  }
}

// StepperDopr5::prepare_dense / dense_out (StepperDopr5.hpp:195-221) for
// element e of a state whose stage vectors are nv apart in ks (stage 7 = f at
// the new state): the step went from y to yout over st->dt_acc ending at
// st->t; writes out[row * row_stride + e] for the pass's rows.
__device__ inline void dopri5_dense_rows(const Dopri5Dense& D, const SolverState* st,
                                         const double* ks, size_t nv, size_t e,
                                         const double* y, const double* yout,
                                         const double* times, double* out,
                                         size_t row_stride) {
  const double h = st->dt_acc;
  const double yi = y[e];
  const double ydiff = yout[e] - yi;
  const double bspl = h * ks[e] - ydiff;
  const double rc3 = ydiff - h * ks[6 * nv + e] - bspl;
  double acc = D.d[0] * ks[static_cast<size_t>(D.idx[0]) * nv + e];
  for (int m = 1; m < 6; ++m) acc = acc + D.d[m] * ks[static_cast<size_t>(D.idx[m]) * nv + e];
  const double rc4 = h * acc;
  const double t_end = st->t;
  const long hi = st->dense_hi;
  for (long row = st->dense_lo; row < hi; ++row) {
    const double s = (times[row] - (t_end - h)) / h;
    const double s1 = 1.0 - s;
    out[static_cast<size_t>(row) * row_stride + e] =
        yi + s * (ydiff + s1 * (bspl + s * (rc3 + s1 * rc4)));
  }
}

// estimate_initial_step (RKEmbeddedParallel.hpp:183-185): d0, d1, dt0 from
// the sums of squares of x0/sc and f0/sc.
__device__ inline void ctrl_init_step_a(SolverState* st, double sa, double sb,
                                        size_t n_var) {
  const double nv = static_cast<double>(static_cast<long>(n_var));
  const double d0 = std::sqrt(sa / nv);
  const double d1 = std::sqrt(sb / nv);
  const double dt0 = (d0 < 1.0e-5 or d1 < 1.0e-5) ? 1.0e-6 : 0.01 * (d0 / d1);
  st->dt0 = dt0;
  st->d1 = d1;
}

// (:193-199): d2, dt1, the initial step, from the sum of squares of
// (f1 - f0)/sc.
__device__ inline void ctrl_init_step_b(SolverState* st, const Tableau& T,
                                        double sa, size_t n_var) {
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

}  // namespace odehip
