// CPU ORACLE -- test infrastructure, not product code.
//
// A from-scratch C++ restatement of the reference's adaptive embedded
// Runge-Kutta path, used to check the HIP library.  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load or run
// anything under oracle/; the product never does.  Build with
// -ffp-contract=off: the reference's golden values are reproduced bit for bit
// only without FMA contraction.
//
// What each part restates (reference file:line):
//   Executor::parallel_for      src/par_exec/SingleThreadedExecutor.hpp:8-15
//   Executor::transform_reduce  sequential: SingleThreadedExecutor.hpp:17-29
//                               chunked:    ThreadPoolExecutor.hpp:81-114
//   NBody                       src/ode/NBodyODE.hpp:31-39,51-56,68-133 in the
//                               gather order of src/ode/NBodySimpleODE.hpp:38-60
//                               (per particle, partners in increasing index);
//                               optional per-particle masses follow the
//                               commented-out mass code NBodyODE.hpp:104-124
//   Exponential / VanDerPol     src/ode/ExponentialODE.hpp:14-18,
//                               src/ode/VanDerPolODE.hpp:16-25
//   integrate                   src/integrator/RKEmbeddedParallel/RKEmbeddedParallel.hpp:13-81
//   stage / update / target /   RKEmbeddedParallel.hpp:91-149
//     norm kernels
//   estimate_initial_step       RKEmbeddedParallel.hpp:151-200
//   nr_dop853                   NumericalRecipes/ODEIntegrator.hpp:66-88,
//                               NumericalRecipes/StepperDopr853.hpp:93-197,291-342
//                               (dense output :200-289, Output.hpp:63-147)
//   nr_dopr5                    NumericalRecipes/StepperDopr5.hpp:80-270
//   energy                      scripts/nbody_physics/energy.py:118-192,224-241
//                               (softened PE + KE), generalised to masses
//
// Parity pins (see tests/): the integrator, the five reference tableaux and
// the unit-mass N-body RHS are pinned by the reference's golden consistency
// tests (src/test/SingleThreadedRKEmbeddedTest.cpp) and by the reference code
// itself compiled from its own headers into oracle/_ref; DOP853's constants by
// the NumericalRecipes goldens through nr_dop853 and by the order conditions;
// masses by the two-body closed form, momentum conservation and the exact
// reduction to the unit-mass case; energy/momentum by closed forms.
#pragma once

#include <quadmath.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <functional>
#include <limits>
#include <stdexcept>
#include <thread>
#include <vector>

#include "tableaus.hpp"

namespace oracle {

// ---------------------------------------------------------------- executor --

enum class ReduceOrder : int {
  Sequential = 0,  // SingleThreadedExecutor: init, then + t_0 + t_1 + ...
  Chunked = 1,     // fixed chunk size C: each chunk folded from 0.0 in index
                   // order, then init + p_0 + p_1 + ... (the HIP library's
                   // canonical order; with C = ceil(n/T) it is the order of
                   // ThreadPoolExecutor with T threads)
  ThreadPool = 2,  // ThreadPoolExecutor(T): C = ceil(n/T)
  HipGroups = 3,   // the HIP library's order (include/odehip.h): groups of
                   // 256 zero-padded items; each 64-item quad reduced with
                   // halving strides 32..1 (v[i] += v[i+s]); group sum
                   // (q0 + q1) + (q2 + q3); groups folded in order from init
};

struct Executor {
  ReduceOrder order = ReduceOrder::Sequential;
  long param = 0;   // chunk size (Chunked) or thread count (ThreadPool)
  int threads = 1;  // worker threads for elementwise/particle loops; results
                    // do not depend on it (every item is written by one worker)

  template <class F>
  void parallel_for(size_t n, F&& f) const {
    int T = threads;
    if (T <= 1 || n < 2048) {
      for (size_t i = 0; i < n; ++i) f(i);
      return;
    }
    std::vector<std::thread> pool;
    size_t per = (n + T - 1) / T;
    for (int t = 0; t < T; ++t) {
      size_t lo = t * per, hi = std::min(n, lo + per);
      if (lo >= hi) break;
      pool.emplace_back([&, lo, hi] {
        for (size_t i = lo; i < hi; ++i) f(i);
      });
    }
    for (auto& th : pool) th.join();
  }

  template <class F>
  double transform_reduce(double init, size_t n, F&& f) const {
    if (order == ReduceOrder::Sequential) {
      double r = init;
      for (size_t i = 0; i < n; ++i) r = r + f(i);
      return r;
    }
    if (order == ReduceOrder::HipGroups) {
      double r = init;
      for (size_t g = 0; g * 256 < n; ++g) {
        double q[4];
        for (int w = 0; w < 4; ++w) {
          double v[64];
          for (int l = 0; l < 64; ++l) {
            size_t i = g * 256 + w * 64 + l;
            v[l] = i < n ? f(i) : 0.0;
          }
          for (int st = 32; st >= 1; st >>= 1)
            for (int l = 0; l < st; ++l) v[l] = v[l] + v[l + st];
          q[w] = v[0];
        }
        r = r + ((q[0] + q[1]) + (q[2] + q[3]));
      }
      return r;
    }
    size_t chunk;
    size_t n_chunks;
    if (order == ReduceOrder::ThreadPool) {
      size_t T = static_cast<size_t>(param);
      chunk = (n + T - 1) / T;
      n_chunks = T;  // threads past the end contribute a 0.0 partial
    } else {
      chunk = static_cast<size_t>(param);
      n_chunks = (n + chunk - 1) / chunk;
    }
    double r = init;
    for (size_t c = 0; c < n_chunks; ++c) {
      double part = 0.0;
      for (size_t i = c * chunk; i < (c + 1) * chunk && i < n; ++i)
        part = part + f(i);
      r = r + part;
    }
    return r;
  }
};

// --------------------------------------------------------------------- pow --

enum class PowMode : int {
  Libm = 0,     // std::pow, as the reference (glibc)
  Correct = 1,  // correctly rounded via binary128 (powq), platform independent
};

inline double opow(double x, double y, PowMode m) {
  if (m == PowMode::Libm) return std::pow(x, y);
  return static_cast<double>(powq(static_cast<__float128>(x),
                                  static_cast<__float128>(y)));
}

// -------------------------------------------------------------------- ODEs --

using Rhs = std::function<void(const double* x, double* dxdt, const Executor&)>;

// N-body gravity, G = 1.  State x = [pos xyz-interleaved (3N) | vel (3N)],
// dxdt = [vel | acc].  Per particle i the partners j != i are visited in
// increasing j: this is the order in which the reference's single-threaded
// pair loop adds into acc[i] (pairs (m,i), m<i, come before pairs (i,m)), and
// -( (x_i - x_m)/den ) == (x_m - x_i)/den exactly, so the gather form is bit
// identical to NBodyODE with SingleThreadedExecutor.
struct NBody {
  size_t n = 0;
  double softening = 0.0;
  const double* masses = nullptr;  // nullptr: unit masses (the reference)

  void operator()(const double* x, double* dxdt, const Executor& exe) const {
    const size_t n_dof = 3 * n;
    // nbody_init_dxdt_kernel (NBodyODE.hpp:51-56)
    for (size_t i = 0; i < n_dof; ++i) {
      dxdt[i] = x[i + n_dof];
      dxdt[i + n_dof] = 0.0;
    }
    const double softening_sq = softening * softening;  // NBodyODE.hpp:38
    const double* pos = x;
    double* acc = dxdt + n_dof;
    const double* m = masses;
    const size_t np = n;
    exe.parallel_for(np, [=](size_t i) {
      double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
      double ax_i = acc[3 * i], ay_i = acc[3 * i + 1], az_i = acc[3 * i + 2];
      for (size_t j = 0; j < np; ++j) {
        if (j == i) continue;
        double dx = pos[3 * j] - xi;
        double dy = pos[3 * j + 1] - yi;
        double dz = pos[3 * j + 2] - zi;
        double dist_sq = dx * dx + dy * dy + dz * dz + softening_sq;
        double dist = std::sqrt(dist_sq);
        double den = dist * dist_sq;
        double ax = dx / den, ay = dy / den, az = dz / den;
        if (m) {
          ax *= m[j];
          ay *= m[j];
          az *= m[j];
        }
        ax_i += ax;
        ay_i += ay;
        az_i += az;
      }
      acc[3 * i] = ax_i;
      acc[3 * i + 1] = ay_i;
      acc[3 * i + 2] = az_i;
    });
  }
};

// ExponentialODE.hpp:14-18
struct Exponential {
  size_t n = 0;
  void operator()(const double* x, double* dxdt, const Executor&) const {
    for (size_t i = 0; i < n; ++i) dxdt[i] = x[i];
  }
};

// VanDerPolODE.hpp:16-25 (eps = 1)
struct VanDerPol {
  void operator()(const double* x, double* dxdt, const Executor&) const {
    dxdt[0] = x[1];
    constexpr double eps = 1.0;
    dxdt[1] = eps * (1.0 - x[0] * x[0]) * x[1] - x[0];
  }
};

// ----------------------------------------------------------------- results --

enum Status : int {
  kReachedEnd = 0,     // t >= tf (the reference's only exit)
  kMaxAccepted = 1,    // accepted-step budget reached (build extension)
  kNonFinite = 2,      // error norm not finite (the reference loops forever)
  kMaxAttempts = 3,    // attempt budget exhausted
};

// Step-size controllers: the reference's (RKEmbeddedParallel.hpp:64-74) and
// a PI controller (Gustafsson's PI.3.4; Hairer & Wanner, Solving ODEs II,
// IV.2): dt * (0.38/err)^(0.3/k) * (err_prev/err)^(0.4/k), k = 1 + min(p, pt)
// -- the same steady-state error 0.38 as the reference's controller --,
// err_prev the error of the last accepted step (1e-4 at first, floored at
// 1e-4), no growth right after a rejection; same clamps, floor and clip.
enum class Controller { Reference = 0, PI = 1 };

struct Options {
  Controller controller = Controller::Reference;
  long max_accepted = 0;  // 0: unlimited
  long max_attempts = 0;  // 0: unlimited
  PowMode pow_mode = PowMode::Libm;
  bool store_states = true;  // keep every accepted state (RawOutput)
  bool store_trace = false;  // per-attempt (err, dt, accepted)
  // DOP853 only: states at these times (increasing, after t0) from each
  // accepted step's continuous extension, the reference's NR dense output
  // (StepperDopr853.hpp:200-289, Output.hpp:99-114): a time is written by the
  // first accepted step that ends strictly after it.
  std::vector<double> dense_times;
};

struct Result {
  std::vector<double> times;   // RawOutput::times()
  std::vector<double> states;  // RawOutput::states(), flattened (if stored)
  std::vector<double> x_final;
  std::vector<double> trace_err, trace_dt;
  std::vector<int> trace_acc;
  long accepted = 0, rejected = 0, rhs_evals = 0;
  std::vector<double> dense_states;  // n_dense x n_var
  long n_dense = 0, dense_steps = 0, dense_rhs_evals = 0;
  int status = kReachedEnd;
  double t = 0.0, dt = 0.0, dt_initial = 0.0;
  size_t n_var = 0;
};

// -------------------------------------------------------------- integrator --

namespace detail {

inline double rk_norm(const Executor& exe, const double* v, const double* sc,
                      size_t n_var) {
  // rk_norm + scaled_value_squared_kernel (RKEmbeddedParallel.hpp:132-149)
  double s = exe.transform_reduce(0.0, n_var, [&](size_t i) {
    double scaled = v[i] / sc[i];
    return scaled * scaled;
  });
  return std::sqrt(s / static_cast<double>(static_cast<std::ptrdiff_t>(n_var)));
}

// RKEmbeddedParallel.hpp:170-200
inline double estimate_initial_step(const Tableau& bt, const Executor& exe,
                                    const double* x0, const double* atol,
                                    const double* rtol, size_t n_var,
                                    const Rhs& ode, PowMode pm, long& rhs_evals) {
  std::vector<double> error_target(n_var), f0(n_var), x1(n_var), df(n_var);
  for (size_t i = 0; i < n_var; ++i)
    error_target[i] = std::abs(x0[i]) * rtol[i] + atol[i];
  ode(x0, f0.data(), exe);
  ++rhs_evals;
  double d0 = rk_norm(exe, x0, error_target.data(), n_var);
  double d1 = rk_norm(exe, f0.data(), error_target.data(), n_var);
  double dt0 = (d0 < 1.0e-5 or d1 < 1.0e-5) ? 1.0e-6 : 0.01 * (d0 / d1);
  for (size_t i = 0; i < n_var; ++i) x1[i] = x0[i] + f0[i] * dt0;
  ode(x1.data(), df.data(), exe);
  ++rhs_evals;
  for (size_t i = 0; i < n_var; ++i) df[i] -= f0[i];
  double d2 = rk_norm(exe, df.data(), error_target.data(), n_var) / dt0;
  const int p = bt.p;
  double dt1 = (std::max(d1, d2) <= 1.0e-15)
                   ? std::max(1.0e-6, dt0 * 1.0e-3)
                   : opow(0.01 / std::max(d1, d2), (1.0 / (1.0 + p)), pm);
  return std::min(100.0 * dt0, dt1);
}

// prepare_dense (StepperDopr853.hpp:200-276) for one accepted step of length
// h from y to yout: the eight polynomial coefficients rc (8 x n).  k[0..12]
// are the step's stage vectors (k[1..4] are not read; k[12] = f(yout)); the
// three extra stages are evaluated by f(stage, ytemp, out) into `extra`
// (3 x n).  Every sum runs left to right from its first term, as the
// reference writes them.
template <class F>
void dop853_prepare_dense(const Dop853Dense& D, size_t n, double h, const double* y,
                          const double* yout, const double* const* k, F&& f,
                          std::vector<double>& rc, std::vector<double>& extra) {
  rc.assign(8 * n, 0.0);
  extra.assign(3 * n, 0.0);
  std::vector<double> ytemp(n);
  const double* kv[16];
  for (int j = 0; j < 13; ++j) kv[j] = k[j];
  for (int e = 0; e < 3; ++e) {
    for (size_t i = 0; i < n; ++i) {
      double s = D.extra_w[e][0] * kv[D.extra_idx[e][0]][i];
      for (int m = 1; m < 8; ++m) s = s + D.extra_w[e][m] * kv[D.extra_idx[e][m]][i];
      ytemp[i] = y[i] + h * s;
    }
    f(e, ytemp.data(), extra.data() + e * n);
    kv[13 + e] = extra.data() + e * n;
  }
  for (size_t i = 0; i < n; ++i) {
    const double ydiff = yout[i] - y[i];
    const double bspl = h * kv[0][i] - ydiff;
    rc[i] = y[i];
    rc[n + i] = ydiff;
    rc[2 * n + i] = bspl;
    rc[3 * n + i] = ydiff - h * kv[12][i] - bspl;
    for (int r = 0; r < 4; ++r) {
      double s = D.d_w[r][0] * kv[D.d_idx[0]][i];
      for (int m = 1; m < 12; ++m) s = s + D.d_w[r][m] * kv[D.d_idx[m]][i];
      rc[(4 + r) * n + i] = h * s;
    }
  }
}

// dense_out (StepperDopr853.hpp:278-289): component i at xout inside the step
// that ends at x_end and has length h.
inline double dop853_dense_eval(const std::vector<double>& rc, size_t n, size_t i,
                                double xout, double x_end, double h) {
  const double s = (xout - (x_end - h)) / h;
  const double s1 = 1.0 - s;
  const double* r = rc.data();
  return r[i] + s * (r[n + i] +
                     s1 * (r[2 * n + i] +
                           s * (r[3 * n + i] +
                                s1 * (r[4 * n + i] +
                                      s * (r[5 * n + i] +
                                           s1 * (r[6 * n + i] + s * r[7 * n + i]))))));
}

// StepperDopr5::prepare_dense (StepperDopr5.hpp:181-208) for one accepted
// step of length h from y to yout: five coefficients rc (5 x n) from the
// step's stages k[0..6] (k[1] is not read; k[6] = f(yout)).
inline void dopri5_prepare_dense(const Dopri5Dense& D, size_t n, double h, const double* y,
                                 const double* yout, const double* const* k,
                                 std::vector<double>& rc) {
  rc.assign(5 * n, 0.0);
  for (size_t i = 0; i < n; ++i) {
    const double ydiff = yout[i] - y[i];
    const double bspl = h * k[0][i] - ydiff;
    rc[i] = y[i];
    rc[n + i] = ydiff;
    rc[2 * n + i] = bspl;
    rc[3 * n + i] = ydiff - h * k[6][i] - bspl;
    double s = D.d[0] * k[D.idx[0]][i];
    for (int m = 1; m < 6; ++m) s = s + D.d[m] * k[D.idx[m]][i];
    rc[4 * n + i] = h * s;
  }
}

// StepperDopr5::dense_out (:210-221).
inline double dopri5_dense_eval(const std::vector<double>& rc, size_t n, size_t i,
                                double xout, double x_end, double h) {
  const double s = (xout - (x_end - h)) / h;
  const double s1 = 1.0 - s;
  const double* r = rc.data();
  return r[i] + s * (r[n + i] + s1 * (r[2 * n + i] + s * (r[3 * n + i] + s1 * r[4 * n + i])));
}

}  // namespace detail

// RKEmbeddedParallel<BT>::integrate (RKEmbeddedParallel.hpp:13-81), with
// runtime sizes, counters, optional accepted-step / attempt budgets and a
// non-finite-norm exit (the reference has none of these; with the budgets at
// 0 and a finite norm the step sequence is the reference's).
inline Result integrate(const Tableau& bt, const double* x0_in, size_t n_var,
                        double t0, double tf, const double* atol,
                        const double* rtol, const Rhs& ode, const Executor& exe,
                        const Options& opt) {
  const int S = bt.S;
  const double max_step_scale = 6.0;
  const double min_step_scale = 0.33;
  const int q = std::min(bt.p, bt.pt);
  const double safety_factor = opow(0.38, (1.0 / (1.0 + q)), opt.pow_mode);
  std::vector<double> db(S);
  for (int j = 0; j < S; ++j) db[j] = bt.b[j] - bt.bt[j];

  Result res;
  res.n_var = n_var;
  std::vector<double> ks(static_cast<size_t>(S) * n_var);
  std::vector<double> x0(x0_in, x0_in + n_var);
  double dt = detail::estimate_initial_step(bt, exe, x0.data(), atol, rtol,
                                            n_var, ode, opt.pow_mode,
                                            res.rhs_evals);
  res.dt_initial = dt;
  double t = t0;
  std::vector<double> x = x0, temp_state(n_var), error_estimate(n_var),
                      error_target(n_var);
  auto save = [&](double tt, const std::vector<double>& s) {
    res.times.push_back(tt);
    if (opt.store_states) res.states.insert(res.states.end(), s.begin(), s.end());
  };
  save(t, x);
  const bool dense = !opt.dense_times.empty();
  const bool dense5 = bt.name == "DOPRI5";
  if (dense && bt.name != "DOP853" && !dense5)
    throw std::invalid_argument("dense output needs DOP853 or DOPRI5");
  const Dop853Dense DD = dop853_dense();
  const Dopri5Dense D5 = dopri5_dense();
  size_t dense_next = 0;
  std::vector<double> k13(dense ? n_var : 0), rc, extra;
  long attempts = 0;
  double err_prev = 1.0e-4;
  const double inv_q1 = 1.0 / (1.0 + q);
  while (t < tf) {
    if (opt.max_attempts > 0 && attempts >= opt.max_attempts) {
      res.status = kMaxAttempts;
      break;
    }
    ++attempts;
    // stages (RKEmbeddedParallel.hpp:37-44)
    ode(x0.data(), ks.data(), exe);
    ++res.rhs_evals;
    for (int stage = 1; stage < S; ++stage) {
      exe.parallel_for(n_var, [&](size_t i) {
        double tmp = 0.0;
        for (int j = 0; j < stage; ++j) tmp += bt.A(stage - 1, j) * ks[j * n_var + i];
        temp_state[i] = x0[i] + tmp * dt;
      });
      ode(temp_state.data(), ks.data() + stage * n_var, exe);
      ++res.rhs_evals;
    }
    // update_state_and_error_kernel (:103-123)
    exe.parallel_for(n_var, [&](size_t i) {
      double xs = 0.0, es = 0.0;
      for (int j = 0; j < S; ++j) {
        xs += bt.b[j] * ks[j * n_var + i];
        es += db[j] * ks[j * n_var + i];
      }
      x[i] = x0[i] + xs * dt;
      error_estimate[i] = es * dt;
    });
    // update_error_target_kernel (:125-130)
    exe.parallel_for(n_var, [&](size_t i) {
      error_target[i] = atol[i] + rtol[i] * std::max(std::abs(x[i]), std::abs(x0[i]));
    });
    double scaled_error =
        detail::rk_norm(exe, error_estimate.data(), error_target.data(), n_var);
    if (!std::isfinite(scaled_error)) {
      res.status = kNonFinite;
      if (opt.store_trace) {
        res.trace_err.push_back(scaled_error);
        res.trace_dt.push_back(dt);
        res.trace_acc.push_back(0);
      }
      ++res.rejected;
      break;
    }
    bool accept = scaled_error <= 1.0;
    if (opt.store_trace) {
      res.trace_err.push_back(scaled_error);
      res.trace_dt.push_back(dt);
      res.trace_acc.push_back(accept ? 1 : 0);
    }
    if (accept) {
      t += dt;
      if (dense && dense_next < opt.dense_times.size() && t > opt.dense_times[dense_next]) {
        // the reference's prepare_dense / dense_out on this step
        if (dense5) {
          // DOPRI5's last stage is f at the new state (its stage input
          // equals x bit for bit), so no evaluation is added
          const double* kp[7];
          for (int j = 0; j < 7; ++j) kp[j] = ks.data() + j * n_var;
          detail::dopri5_prepare_dense(D5, n_var, dt, x0.data(), x.data(), kp, rc);
        } else {
          ode(x.data(), k13.data(), exe);
          const double* kp[13];
          for (int j = 0; j < 12; ++j) kp[j] = ks.data() + j * n_var;
          kp[12] = k13.data();
          detail::dop853_prepare_dense(
              DD, n_var, dt, x0.data(), x.data(), kp,
              [&](int, const double* yt, double* out) { ode(yt, out, exe); }, rc, extra);
          res.dense_rhs_evals += 4;
        }
        ++res.dense_steps;
        while (dense_next < opt.dense_times.size() && t > opt.dense_times[dense_next]) {
          for (size_t i = 0; i < n_var; ++i)
            res.dense_states.push_back(
                dense5 ? detail::dopri5_dense_eval(rc, n_var, i, opt.dense_times[dense_next], t, dt)
                       : detail::dop853_dense_eval(rc, n_var, i, opt.dense_times[dense_next], t, dt));
          ++dense_next;
          ++res.n_dense;
        }
      }
      x0 = x;
      save(t, x);
      ++res.accepted;
    } else {
      ++res.rejected;
    }
    // controller (:65-74)
    double dtnew;
    if (opt.controller == Controller::PI) {
      double fac = opow(0.38 / scaled_error, 0.3 * inv_q1, opt.pow_mode) *
                   opow(err_prev / scaled_error, 0.4 * inv_q1, opt.pow_mode);
      if (!accept && fac > 1.0) fac = 1.0;
      dtnew = dt * fac;
    } else {
      dtnew = dt * safety_factor / opow(scaled_error, 1.0 / (1.0 + q), opt.pow_mode);
    }
    if (accept) err_prev = scaled_error < 1.0e-4 ? 1.0e-4 : scaled_error;
    if (std::abs(dtnew) > max_step_scale * std::abs(dt)) {
      dt *= max_step_scale;
    } else if (std::abs(dtnew) < min_step_scale * std::abs(dt)) {
      dt *= min_step_scale;
    } else if (dtnew / (tf - t0) < 1.0e-12) {
      dt = (tf - t0) * 1.0e-12;
    } else {
      dt = dtnew;
    }
    // last-step clip (:77-79)
    if (t + dt > tf) dt = tf - t;
    if (accept && opt.max_accepted > 0 && res.accepted >= opt.max_accepted) {
      res.status = t < tf ? kMaxAccepted : kReachedEnd;
      break;
    }
  }
  res.t = t;
  res.dt = dt;
  res.x_final = x0;
  return res;
}

// ---------------------------------------------------------------- leapfrog --
// ParallelLeapfrogIntegrator::integrate (LeapfrogParallel.hpp:8-41, kernels
// :57-73): velocity Verlet on a [pos | vel] state with the N-body
// accelerations (NBodyODE::acc, NBodyODE.hpp:41-46 -- the RHS's second half);
// fixed dt, no clip of the last step, every state saved.
inline Result leapfrog(const double* x0_in, size_t n_var, double t0, double tf,
                       double dt, const Rhs& ode, const Executor& exe,
                       bool store_states) {
  Result res;
  res.n_var = n_var;
  const size_t n_dof = n_var / 2;
  std::vector<double> x(x0_in, x0_in + n_var), d0(n_var), d1(n_var);
  auto save = [&](double tt) {
    res.times.push_back(tt);
    if (store_states) res.states.insert(res.states.end(), x.begin(), x.end());
  };
  double t = t0;
  save(t);
  ode(x.data(), d0.data(), exe);  // acc0 = d0[n_dof:]
  ++res.rhs_evals;
  while (t < tf) {
    for (size_t i = 0; i < n_dof; ++i)
      x[i] += dt * (x[n_dof + i] + 0.5 * dt * d0[n_dof + i]);
    ode(x.data(), d1.data(), exe);
    ++res.rhs_evals;
    for (size_t i = 0; i < n_dof; ++i)
      x[n_dof + i] += 0.5 * dt * (d0[n_dof + i] + d1[n_dof + i]);
    std::swap(d0, d1);  // acc0 = acc1
    t += dt;
    ++res.accepted;
    save(t);
  }
  res.t = t;
  res.dt = dt;
  res.x_final = x;
  return res;
}

// ------------------------------------------------- NR-compatible DOP853 ----
// ODEIntegrator<StepperDopr853<D, RawOutput>> (NumericalRecipes/
// ODEIntegrator.hpp:66-88, StepperDopr853.hpp:93-197, 291-342): the NR
// stepper's own stage expressions, two-estimator error norm and controller.
// Used only to pin the DOP853 constants against the NR goldens.
struct NrResult {
  std::vector<double> xs;                 // saved x values
  std::vector<std::vector<double>> ys;    // saved y values
  std::vector<double> y_final;            // ystart after a normal exit
  int status = 0;                         // 0 ok, 1 too many steps, 2 underflow
};

using NrDerivs = std::function<void(double, const std::vector<double>&,
                                    std::vector<double>&)>;

// dense_intervals > 0 selects the DenseOutput policy (Output.hpp:63-147):
// the first save records (x1, ystart); every accepted step then prepares its
// continuous extension and records the points x1 + k (x2 - x1)/intervals
// (accumulated by repeated addition) that lie strictly before its end.
inline NrResult nr_dop853(std::vector<double> ystart, double x1, double x2,
                          double atol, double rtol, double h1, double hmin,
                          const NrDerivs& derivs, int dense_intervals = 0) {
  const Tableau T = dop853();
  const Dop853Extras E = dop853_extras();
  auto a = [&](int i, int j) { return T.A(i - 2, j - 1); };  // a_ij, 1-based
  const int n = static_cast<int>(ystart.size());
  const double eps = std::numeric_limits<double>::epsilon();
  std::vector<double> y = ystart, dydx(n), yout = ystart, yerr(n), yerr2(n);
  std::vector<double> k2(n), k3(n), k4(n), k5(n), k6(n), k7(n), k8(n), k9(n),
      k10(n), dydxnew(n), ytemp(n);
  double x = x1;
  double hnext = 0.0, errold = 1.0e-4;
  bool reject = false;
  NrResult out;
  const Dop853Dense DD = dop853_dense();
  const double width = dense_intervals > 0 ? (x2 - x1) / dense_intervals : 0.0;
  double next_x = x1 + width;
  double hdid = 0.0;
  bool first_save = true;
  std::vector<double> rc, extra;
  auto save = [&] {
    if (dense_intervals <= 0 || first_save) {
      first_save = false;
      out.xs.push_back(x);
      out.ys.push_back(yout);
      return;
    }
    const double* kp[13] = {dydx.data(), nullptr, nullptr, nullptr, nullptr,
                            k6.data(), k7.data(), k8.data(), k9.data(), k10.data(),
                            k2.data(), k3.data(), dydxnew.data()};
    detail::dop853_prepare_dense(
        DD, static_cast<size_t>(n), hdid, y.data(), yout.data(), kp,
        [&](int e, const double* yt, double* o) {
          std::vector<double> yv(yt, yt + n), ov(n);
          derivs(x + DD.c[e] * hdid, yv, ov);
          std::copy(ov.begin(), ov.end(), o);
        },
        rc, extra);
    while ((x - next_x) * (x2 - x1) > 0.0) {
      std::vector<double> yd(n);
      for (int i = 0; i < n; ++i)
        yd[i] = detail::dop853_dense_eval(rc, static_cast<size_t>(n), i, next_x, x, hdid);
      out.xs.push_back(next_x);
      out.ys.push_back(yd);
      next_x += width;
    }
  };
  auto dy = [&](double h) {
    int i;
    for (i = 0; i < n; ++i) ytemp[i] = y[i] + h * a(2, 1) * dydx[i];
    derivs(x + T.c[1] * h, ytemp, k2);
    for (i = 0; i < n; ++i) ytemp[i] = y[i] + h * (a(3, 1) * dydx[i] + a(3, 2) * k2[i]);
    derivs(x + T.c[2] * h, ytemp, k3);
    for (i = 0; i < n; ++i) ytemp[i] = y[i] + h * (a(4, 1) * dydx[i] + a(4, 3) * k3[i]);
    derivs(x + T.c[3] * h, ytemp, k4);
    for (i = 0; i < n; ++i)
      ytemp[i] = y[i] + h * (a(5, 1) * dydx[i] + a(5, 3) * k3[i] + a(5, 4) * k4[i]);
    derivs(x + T.c[4] * h, ytemp, k5);
    for (i = 0; i < n; ++i)
      ytemp[i] = y[i] + h * (a(6, 1) * dydx[i] + a(6, 4) * k4[i] + a(6, 5) * k5[i]);
    derivs(x + T.c[5] * h, ytemp, k6);
    for (i = 0; i < n; ++i)
      ytemp[i] = y[i] + h * (a(7, 1) * dydx[i] + a(7, 4) * k4[i] + a(7, 5) * k5[i] +
                             a(7, 6) * k6[i]);
    derivs(x + T.c[6] * h, ytemp, k7);
    for (i = 0; i < n; ++i)
      ytemp[i] = y[i] + h * (a(8, 1) * dydx[i] + a(8, 4) * k4[i] + a(8, 5) * k5[i] +
                             a(8, 6) * k6[i] + a(8, 7) * k7[i]);
    derivs(x + T.c[7] * h, ytemp, k8);
    for (i = 0; i < n; ++i)
      ytemp[i] = y[i] + h * (a(9, 1) * dydx[i] + a(9, 4) * k4[i] + a(9, 5) * k5[i] +
                             a(9, 6) * k6[i] + a(9, 7) * k7[i] + a(9, 8) * k8[i]);
    derivs(x + T.c[8] * h, ytemp, k9);
    for (i = 0; i < n; ++i)
      ytemp[i] = y[i] + h * (a(10, 1) * dydx[i] + a(10, 4) * k4[i] +
                             a(10, 5) * k5[i] + a(10, 6) * k6[i] +
                             a(10, 7) * k7[i] + a(10, 8) * k8[i] + a(10, 9) * k9[i]);
    derivs(x + T.c[9] * h, ytemp, k10);
    for (i = 0; i < n; ++i)
      ytemp[i] = y[i] + h * (a(11, 1) * dydx[i] + a(11, 4) * k4[i] +
                             a(11, 5) * k5[i] + a(11, 6) * k6[i] +
                             a(11, 7) * k7[i] + a(11, 8) * k8[i] +
                             a(11, 9) * k9[i] + a(11, 10) * k10[i]);
    derivs(x + T.c[10] * h, ytemp, k2);
    double xph = x + h;
    for (i = 0; i < n; ++i)
      ytemp[i] = y[i] + h * (a(12, 1) * dydx[i] + a(12, 4) * k4[i] +
                             a(12, 5) * k5[i] + a(12, 6) * k6[i] +
                             a(12, 7) * k7[i] + a(12, 8) * k8[i] +
                             a(12, 9) * k9[i] + a(12, 10) * k10[i] +
                             a(12, 11) * k2[i]);
    derivs(xph, ytemp, k3);
    const auto& b = T.b;
    for (i = 0; i < n; ++i) {
      k4[i] = b[0] * dydx[i] + b[5] * k6[i] + b[6] * k7[i] + b[7] * k8[i] +
              b[8] * k9[i] + b[9] * k10[i] + b[10] * k2[i] + b[11] * k3[i];
      yout[i] = y[i] + h * k4[i];
    }
    for (i = 0; i < n; ++i) {
      yerr[i] = k4[i] - E.bhh1 * dydx[i] - E.bhh2 * k9[i] - E.bhh3 * k3[i];
      yerr2[i] = E.er[0] * dydx[i] + E.er[5] * k6[i] + E.er[6] * k7[i] +
                 E.er[7] * k8[i] + E.er[8] * k9[i] + E.er[9] * k10[i] +
                 E.er[10] * k2[i] + E.er[11] * k3[i];
    }
  };
  auto error = [&](double h) {
    double err = 0.0, err2 = 0.0;
    for (int i = 0; i < n; ++i) {
      double sk = atol + rtol * std::max(std::abs(y[i]), std::abs(yout[i]));
      double e1 = yerr[i] / sk, e2 = yerr2[i] / sk;
      err2 += e1 * e1;
      err += e2 * e2;
    }
    double deno = err + 0.01 * err2;
    if (deno <= 0.0) deno = 1.0;
    return std::abs(h) * err * std::sqrt(1.0 / (n * deno));
  };
  auto success = [&](double err, double& h) {
    const double beta = 0.0, alpha = 1.0 / 8.0 - beta * 0.2, safe = 0.9,
                 minscale = 0.333, maxscale = 6.0;
    double scale;
    if (err <= 1.0) {
      if (err == 0.0) {
        scale = maxscale;
      } else {
        scale = safe * std::pow(err, -alpha) * std::pow(errold, beta);
        if (scale < minscale) scale = minscale;
        if (scale > maxscale) scale = maxscale;
      }
      hnext = reject ? h * std::min(scale, 1.0) : h * scale;
      errold = std::max(err, 1.0e-4);
      reject = false;
      return true;
    }
    scale = std::max(safe * std::pow(err, -alpha), minscale);
    h *= scale;
    reject = true;
    return false;
  };
  double h = x2 - x1 > 0.0 ? std::fabs(h1) : -std::fabs(h1);
  derivs(x, y, dydx);
  save();
  const int max_step = 50000;
  for (int nstp = 0; nstp < max_step; ++nstp) {
    if ((x + h * 1.0001 - x2) * (x2 - x1) > 0.0) h = x2 - x;
    // step(h)
    double hh = h;
    for (;;) {
      dy(hh);
      double err = error(hh);
      if (success(err, hh)) break;
      if (std::abs(hh) <= std::abs(x) * eps) {
        out.status = 2;
        return out;
      }
    }
    derivs(x + hh, yout, dydxnew);
    x += hh;
    hdid = hh;
    save();
    dydx = dydxnew;
    y = yout;
    h = hnext;
    if ((x - x2) * (x2 - x1) >= 0.0) {
      out.y_final = y;  // ystart = y (ODEIntegrator.hpp:80)
      return out;
    }
    if (std::abs(h) <= hmin) {
      out.status = 2;
      return out;
    }
  }
  out.status = 1;
  return out;
}

// ------------------------------------------------- NR-compatible DOPRI5 ----
// ODEIntegrator<StepperDopr5<D, OP>> (NumericalRecipes/ODEIntegrator.hpp:66-88;
// StepperDopr5.hpp:80-98 step, :109-177 dy, :224-231 error, :236-270
// controller, :181-221 dense output), with the RawOutput or DenseOutput
// policy (dense_intervals > 0, as nr_dop853).  Used to pin DOPRI5's
// continuous extension against the NR goldens.
inline NrResult nr_dopr5(std::vector<double> ystart, double x1, double x2, double atol,
                         double rtol, double h1, double hmin, const NrDerivs& derivs,
                         int dense_intervals = 0) {
  const Tableau T = dopri5();
  const std::vector<double> e = dopri5_nr_error_weights();
  const Dopri5Dense D5 = dopri5_dense();
  auto a = [&](int i, int j) { return T.A(i - 2, j - 1); };  // a_ij, 1-based
  const double c[5] = {0.0, 0.2, 0.3, 0.8, 8.0 / 9.0};  // nodes c_1..c_5
  const int n = static_cast<int>(ystart.size());
  const double eps = std::numeric_limits<double>::epsilon();
  std::vector<double> y = ystart, dydx(n), yout = ystart, yerr(n);
  std::vector<double> k2(n), k3(n), k4(n), k5(n), k6(n), dydxnew(n), ytemp(n);
  double x = x1, hnext = 0.0, errold = 1.0e-4, hdid = 0.0;
  bool reject = false;
  NrResult out;
  const double width = dense_intervals > 0 ? (x2 - x1) / dense_intervals : 0.0;
  double next_x = x1 + width;
  bool first_save = true;
  std::vector<double> rc;
  auto save = [&] {
    if (dense_intervals <= 0 || first_save) {
      first_save = false;
      out.xs.push_back(x);
      out.ys.push_back(yout);
      return;
    }
    const double* kp[7] = {dydx.data(), nullptr, k3.data(), k4.data(),
                           k5.data(), k6.data(), dydxnew.data()};
    detail::dopri5_prepare_dense(D5, static_cast<size_t>(n), hdid, y.data(), yout.data(), kp,
                                 rc);
    while ((x - next_x) * (x2 - x1) > 0.0) {
      std::vector<double> yd(n);
      for (int i = 0; i < n; ++i)
        yd[i] = detail::dopri5_dense_eval(rc, static_cast<size_t>(n), i, next_x, x, hdid);
      out.xs.push_back(next_x);
      out.ys.push_back(yd);
      next_x += width;
    }
  };
  auto dy = [&](double h) {
    int i;
    for (i = 0; i < n; ++i) ytemp[i] = y[i] + h * a(2, 1) * dydx[i];
    derivs(x + c[1] * h, ytemp, k2);
    for (i = 0; i < n; ++i) ytemp[i] = y[i] + h * (a(3, 1) * dydx[i] + a(3, 2) * k2[i]);
    derivs(x + c[2] * h, ytemp, k3);
    for (i = 0; i < n; ++i)
      ytemp[i] = y[i] + h * (a(4, 1) * dydx[i] + a(4, 2) * k2[i] + a(4, 3) * k3[i]);
    derivs(x + c[3] * h, ytemp, k4);
    for (i = 0; i < n; ++i)
      ytemp[i] = y[i] + h * (a(5, 1) * dydx[i] + a(5, 2) * k2[i] + a(5, 3) * k3[i] +
                             a(5, 4) * k4[i]);
    derivs(x + c[4] * h, ytemp, k5);
    for (i = 0; i < n; ++i)
      ytemp[i] = y[i] + h * (a(6, 1) * dydx[i] + a(6, 2) * k2[i] + a(6, 3) * k3[i] +
                             a(6, 4) * k4[i] + a(6, 5) * k5[i]);
    const double xph = x + h;
    derivs(xph, ytemp, k6);
    for (i = 0; i < n; ++i)
      yout[i] = y[i] + h * (a(7, 1) * dydx[i] + a(7, 3) * k3[i] + a(7, 4) * k4[i] +
                            a(7, 5) * k5[i] + a(7, 6) * k6[i]);
    derivs(xph, yout, dydxnew);
    for (i = 0; i < n; ++i)
      yerr[i] = h * (e[0] * dydx[i] + e[2] * k3[i] + e[3] * k4[i] + e[4] * k5[i] +
                     e[5] * k6[i] + e[6] * dydxnew[i]);
  };
  auto error = [&] {
    double err = 0.0;
    for (int i = 0; i < n; ++i) {
      const double sk = atol + rtol * std::max(std::fabs(y[i]), std::fabs(yout[i]));
      err += std::pow(yerr[i] / sk, 2);
    }
    return std::sqrt(err / n);
  };
  auto success = [&](double err, double& h) {
    const double beta = 0.0, alpha = 0.2 - beta * 0.75, safe = 0.9, minscale = 0.2,
                 maxscale = 10.0;
    double scale;
    if (err <= 1.0) {
      if (err == 0.0) {
        scale = maxscale;
      } else {
        scale = safe * std::pow(err, -alpha) * std::pow(errold, beta);
        if (scale < minscale) scale = minscale;
        if (scale > maxscale) scale = maxscale;
      }
      hnext = reject ? h * std::min(scale, 1.0) : h * scale;
      errold = std::max(err, 1.0e-4);
      reject = false;
      return true;
    }
    scale = std::max(safe * std::pow(err, -alpha), minscale);
    h *= scale;
    reject = true;
    return false;
  };
  double h = x2 - x1 > 0.0 ? std::fabs(h1) : -std::fabs(h1);
  derivs(x, y, dydx);
  save();
  const int max_step = 50000;
  for (int nstp = 0; nstp < max_step; ++nstp) {
    if ((x + h * 1.0001 - x2) * (x2 - x1) > 0.0) h = x2 - x;
    double hh = h;
    for (;;) {
      dy(hh);
      const double err = error();
      if (success(err, hh)) break;
      if (std::abs(hh) <= std::abs(x) * eps) {
        out.status = 2;
        return out;
      }
    }
    x += hh;
    hdid = hh;
    save();
    dydx = dydxnew;
    y = yout;
    h = hnext;
    if ((x - x2) * (x2 - x1) >= 0.0) {
      out.y_final = y;
      return out;
    }
    if (std::abs(h) <= hmin) {
      out.status = 2;
      return out;
    }
  }
  out.status = 1;
  return out;
}

// ------------------------------------------------------ conserved quantities --
// Softened total energy, G = 1: KE = sum m_i |v_i|^2 / 2,
// PE = -sum_{i<j} m_i m_j / sqrt(|r_ij|^2 + eps^2).  With unit masses this is
// scripts/nbody_physics/energy.py:224-241 (PE) + :239-241 (KE).  Accumulated in
// long double to make it an accuracy reference rather than an order model.
inline void energy(const double* x, size_t n, double softening,
                   const double* masses, double* ke_out, double* pe_out) {
  const double* pos = x;
  const double* vel = x + 3 * n;
  long double ke = 0.0L, pe = 0.0L;
  for (size_t i = 0; i < n; ++i) {
    long double mi = masses ? masses[i] : 1.0L;
    long double v2 = 0.0L;
    for (int c = 0; c < 3; ++c) v2 += (long double)vel[3 * i + c] * vel[3 * i + c];
    ke += 0.5L * mi * v2;
  }
  for (size_t i = 0; i < n; ++i) {
    long double mi = masses ? masses[i] : 1.0L;
    for (size_t j = i + 1; j < n; ++j) {
      long double mj = masses ? masses[j] : 1.0L;
      long double dr2 = 0.0L;
      for (int c = 0; c < 3; ++c) {
        long double d = (long double)pos[3 * j + c] - pos[3 * i + c];
        dr2 += d * d;
      }
      pe -= mi * mj / sqrtl(dr2 + (long double)softening * softening);
    }
  }
  *ke_out = static_cast<double>(ke);
  *pe_out = static_cast<double>(pe);
}

// Total linear momentum P = sum m_i v_i (absent from the reference; pairwise
// antisymmetric forces conserve it exactly in exact arithmetic).
inline void momentum(const double* x, size_t n, const double* masses,
                     double* p_out) {
  long double p[3] = {0.0L, 0.0L, 0.0L};
  const double* vel = x + 3 * n;
  for (size_t i = 0; i < n; ++i) {
    long double mi = masses ? masses[i] : 1.0L;
    for (int c = 0; c < 3; ++c) p[c] += mi * vel[3 * i + c];
  }
  for (int c = 0; c < 3; ++c) p_out[c] = static_cast<double>(p[c]);
}

}  // namespace oracle
