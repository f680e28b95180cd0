// CPU ORACLE -- test infrastructure.  extern "C" wrapper over oracle.hpp so
// the Python tests can drive the restatement through ctypes.  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg load this library.
#include <cstring>
#include <numbers>
#include <random>
#include <memory>
#include <string>

#include "mlp.hpp"
#include "oracle.hpp"

using namespace oracle;

namespace {
constexpr int kRhsNBody = 0;
constexpr int kRhsExponential = 1;
constexpr int kRhsVanDerPol = 2;

Rhs make_rhs(int kind, size_t n_var, double softening, const double* masses) {
  if (kind == kRhsNBody) {
    NBody nb{n_var / 6, softening, masses};
    return [nb](const double* x, double* d, const Executor& e) { nb(x, d, e); };
  }
  if (kind == kRhsExponential) {
    Exponential ex{n_var};
    return [ex](const double* x, double* d, const Executor& e) { ex(x, d, e); };
  }
  if (kind == kRhsVanDerPol) {
    return [](const double* x, double* d, const Executor& e) {
      VanDerPol{}(x, d, e);
    };
  }
  throw std::invalid_argument("unknown rhs kind");
}

Executor make_exe(int reduce_kind, long reduce_param, int threads) {
  Executor e;
  e.order = static_cast<ReduceOrder>(reduce_kind);
  e.param = reduce_param;
  e.threads = threads;
  return e;
}
}  // namespace

extern "C" {

// One RHS evaluation.
int oracle_rhs(int kind, const double* x, double* dxdt, size_t n_var,
               double softening, const double* masses, int threads) {
  try {
    Executor e = make_exe(0, 0, threads);
    make_rhs(kind, n_var, softening, masses)(x, dxdt, e);
    return 0;
  } catch (...) {
    return -1;
  }
}

// rk_norm with a chosen reduction order.
double oracle_rk_norm(const double* v, const double* sc, size_t n,
                      int reduce_kind, long reduce_param) {
  Executor e = make_exe(reduce_kind, reduce_param, 1);
  return detail::rk_norm(e, v, sc, n);
}

// Full integration; returns an opaque result handle (nullptr on error).
void* oracle_integrate(const char* tableau, int rhs_kind, const double* x0,
                       size_t n_var, double t0, double tf, const double* atol,
                       const double* rtol, double softening,
                       const double* masses, int reduce_kind, long reduce_param,
                       int pow_mode, long max_accepted, long max_attempts,
                       int store_states, int store_trace, int threads,
                       int controller, const double* dense_times, long n_dense) {
  try {
    Tableau bt = tableau_by_name(tableau);
    Executor e = make_exe(reduce_kind, reduce_param, threads);
    Options o;
    o.controller = static_cast<Controller>(controller);
    o.max_accepted = max_accepted;
    o.max_attempts = max_attempts;
    o.pow_mode = static_cast<PowMode>(pow_mode);
    o.store_states = store_states != 0;
    o.store_trace = store_trace != 0;
    if (dense_times && n_dense > 0) o.dense_times.assign(dense_times, dense_times + n_dense);
    auto* r = new Result(integrate(bt, x0, n_var, t0, tf, atol, rtol,
                                   make_rhs(rhs_kind, n_var, softening, masses),
                                   e, o));
    return r;
  } catch (...) {
    return nullptr;
  }
}

// Leapfrog (velocity Verlet) of an N-body state; the result handle is read
// with the oracle_result_* accessors (accepted = steps).
void* oracle_leapfrog(const double* x0, size_t n_var, double t0, double tf,
                      double dt, double softening, const double* masses,
                      int store_states, int threads) {
  try {
    Executor e = make_exe(0, 0, threads);
    return new Result(leapfrog(x0, n_var, t0, tf, dt,
                               make_rhs(kRhsNBody, n_var, softening, masses), e,
                               store_states != 0));
  } catch (...) {
    return nullptr;
  }
}

long oracle_result_count(void* h) {
  return static_cast<long>(static_cast<Result*>(h)->times.size());
}
void oracle_result_times(void* h, double* out) {
  auto* r = static_cast<Result*>(h);
  std::memcpy(out, r->times.data(), r->times.size() * sizeof(double));
}
int oracle_result_state(void* h, long k, double* out) {
  auto* r = static_cast<Result*>(h);
  if (r->states.empty() || k < 0 || k >= static_cast<long>(r->times.size()))
    return -1;
  std::memcpy(out, r->states.data() + k * r->n_var, r->n_var * sizeof(double));
  return 0;
}
void oracle_result_final(void* h, double* out) {
  auto* r = static_cast<Result*>(h);
  std::memcpy(out, r->x_final.data(), r->n_var * sizeof(double));
}
// stats: accepted, rejected, rhs_evals, status; dstats: t, dt, dt_initial
void oracle_result_stats(void* h, long* stats, double* dstats) {
  auto* r = static_cast<Result*>(h);
  stats[0] = r->accepted;
  stats[1] = r->rejected;
  stats[2] = r->rhs_evals;
  stats[3] = r->status;
  dstats[0] = r->t;
  dstats[1] = r->dt;
  dstats[2] = r->dt_initial;
}
long oracle_result_trace_len(void* h) {
  return static_cast<long>(static_cast<Result*>(h)->trace_err.size());
}
void oracle_result_trace(void* h, double* err, double* dt, int* acc) {
  auto* r = static_cast<Result*>(h);
  size_t n = r->trace_err.size();
  std::memcpy(err, r->trace_err.data(), n * sizeof(double));
  std::memcpy(dt, r->trace_dt.data(), n * sizeof(double));
  std::memcpy(acc, r->trace_acc.data(), n * sizeof(int));
}
void oracle_result_free(void* h) { delete static_cast<Result*>(h); }

// NR-compatible DOP853 on the two NR test problems.  problem 0: harmonic
// oscillator y'' = -y; problem 1: Van der Pol with stiffness parameter mu_eps
// (dydx[1] = ((1 - y0^2) y1 - y0) / mu_eps).  Writes the count of saved
// points, and x/y at indices 0, count/2 and count-1 into sel[0..8]
// (x0,y00,y01, xm,ym0,ym1, xl,yl0,yl1).
namespace {
NrDerivs nr_problem(int problem, double mu_eps) {
  NrDerivs d;
  if (problem == 0) {
    d = [](double, const std::vector<double>& y, std::vector<double>& dydx) {
      dydx[0] = y[1];
      dydx[1] = -y[0];
    };
  } else {
    d = [mu_eps](double, const std::vector<double>& y,
                 std::vector<double>& dydx) {
      dydx[0] = y[1];
      dydx[1] = ((1.0 - y[0] * y[0]) * y[1] - y[0]) / mu_eps;
    };
  }
  return d;
}
}  // namespace

int oracle_nr_dop853(int problem, double y0a, double y0b, double x1, double x2,
                     double atol, double rtol, double h1, double hmin,
                     double mu_eps, long* count, double* sel) {
  NrResult r = nr_dop853({y0a, y0b}, x1, x2, atol, rtol, h1, hmin,
                         nr_problem(problem, mu_eps));
  long c = static_cast<long>(r.xs.size());
  *count = c;
  long idx[3] = {0, c / 2, c - 1};
  for (int k = 0; k < 3; ++k) {
    sel[3 * k] = r.xs[idx[k]];
    sel[3 * k + 1] = r.ys[idx[k]][0];
    sel[3 * k + 2] = r.ys[idx[k]][1];
  }
  return r.status;
}

// The same with the DenseOutput policy (`intervals` equal intervals): writes
// up to cap saved points (x, y0, y1) and the final state; returns the status
// and the number of saved points in *count.
int oracle_nr_dop853_dense(int problem, double y0a, double y0b, double x1, double x2,
                           double atol, double rtol, double h1, double hmin,
                           double mu_eps, int intervals, long cap, long* count,
                           double* xs, double* y0s, double* y1s, double* y_final) {
  NrResult r = nr_dop853({y0a, y0b}, x1, x2, atol, rtol, h1, hmin,
                         nr_problem(problem, mu_eps), intervals);
  const long c = static_cast<long>(r.xs.size());
  *count = c;
  for (long k = 0; k < c && k < cap; ++k) {
    xs[k] = r.xs[k];
    y0s[k] = r.ys[k][0];
    y1s[k] = r.ys[k][1];
  }
  if (r.y_final.size() == 2) {
    y_final[0] = r.y_final[0];
    y_final[1] = r.y_final[1];
  }
  return r.status;
}

// NR-compatible DOPRI5 (StepperDopr5 + RawOutput / DenseOutput), as
// oracle_nr_dop853_dense; intervals = 0 selects RawOutput.
int oracle_nr_dopr5(int problem, double y0a, double y0b, double x1, double x2, double atol,
                    double rtol, double h1, double hmin, double mu_eps, int intervals, long cap,
                    long* count, double* xs, double* y0s, double* y1s, double* y_final) {
  NrResult r = nr_dopr5({y0a, y0b}, x1, x2, atol, rtol, h1, hmin, nr_problem(problem, mu_eps),
                        intervals);
  const long c = static_cast<long>(r.xs.size());
  *count = c;
  for (long k = 0; k < c && k < cap; ++k) {
    xs[k] = r.xs[k];
    y0s[k] = r.ys[k][0];
    y1s[k] = r.ys[k][1];
  }
  if (r.y_final.size() == 2) {
    y_final[0] = r.y_final[0];
    y_final[1] = r.y_final[1];
  }
  return r.status;
}

// One DOPRI5 step of length h from y (RK form) followed by its continuous
// extension at the fractions s_k of the step, as oracle_dop853_step_dense.
int oracle_dopri5_step_dense(int rhs_kind, const double* y, size_t n_var, double h,
                             double softening, const double* masses, const double* s_list,
                             long n_s, double* yout, double* out) {
  try {
    Executor e = make_exe(0, 0, 1);
    Rhs f = make_rhs(rhs_kind, n_var, softening, masses);
    Tableau bt = dopri5();
    const size_t n = n_var;
    std::vector<double> ks(7 * n), tmp(n), rc;
    f(y, ks.data(), e);
    for (int stage = 1; stage < 7; ++stage) {
      for (size_t i = 0; i < n; ++i) {
        double acc = 0.0;
        for (int j = 0; j < stage; ++j) acc += bt.A(stage - 1, j) * ks[j * n + i];
        tmp[i] = y[i] + acc * h;
      }
      f(tmp.data(), ks.data() + stage * n, e);
    }
    for (size_t i = 0; i < n; ++i) {
      double xs = 0.0;
      for (int j = 0; j < 7; ++j) xs += bt.b[j] * ks[j * n + i];
      yout[i] = y[i] + xs * h;
    }
    const double* kp[7];
    for (int j = 0; j < 7; ++j) kp[j] = ks.data() + j * n;
    detail::dopri5_prepare_dense(dopri5_dense(), n, h, y, yout, kp, rc);
    for (long k = 0; k < n_s; ++k)
      for (size_t i = 0; i < n; ++i)
        out[k * n + i] = detail::dopri5_dense_eval(rc, n, i, s_list[k] * h, h, h);
    return 0;
  } catch (...) {
    return -1;
  }
}

// One DOP853 step of length h from y (RK form: the twelve stages and the
// 8th-order update of RKEmbeddedParallel.hpp:37-55) followed by its
// continuous extension evaluated at the fractions s_k of the step
// (xout = s_k h with the step ending at x_end = h): out is n_s x n_var.
int oracle_dop853_step_dense(int rhs_kind, const double* y, size_t n_var, double h,
                             double softening, const double* masses, const double* s_list,
                             long n_s, double* yout, double* out) {
  try {
    Executor e = make_exe(0, 0, 1);
    Rhs f = make_rhs(rhs_kind, n_var, softening, masses);
    Tableau bt = dop853();
    const size_t n = n_var;
    std::vector<double> ks(12 * n), tmp(n), k13(n), rc, extra;
    f(y, ks.data(), e);
    for (int stage = 1; stage < 12; ++stage) {
      for (size_t i = 0; i < n; ++i) {
        double acc = 0.0;
        for (int j = 0; j < stage; ++j) acc += bt.A(stage - 1, j) * ks[j * n + i];
        tmp[i] = y[i] + acc * h;
      }
      f(tmp.data(), ks.data() + stage * n, e);
    }
    for (size_t i = 0; i < n; ++i) {
      double xs = 0.0;
      for (int j = 0; j < 12; ++j) xs += bt.b[j] * ks[j * n + i];
      yout[i] = y[i] + xs * h;
    }
    f(yout, k13.data(), e);
    const double* kp[13];
    for (int j = 0; j < 12; ++j) kp[j] = ks.data() + j * n;
    kp[12] = k13.data();
    detail::dop853_prepare_dense(
        dop853_dense(), n, h, y, yout, kp,
        [&](int, const double* yt, double* o) { f(yt, o, e); }, rc, extra);
    for (long k = 0; k < n_s; ++k)
      for (size_t i = 0; i < n; ++i)
        out[k * n + i] = detail::dop853_dense_eval(rc, n, i, s_list[k] * h, h, h);
    return 0;
  } catch (...) {
    return -1;
  }
}

// Dense-output results of an oracle_integrate handle.
long oracle_result_n_dense(void* h) { return static_cast<Result*>(h)->n_dense; }
void oracle_result_dense(void* h, double* out, long* counts) {
  auto* r = static_cast<Result*>(h);
  std::memcpy(out, r->dense_states.data(), r->dense_states.size() * sizeof(double));
  counts[0] = r->n_dense;
  counts[1] = r->dense_steps;
  counts[2] = r->dense_rhs_evals;
}

void oracle_energy(const double* x, size_t n, double softening,
                   const double* masses, double* ke, double* pe) {
  energy(x, n, softening, masses, ke, pe);
}

void oracle_momentum(const double* x, size_t n, const double* masses,
                     double* p) {
  momentum(x, n, masses, p);
}

// Registers a tableau under `name` (a: (S-1) x (S-1) row-major, entries
// j >= row + 1 ignored; b, bt: S), as odehip_tableau_register does for the
// library; later integrate calls take it by that name.
int oracle_register_tableau(const char* name, int S, int p, int pt, const double* a,
                            const double* b, const double* bt) {
  if (!name || S < 2 || S > 13 || !a || !b || !bt) return -1;
  Tableau t;
  t.name = name;
  t.S = S;
  t.p = p;
  t.pt = pt;
  t.a.assign(static_cast<size_t>(S - 1) * (S - 1), 0.0);
  for (int r = 0; r < S - 1; ++r)
    for (int j = 0; j <= r; ++j) t.a[r * (S - 1) + j] = a[r * (S - 1) + j];
  t.b.assign(b, b + S);
  t.bt.assign(bt, bt + S);
  registered_tableaus()[t.name] = t;
  return 0;
}

// Tableau access for tests (order conditions, row sums).
int oracle_tableau(const char* name, int* S, int* p, int* pt, double* a,
                   double* b, double* bt) {
  try {
    Tableau t = tableau_by_name(name);
    *S = t.S;
    *p = t.p;
    *pt = t.pt;
    if (a) std::memcpy(a, t.a.data(), t.a.size() * sizeof(double));
    if (b) std::memcpy(b, t.b.data(), t.b.size() * sizeof(double));
    if (bt) std::memcpy(bt, t.bt.data(), t.bt.size() * sizeof(double));
    return 0;
  } catch (...) {
    return -1;
  }
}

// SpinningCubeNBodyTest<N> initial state and parameters
// (src/test/SpinningCubeNBodyTest.hpp:55-103): mt19937{0} with libstdc++'s
// uniform_real_distribution(0, L) for positions, spin about the vertical axis.
// params: softening, tf, tol.
void oracle_spinning_cube_ic(long n, double* x, double* params) {
  const double L = 1.0;
  const long n_var = 6 * n;
  const double tf = std::sqrt(L * L * L / n);
  const double omega = 2.0 * std::sqrt(std::numbers::pi / 3.0) / std::sqrt(L * L * L / n);
  const double softening = 4.0 * L / std::sqrt(static_cast<double>(n));
  auto gen = std::mt19937{0};
  auto dist = std::uniform_real_distribution<double>(0.0, L);
  for (long i = 0; i < n_var / 2; ++i) x[i] = dist(gen);
  for (long i = 0; i < n; ++i) {
    const double xx = x[3 * i] - L / 2.0;
    const double yy = x[3 * i + 1] - L / 2.0;
    x[n_var / 2 + 3 * i] = -omega * yy;
    x[n_var / 2 + 3 * i + 1] = omega * xx;
    x[n_var / 2 + 3 * i + 2] = 0.0;
  }
  params[0] = softening;
  params[1] = tf;
  params[2] = softening;
}

double oracle_pow(double x, double y, int mode) {
  return opow(x, y, static_cast<PowMode>(mode));
}

// Learned-controller network (mlp.hpp): outputs for n samples of raw
// features [n x 10]; bf16 != 0 rounds weights and layer inputs to bfloat16.
void oracle_mlp_forward(const double* params, const double* mean,
                        const double* stdv, const double* raw, long n, int bf16,
                        double* out) {
  for (long s = 0; s < n; ++s)
    out[s] = mlp_forward_one(params, mean, stdv, raw + s * kIn, bf16 != 0);
}

double oracle_mlp_loss_grad(const double* params, const double* mean,
                            const double* stdv, const double* raw,
                            const double* target, long n, double* grad) {
  return mlp_loss_grad(params, mean, stdv, raw, target, n, grad);
}

}  // extern "C"
