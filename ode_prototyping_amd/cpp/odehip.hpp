// C++20 host-side mirror of the reference's integration interfaces over the
// libodehip C ABI (include/odehip.h).  A reference user switches the
// execution policy and state container and keeps the rest:
//
//   reference (CUDA):                         this library (HIP, gfx950):
//   CudaExecutor exe;                         odehip::HipExecutor exe;
//   CudaState<double, N> x0{data};            odehip::HipState<double, N> x0{data};
//   RKEmbeddedParallel<BTDOPRI5>::integrate(  odehip::RKEmbeddedParallel<odehip::BTDOPRI5>::integrate(
//       x0, t0, tf, atol, rtol,                   x0, t0, tf, atol, rtol,
//       NBodyODE<double, N>{eps}, out, exe);      odehip::NBodyODE<double, N>{eps}, out, exe);
//
// Mirrored interfaces (reference file:line):
//   HipState<T, N>         src/state/CudaState.cuh:10-64 (ODEState concept,
//                          src/state/ODEState.hpp:25-43)
//   HipExecutor            src/par_exec/CudaExecutor.cuh:80-126
//   NBodyODE / VanDerPolODE / ExponentialODE   src/ode/*.hpp
//   RawOutput              src/output/RawOutput.hpp:8-27
//   RKEmbeddedParallel<BT>::integrate   src/integrator/RKEmbeddedParallel/
//                          RKEmbeddedParallel.hpp:13-81 (host-driven: every
//                          reference kernel launch is one C-ABI call, the
//                          controller runs on the host with std::pow, exactly
//                          the CUDA executor's control flow);
//                          integrate_on_device: the same solve with the attempt
//                          loop and controller resident on the GPU.
//   fill                   src/state/StateUtils.hpp:5-16
//   ParallelLeapfrogIntegrator::integrate   src/integrator/leapfrog/
//                          LeapfrogParallel.hpp:8-41 (device loop, every
//                          state returned to the Output)
//   solve_ensemble         many RKEmbeddedParallel::integrate calls at once
//                          (odehip_ensemble_solve; no reference counterpart)
// Errors: the reference prints CUDA errors and continues
// (src/cuda_util/CudaErrorCheck.cuh:8-16); here every failing C-ABI call
// throws odehip::Error with the library's message.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <concepts>
#include <cstddef>
#include <span>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "odehip.h"

namespace odehip {

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// HipExecutor{} wraps the process-wide context (one HIP stream) that HipState
// also allocates and copies on, as CudaExecutor{} and CudaState share the CUDA
// runtime's implicit context and legacy stream (CudaExecutor.cuh:80-126,
// CudaState.cuh:15-50): kernels and state copies are ordered.  HipExecutor(d)
// creates a separate context (own stream) on device d for callers that order
// their own work.
class HipExecutor {
 public:
  HipExecutor() : ctx_{shared_context()}, owned_{false} {}
  explicit HipExecutor(int device) : owned_{true} {
    if (odehip_create(device, &ctx_) != ODEHIP_OK)
      throw Error("odehip_create failed: no usable HIP device");
  }
  HipExecutor(const HipExecutor&) = delete;
  HipExecutor& operator=(const HipExecutor&) = delete;
  ~HipExecutor() {
    if (owned_) odehip_destroy(ctx_);
  }

  odehip_context* context() const { return ctx_; }
  void check(int rc, const char* what) const {
    if (rc != ODEHIP_OK)
      throw Error(std::string(what) + ": " + odehip_last_error(ctx_));
  }
  void synchronize() const { check(odehip_synchronize(ctx_), "synchronize"); }
  // Mixed RHS near field: pairs with r^2 < coefficient / n are FP64.
  void set_near_field(double coefficient) const {
    check(odehip_set_near_field(ctx_, coefficient), "set_near_field");
  }
  double near_field() const {
    double c = 0.0;
    check(odehip_get_near_field(ctx_, &c), "get_near_field");
    return c;
  }

  static odehip_context* shared_context() {
    static odehip_context* ctx = [] {
      odehip_context* c = nullptr;
      if (odehip_create(0, &c) != ODEHIP_OK)
        throw Error("odehip_create failed: no usable HIP device");
      return c;
    }();
    return ctx;
  }

 private:
  odehip_context* ctx_ = nullptr;
  bool owned_ = false;
};

// Owning device buffer of N values of T (T = double); CudaState semantics:
// zero-initialised default, construction from a host span, deep copy on copy
// construction/assignment, copy_to_span to the host.
template <typename T, std::size_t N>
class HipState {
  static_assert(std::is_same_v<T, double>, "libodehip states are FP64");

 public:
  static constexpr auto size() { return N; }

  HipState() { alloc(); }
  explicit HipState(std::span<const T, N> host) {
    alloc();
    exe().check(odehip_copy_h2d(exe().context(), ptr_, host.data(), N), "copy_h2d");
  }
  HipState(const HipState& other) {
    alloc();
    exe().check(odehip_copy_d2d(exe().context(), ptr_, other.ptr_, N), "copy_d2d");
  }
  HipState(HipState&& other) noexcept : ptr_{other.ptr_} { other.ptr_ = nullptr; }
  HipState& operator=(const HipState& other) {
    if (this != &other)
      exe().check(odehip_copy_d2d(exe().context(), ptr_, other.ptr_, N), "copy_d2d");
    return *this;
  }
  HipState& operator=(HipState&& other) noexcept {
    std::swap(ptr_, other.ptr_);
    return *this;
  }
  ~HipState() {
    if (ptr_) odehip_free(exe().context(), ptr_);
  }

  T* data() { return ptr_; }
  const T* data() const { return ptr_; }

  void copy_to_span(std::span<T, N> dest) const {
    exe().check(odehip_copy_d2h(exe().context(), dest.data(), ptr_, N), "copy_d2h");
  }
  friend auto span(HipState& s) { return std::span<T, N>{s.data(), N}; }
  friend auto span(const HipState& s) { return std::span<const T, N>{s.data(), N}; }

 private:
  struct Ctx {
    odehip_context* context() const { return HipExecutor::shared_context(); }
    void check(int rc, const char* what) const {
      if (rc != ODEHIP_OK)
        throw Error(std::string(what) + ": " + odehip_last_error(context()));
    }
  };
  static Ctx exe() { return {}; }
  void alloc() {
    exe().check(odehip_alloc(exe().context(), N, &ptr_), "alloc");
  }
  T* ptr_ = nullptr;
};

template <typename R>
HipState(R&&) -> HipState<typename std::remove_reference_t<R>::value_type,
                          std::tuple_size_v<std::remove_reference_t<R>>>;

// StateUtils fill (src/state/StateUtils.hpp:13-16).
template <typename T, std::size_t N>
void fill(HipExecutor& exe, HipState<T, N>& s, T value) {
  exe.check(odehip_fill(exe.context(), s.data(), value, N), "fill");
}

// ------------------------------------------------------------------ tableaux
// Tag types carrying the library's tableau id and (p, pt, S), named as the
// reference's butcher_tableaus/*.hpp structs.
template <int Id, int P, int PT, int S>
struct TableauTag {
  static constexpr int id = Id;
  static constexpr int p = P;
  static constexpr int pt = PT;
  static constexpr int n_stages = S;
};
using BTHE21 = TableauTag<ODEHIP_TABLEAU_HE21, 2, 1, 2>;
using BTRKF45 = TableauTag<ODEHIP_TABLEAU_RKF45, 4, 5, 6>;
using BTDOPRI5 = TableauTag<ODEHIP_TABLEAU_DOPRI5, 5, 4, 7>;
using BTDVERK = TableauTag<ODEHIP_TABLEAU_DVERK, 6, 5, 8>;
using BTRKF78 = TableauTag<ODEHIP_TABLEAU_RKF78, 7, 8, 13>;
using BTDOP853 = TableauTag<ODEHIP_TABLEAU_DOP853, 8, 5, 12>;

// A tableau type in the reference's own format (butcher_tableaus/BTDOPRI5.hpp:
// 5-26): static a (n_stages - 1 rows of stage weights), b, bt, p, pt, n_stages.
template <typename BT>
concept ReferenceTableau = requires {
  BT::a[0][0];
  BT::b[0];
  BT::bt[0];
  { BT::p } -> std::convertible_to<int>;
  { BT::pt } -> std::convertible_to<int>;
  { BT::n_stages } -> std::convertible_to<int>;
};

template <typename BT>
concept LibraryTableau = requires {
  { BT::id } -> std::convertible_to<int>;
};

// The library's id of tableau type BT: a tag's id, or a reference-format
// struct registered with the library on first use (odehip_tableau_register;
// the same coefficients always give the same id).
template <typename BT>
int tableau_id(HipExecutor& exe) {
  if constexpr (LibraryTableau<BT>) {
    return BT::id;
  } else {
    static_assert(ReferenceTableau<BT>,
                  "a tableau is a library tag or a struct with a, b, bt, p, pt, n_stages");
    constexpr int S = BT::n_stages;
    std::array<double, (S - 1) * (S - 1)> a{};
    std::array<double, S> b{}, bt{};
    for (int r = 0; r < S - 1; ++r)
      for (int j = 0; j <= r; ++j) a[r * (S - 1) + j] = BT::a[r][j];
    for (int j = 0; j < S; ++j) {
      b[j] = BT::b[j];
      bt[j] = BT::bt[j];
    }
    int id = -1;
    exe.check(odehip_tableau_register(exe.context(), S, BT::p, BT::pt, a.data(), b.data(),
                                      bt.data(), &id),
              "tableau_register");
    return id;
  }
}

// ---------------------------------------------------------------------- ODEs
enum class NBodyArithmetic { exact = ODEHIP_NBODY_EXACT, fast = ODEHIP_NBODY_FAST,
                             mixed = ODEHIP_NBODY_MIXED };

// src/ode/NBodyODE.hpp: softening as the reference's aggregate member; masses
// (device pointer, N entries) and arithmetic are extensions.
template <typename ValueType, std::size_t n_var>
struct NBodyODE {
  static constexpr auto n_particles = n_var / 6;
  double softening{};
  const double* masses = nullptr;
  NBodyArithmetic arithmetic = NBodyArithmetic::exact;

  void operator()(HipExecutor& exe, std::span<const ValueType, n_var> x,
                  std::span<ValueType, n_var> dxdt) const {
    exe.check(odehip_nbody_rhs(exe.context(), x.data(), dxdt.data(), n_particles,
                               softening, masses, static_cast<int>(arithmetic)),
              "nbody_rhs");
  }
  // NBodyODE::acc (NBodyODE.hpp:41-46)
  void acc(HipExecutor& exe, std::span<const ValueType, n_var> pos_vel,
           std::span<ValueType, n_var / 2> a) const {
    exe.check(odehip_nbody_acc(exe.context(), pos_vel.data(), a.data(), n_particles,
                               softening, masses, static_cast<int>(arithmetic)),
              "nbody_acc");
  }
  odehip_ode describe() const {
    return {ODEHIP_RHS_NBODY, static_cast<int>(arithmetic), n_var, softening, masses};
  }
};

template <typename ValueType>
struct VanDerPolODE {
  static constexpr auto n_var = 2;
  void operator()(HipExecutor& exe, std::span<const ValueType, 2> x,
                  std::span<ValueType, 2> dxdt) const {
    exe.check(odehip_ode_rhs(exe.context(), ODEHIP_RHS_VANDERPOL, x.data(), dxdt.data(), 2),
              "vdp_rhs");
  }
  odehip_ode describe() const { return {ODEHIP_RHS_VANDERPOL, 0, 2, 0.0, nullptr}; }
};

template <typename ValueType, std::size_t n_var>
struct ExponentialODE {
  void operator()(HipExecutor& exe, std::span<const ValueType, n_var> x,
                  std::span<ValueType, n_var> dxdt) const {
    exe.check(odehip_ode_rhs(exe.context(), ODEHIP_RHS_EXPONENTIAL, x.data(), dxdt.data(),
                             n_var),
              "exp_rhs");
  }
  odehip_ode describe() const { return {ODEHIP_RHS_EXPONENTIAL, 0, n_var, 0.0, nullptr}; }
};

// -------------------------------------------------------------------- output
// src/output/RawOutput.hpp:8-27 (host copies of every saved state).
template <std::size_t N>
class RawOutput {
 public:
  template <typename StateType>
  void save_state(double t, const StateType& x) {
    times_.push_back(t);
    states_.emplace_back();
    x.copy_to_span(std::span<double, N>{states_.back()});
  }
  const auto& times() const { return times_; }
  const auto& states() const { return states_; }

 private:
  std::vector<double> times_{};
  std::vector<std::array<double, N>> states_{};
};

// NumericalRecipes/Output.hpp:63-147 (the DenseOutput policy) for the device
// DOP853 or DOPRI5 solve: the initial state, then the state at x1 + k (x2 - x1) / n
// (accumulated by repeated addition, as the reference does) for every such
// point strictly before the end of the integration, from the continuous
// extension of the step that covers it.
template <std::size_t N>
class DenseOutput {
 public:
  static constexpr int default_n_intervals = 20;

  void set_n_intervals(int n_intervals) {
    if (n_intervals < 1) throw std::invalid_argument{"n_intervals must be at least 1."};
    n_intervals_ = n_intervals;
  }
  // The candidate points after x1 (those before x2).
  std::vector<double> output_times(double x1, double x2) const {
    std::vector<double> out;
    const double width = (x2 - x1) / n_intervals_;
    for (double x = x1 + width; x < x2 && out.size() <= static_cast<std::size_t>(n_intervals_);
         x += width)
      out.push_back(x);
    return out;
  }
  void save(double x, const double* y) {
    x_values_.push_back(x);
    for (std::size_t i = 0; i < N; ++i) y_values_[i].push_back(y[i]);
  }
  auto n_steps() const { return x_values_.size(); }
  const auto& x_values() const { return x_values_; }
  const auto& y_values() const { return y_values_; }

 private:
  int n_intervals_{default_n_intervals};
  std::vector<double> x_values_{};
  std::array<std::vector<double>, N> y_values_{};
};

// ---------------------------------------------------------------- integrator
struct DeviceSolveStats {
  long accepted = 0, rejected = 0, rhs_evals = 0;
  int status = 0;
  double t = 0.0;
};

template <typename BT>
struct RKEmbeddedParallel {
  // RKEmbeddedParallel.hpp:13-81, driven from the host: one C-ABI call per
  // reference kernel launch; the error norm is read back each attempt
  // (as CudaExecutor::transform_reduce does) and the controller runs here.
  template <typename ODE, typename Output, std::size_t N>
  static void integrate(HipState<double, N> x0, double t0, double tf,
                        HipState<double, N> atol, HipState<double, N> rtol, ODE ode,
                        Output& output, HipExecutor& exe) {
    constexpr auto n_var = N;
    constexpr auto max_step_scale = 6.0;
    constexpr auto min_step_scale = 0.33;
    constexpr auto q = BT::p < BT::pt ? BT::p : BT::pt;
    static const auto safety_factor = std::pow(0.38, (1.0 / (1.0 + q)));
    HipState<double, BT::n_stages * n_var> ks{};
    auto* ctx = exe.context();
    const int tab = tableau_id<BT>(exe);
    double dt = estimate_initial_step(exe, x0, atol, rtol, ode);
    double t = t0;
    auto x = x0;
    HipState<double, N> temp_state{}, error_estimate{}, error_target{};
    output.save_state(t, x);
    while (t < tf) {
      ode(exe, span(std::as_const(x0)), std::span<double, n_var>{ks.data(), n_var});
      for (int stage = 1; stage < BT::n_stages; ++stage) {
        exe.check(odehip_rk_stage(ctx, tab, stage, dt, temp_state.data(), ks.data(),
                                  x0.data(), n_var),
                  "rk_stage");
        ode(exe, span(std::as_const(temp_state)),
            std::span<double, n_var>{ks.data() + stage * n_var, n_var});
      }
      exe.check(odehip_rk_update_state_and_error(ctx, tab, dt, x.data(),
                                                 error_estimate.data(), ks.data(),
                                                 x0.data(), n_var),
                "rk_update_state_and_error");
      exe.check(odehip_rk_update_error_target(ctx, error_target.data(), atol.data(),
                                              rtol.data(), x.data(), x0.data(), n_var),
                "rk_update_error_target");
      double scaled_error = 0.0;
      exe.check(odehip_rk_norm(ctx, error_estimate.data(), error_target.data(), n_var,
                               &scaled_error),
                "rk_norm");
      if (!std::isfinite(scaled_error))
        throw Error("non-finite error norm (the reference would loop forever)");
      if (scaled_error <= 1.0) {
        t += dt;
        x0 = x;
        output.save_state(t, x);
      }
      auto dtnew = dt * safety_factor / std::pow(scaled_error, 1.0 / (1.0 + q));
      if (std::abs(dtnew) > max_step_scale * std::abs(dt)) {
        dt *= max_step_scale;
      } else if (std::abs(dtnew) < min_step_scale * std::abs(dt)) {
        dt *= min_step_scale;
      } else if (dtnew / (tf - t0) < 1.0e-12) {
        dt = (tf - t0) * 1.0e-12;
      } else {
        dt = dtnew;
      }
      if (t + dt > tf) dt = tf - t;
    }
  }

  // The same solve with the attempt loop, error norm, accept/reject and
  // controller on the device (odehip_rk_solve); the final state replaces x,
  // and the accepted times are returned.  max_accepted = 0: run to tf.
  template <typename ODE, std::size_t N>
  static DeviceSolveStats integrate_on_device(HipState<double, N>& x, double t0, double tf,
                                              const HipState<double, N>& atol,
                                              const HipState<double, N>& rtol,
                                              const ODE& ode, HipExecutor& exe,
                                              std::vector<double>* times = nullptr,
                                              long max_accepted = 0, bool fsal = false) {
    odehip_ode o = ode.describe();
    odehip_solve_options opt{};
    opt.t0 = t0;
    opt.tf = tf;
    opt.atol = atol.data();
    opt.rtol = rtol.data();
    opt.max_accepted = max_accepted;
    opt.fsal = fsal ? 1 : 0;
    if (times) {
      times->assign(max_accepted > 0 ? max_accepted + 1 : (1 << 20), 0.0);
      opt.times = times->data();
      opt.times_cap = static_cast<long>(times->size());
    }
    odehip_solve_stats st{};
    exe.check(odehip_rk_solve(exe.context(), tableau_id<BT>(exe), &o, &opt, x.data(), &st),
              "rk_solve");
    if (times) times->resize(st.n_times);
    return {st.accepted, st.rejected, st.rhs_evals, st.status, st.t};
  }

  // The device solve with dense output into a DenseOutput policy
  // (StepperDopr853 / StepperDopr5 ::prepare_dense, dense_out,
  // StepperDopr853.hpp:200-289, StepperDopr5.hpp:181-221).
  template <typename ODE, std::size_t N>
  static DeviceSolveStats integrate_on_device(HipState<double, N>& x, double t0, double tf,
                                              const HipState<double, N>& atol,
                                              const HipState<double, N>& rtol,
                                              const ODE& ode, DenseOutput<N>& output,
                                              HipExecutor& exe, bool fsal = true) {
    static_assert(LibraryTableau<BT>, "dense output needs the library's DOP853 or DOPRI5");
    static_assert(BT::id == ODEHIP_TABLEAU_DOP853 || BT::id == ODEHIP_TABLEAU_DOPRI5,
                  "dense output needs DOP853 or DOPRI5");
    std::array<double, N> first{};
    x.copy_to_span(std::span<double, N>{first});
    const std::vector<double> when = output.output_times(t0, tf);
    std::vector<double> rows(when.size() * N);
    odehip_ode o = ode.describe();
    odehip_solve_options opt{};
    opt.t0 = t0;
    opt.tf = tf;
    opt.atol = atol.data();
    opt.rtol = rtol.data();
    opt.fsal = fsal ? 1 : 0;
    opt.dense_times = when.data();
    opt.n_dense = static_cast<long>(when.size());
    opt.dense_states = rows.data();
    odehip_solve_stats st{};
    exe.check(odehip_rk_solve(exe.context(), BT::id, &o, &opt, x.data(), &st), "rk_solve");
    output.save(t0, first.data());
    for (long k = 0; k < st.n_dense; ++k) output.save(when[k], rows.data() + k * N);
    return {st.accepted, st.rejected, st.rhs_evals, st.status, st.t};
  }

 private:
  // detail::estimate_initial_step (RKEmbeddedParallel.hpp:170-200).
  template <typename ODE, std::size_t N>
  static double estimate_initial_step(HipExecutor& exe, const HipState<double, N>& x0,
                                      const HipState<double, N>& atol,
                                      const HipState<double, N>& rtol, const ODE& ode) {
    auto* ctx = exe.context();
    HipState<double, N> error_target{}, f0{}, x1{}, df{};
    exe.check(odehip_rk_compute_error_target(ctx, error_target.data(), x0.data(),
                                             atol.data(), rtol.data(), N),
              "compute_error_target");
    ode(exe, span(x0), span(f0));
    auto norm = [&](const HipState<double, N>& v) {
      double r = 0.0;
      exe.check(odehip_rk_norm(ctx, v.data(), error_target.data(), N, &r), "rk_norm");
      return r;
    };
    auto d0 = norm(x0);
    auto d1 = norm(f0);
    auto dt0 = (d0 < 1.0e-5 or d1 < 1.0e-5) ? 1.0e-6 : 0.01 * (d0 / d1);
    exe.check(odehip_rk_euler_step(ctx, x1.data(), x0.data(), f0.data(), dt0, N),
              "euler_step");
    ode(exe, span(std::as_const(x1)), span(df));
    exe.check(odehip_rk_difference(ctx, df.data(), f0.data(), N), "difference");
    auto d2 = norm(df) / dt0;
    constexpr auto p = BT::p;
    auto dt1 = (std::max(d1, d2) <= 1.0e-15)
                   ? std::max(1.0e-6, dt0 * 1.0e-3)
                   : std::pow(0.01 / std::max(d1, d2), (1.0 / (1.0 + p)));
    return std::min(100.0 * dt0, dt1);
  }
};

// A saved host state in the shape Output::save_state expects (copy_to_span).
template <std::size_t N>
struct HostStateView {
  const double* p;
  void copy_to_span(std::span<double, N> out) const { std::copy(p, p + N, out.begin()); }
};

// ParallelLeapfrogIntegrator::integrate (LeapfrogParallel.hpp:8-41): fixed-step
// velocity Verlet on the device (odehip_leapfrog); output.save_state receives
// the initial state and the state after every step, as in the reference.
struct ParallelLeapfrogIntegrator {
  template <std::size_t N, typename Output>
  static void integrate(HipState<double, N> x0, double t0, double tf, double dt,
                        const NBodyODE<double, N>& ode, Output& output, HipExecutor& exe) {
    long steps = 0;
    for (double t = t0; t < tf;) {  // the reference's time loop
      t += dt;
      ++steps;
    }
    std::vector<double> states(static_cast<std::size_t>(steps + 1) * N), times(steps + 1);
    odehip_ode o = ode.describe();
    odehip_leapfrog_stats st{};
    exe.check(odehip_leapfrog(exe.context(), &o, t0, tf, dt, x0.data(), states.data(),
                              times.data(), steps + 1, &st),
              "leapfrog");
    for (long k = 0; k < st.n_states; ++k)
      output.save_state(times[k], HostStateView<N>{states.data() + k * N});
  }
};

// Many independent N-body systems of different sizes in one device solve
// (odehip_ensemble_solve): states packed system after system, each system's
// result bit-identical to integrate_on_device on it alone.
struct EnsembleResult {
  odehip_ensemble_stats totals{};
  std::vector<odehip_system_stats> systems;
};

template <typename BT>
inline EnsembleResult solve_ensemble(HipExecutor& exe, const std::vector<int>& sizes,
                                     const double* x0_dev, double* x_out_dev,
                                     const odehip_ensemble_options& opt, double softening,
                                     const double* masses_dev = nullptr,
                                     NBodyArithmetic arithmetic = NBodyArithmetic::exact) {
  EnsembleResult r;
  r.systems.resize(sizes.size());
  exe.check(odehip_ensemble_solve(exe.context(), tableau_id<BT>(exe),
                                  static_cast<int>(sizes.size()),
                                  sizes.data(), static_cast<int>(arithmetic), softening,
                                  masses_dev, &opt, x0_dev, x_out_dev, &r.totals,
                                  r.systems.data()),
            "ensemble_solve");
  return r;
}

}  // namespace odehip
