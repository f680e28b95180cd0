// CPU ORACLE -- test infrastructure.  A driver of this repository's own that
// includes the reference's headers where they lie (the include path points
// at the reference tree; nothing is copied) and exposes a few fixed-size
// instantiations through extern "C" so tests can check the oracle restatement
// against the reference code itself, and bench.py can time the reference's
// thread-pool path as the CPU baseline.  Built by `make ref` into oracle/_ref/.
#include <algorithm>
#include <array>
#include <chrono>
#include <cstring>
#include <memory>
#include <span>
#include <vector>

#include "BTDOPRI5.hpp"
#include "BTRKF78.hpp"
#include "HeapState.hpp"
#include "LeapfrogParallel.hpp"
#include "NBodyODE.hpp"
#include "RKEmbeddedParallel.hpp"
#include "SingleThreadedExecutor.hpp"
#include "ThreadPoolExecutor.hpp"

// DOP853 written in the reference's tableau format (the reference ships the
// constants only for its NumericalRecipes stepper).  b = 8th-order weights,
// bt = b - er from the exact decimal difference of the published constants.
struct BTDOP853Ref {
  static constexpr auto a = std::array{
      std::array{5.26001519587677318785587544488e-2, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
      std::array{1.97250569845378994544595329183e-2, 5.91751709536136983633785987549e-2, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
      std::array{2.95875854768068491816892993775e-2, 0.0, 8.87627564304205475450678981324e-2, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
      std::array{2.41365134159266685502369798665e-1, 0.0, -8.84549479328286085344864962717e-1, 9.24834003261792003115737966543e-1, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
      std::array{3.7037037037037037037037037037e-2, 0.0, 0.0, 1.70828608729473871279604482173e-1, 1.25467687566822425016691814123e-1, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
      std::array{3.7109375e-2, 0.0, 0.0, 1.70252211019544039314978060272e-1, 6.02165389804559606850219397283e-2, -1.7578125e-2, 0.0, 0.0, 0.0, 0.0, 0.0},
      std::array{3.70920001185047927108779319836e-2, 0.0, 0.0, 1.70383925712239993810214054705e-1, 1.07262030446373284651809199168e-1, -1.53194377486244017527936158236e-2, 8.27378916381402288758473766002e-3, 0.0, 0.0, 0.0, 0.0},
      std::array{6.24110958716075717114429577812e-1, 0.0, 0.0, -3.36089262944694129406857109825e0, -8.68219346841726006818189891453e-1, 2.75920996994467083049415600797e1, 2.01540675504778934086186788979e1, -4.34898841810699588477366255144e1, 0.0, 0.0, 0.0},
      std::array{4.77662536438264365890433908527e-1, 0.0, 0.0, -2.48811461997166764192642586468e0, -5.90290826836842996371446475743e-1, 2.12300514481811942347288949897e1, 1.52792336328824235832596922938e1, -3.32882109689848629194453265587e1, -2.03312017085086261358222928593e-2, 0.0, 0.0},
      std::array{-9.3714243008598732571704021658e-1, 0.0, 0.0, 5.18637242884406370830023853209e0, 1.09143734899672957818500254654e0, -8.14978701074692612513997267357e0, -1.85200656599969598641566180701e1, 2.27394870993505042818970056734e1, 2.49360555267965238987089396762e0, -3.0467644718982195003823669022e0, 0.0},
      std::array{2.27331014751653820792359768449e0, 0.0, 0.0, -1.05344954667372501984066689879e1, -2.00087205822486249909675718444e0, -1.79589318631187989172765950534e1, 2.79488845294199600508499808837e1, -2.85899827713502369474065508674e0, -8.87285693353062954433549289258e0, 1.23605671757943030647266201528e1, 6.43392746015763530355970484046e-1}};
  static constexpr auto b = std::array{
      5.42937341165687622380535766363e-2, 0.0, 0.0, 0.0, 0.0,
      4.45031289275240888144113950566e0, 1.89151789931450038304281599044e0,
      -5.8012039600105847814672114227e0, 3.1116436695781989440891606237e-1,
      -1.52160949662516078556178806805e-1, 2.01365400804030348374776537501e-1,
      4.47106157277725905176885569043e-2};
  static constexpr auto bt = std::array{
      0.0411736891223738815055525466763, 0.0, 0.0, 0.0, 0.0,
      5.67546933912861332216170925866, 2.38727684897175057456422398564,
      -7.4655811424655713184287418377, 0.66149321570779357609756479137,
      -0.486340068375533557585910690905, 0.119442194318914635909069111371,
      0.0670659235916588857765328353543};
  static constexpr auto p = 8;
  static constexpr auto pt = 5;
  static constexpr auto n_stages = static_cast<int>(b.size());
};

namespace {

// Records every accepted state (like RawOutput) and can stop the solve after
// a budget of accepted steps by throwing out of save_state (the reference has
// no step budget of its own).
template <std::size_t N>
struct CollectOutput {
  std::vector<double> times;
  std::vector<double> last;
  long stop_after = 0;  // accepted steps; 0 = never
  long saves = 0;
  template <typename StateType>
  void save_state(double t, StateType const& x) {
    times.push_back(t);
    last.assign(x.data(), x.data() + N);
    if (stop_after > 0 && saves++ >= stop_after) throw saves;
  }
};

// Unit-mass reference RHS followed by a uniform mass scale m (used only for
// the CPU timing baseline: the reference has no masses).
template <std::size_t NV>
struct ScaledNBody {
  NBodyODE<double, NV> ode;
  double m;
  static constexpr void scale_kernel(std::size_t i, double* acc, double m) {
    acc[i] *= m;
  }
  void operator()(auto& exe, std::span<double const, NV> x,
                  std::span<double, NV> dxdt) {
    ode(exe, x, dxdt);
    if (m != 1.0) {
      call_parallel_kernel<scale_kernel>(exe, NV / 2, dxdt.data() + NV / 2, m);
    }
  }
};

template <typename BT, std::size_t NV, typename Exe>
long run_nbody(Exe& exe, const double* x0, double t0, double tf, double tol,
               double softening, double m, long stop_after, double* x_out,
               double* times_out, long times_cap) {
  auto init = std::make_unique<std::array<double, NV>>();
  std::copy(x0, x0 + NV, init->begin());
  auto tolv = std::make_unique<std::array<double, NV>>();
  tolv->fill(tol);
  auto xs = HeapState<double, NV>{*init};
  auto atol = HeapState<double, NV>{*tolv};
  auto rtol = HeapState<double, NV>{*tolv};
  CollectOutput<NV> out;
  out.stop_after = stop_after;
  ScaledNBody<NV> ode{NBodyODE<double, NV>{softening}, m};
  try {
    RKEmbeddedParallel<BT>::integrate(xs, t0, tf, atol, rtol, ode, out, exe);
  } catch (long) {
  }
  if (x_out) std::copy(out.last.begin(), out.last.end(), x_out);
  long n = static_cast<long>(out.times.size());
  if (times_out)
    std::copy(out.times.begin(), out.times.begin() + std::min(n, times_cap),
              times_out);
  return n;
}

}  // namespace

extern "C" {

// One reference NBodyODE evaluation (unit masses) at N = 64 or 1024.
int ref_nbody_rhs(long n, const double* x, double* dxdt, double softening) {
  SingleThreadedExecutor exe;
  if (n == 64) {
    constexpr std::size_t NV = 64 * 6;
    NBodyODE<double, NV>{softening}(exe, std::span<double const, NV>{x, NV},
                                    std::span<double, NV>{dxdt, NV});
    return 0;
  }
  if (n == 1024) {
    constexpr std::size_t NV = 1024 * 6;
    NBodyODE<double, NV>{softening}(exe, std::span<double const, NV>{x, NV},
                                    std::span<double, NV>{dxdt, NV});
    return 0;
  }
  return -1;
}

// Reference single-threaded integration of an N=64 unit-mass N-body system
// with DOPRI5 (tableau 0) or DOP853-in-RK-form (tableau 1).  Returns the
// number of saved outputs (accepted steps + 1).
long ref_integrate_nbody64(int tableau, const double* x0, double t0, double tf,
                           double tol, double softening, double* x_out,
                           double* times_out, long times_cap) {
  SingleThreadedExecutor exe;
  constexpr std::size_t NV = 64 * 6;
  if (tableau == 0)
    return run_nbody<BTDOPRI5, NV>(exe, x0, t0, tf, tol, softening, 1.0, 0,
                                   x_out, times_out, times_cap);
  return run_nbody<BTDOP853Ref, NV>(exe, x0, t0, tf, tol, softening, 1.0, 0,
                                    x_out, times_out, times_cap);
}

// Reference leapfrog (ParallelLeapfrogIntegrator, single-threaded) of an
// N=64 unit-mass system; returns the number of saved states.
long ref_leapfrog_nbody64(const double* x0, double t0, double tf, double dt,
                          double softening, double* x_out, double* times_out,
                          long times_cap) {
  SingleThreadedExecutor exe;
  constexpr std::size_t NV = 64 * 6;
  auto init = std::make_unique<std::array<double, NV>>();
  std::copy(x0, x0 + NV, init->begin());
  auto xs = HeapState<double, NV>{*init};
  CollectOutput<NV> out;
  NBodyODE<double, NV> ode{softening};
  ParallelLeapfrogIntegrator::integrate(xs, t0, tf, dt, ode, out, exe);
  if (x_out) std::copy(out.last.begin(), out.last.end(), x_out);
  long n = static_cast<long>(out.times.size());
  if (times_out)
    std::copy(out.times.begin(), out.times.begin() + std::min(n, times_cap), times_out);
  return n;
}

// CPU baseline: the reference's ThreadPoolExecutor path (pair-parallel RHS
// with atomic adds, spin-waiting workers) on N=16384 with DOP853 in RK form,
// uniform masses m applied by scaling the unit-mass RHS, stopped after
// `accepted` accepted steps.  Returns wall seconds of the integrate call.
double ref_tp_solve_16384_dop853(int threads, const double* x0, double tol,
                                 double softening, double m, long accepted,
                                 double tf, double* x_out) {
  constexpr std::size_t NV = 16384 * 6;
  ThreadPoolExecutor exe{threads};
  auto t_start = std::chrono::steady_clock::now();
  run_nbody<BTDOP853Ref, NV>(exe, x0, 0.0, tf, tol, softening, m, accepted,
                             x_out, nullptr, 0);
  auto t_end = std::chrono::steady_clock::now();
  return std::chrono::duration<double>(t_end - t_start).count();
}

// Same with N = 64 and DOPRI5 (the CPU-runnable configuration), run to tf.
double ref_solve_64_dopri5(int threads, const double* x0, double tol,
                           double softening, double m, double tf,
                           double* x_out, long* n_saved) {
  constexpr std::size_t NV = 64 * 6;
  auto t_start = std::chrono::steady_clock::now();
  long n;
  if (threads <= 1) {
    SingleThreadedExecutor exe;
    n = run_nbody<BTDOPRI5, NV>(exe, x0, 0.0, tf, tol, softening, m, 0, x_out,
                                nullptr, 0);
  } else {
    ThreadPoolExecutor exe{threads};
    n = run_nbody<BTDOPRI5, NV>(exe, x0, 0.0, tf, tol, softening, m, 0, x_out,
                                nullptr, 0);
  }
  auto t_end = std::chrono::steady_clock::now();
  if (n_saved) *n_saved = n;
  return std::chrono::duration<double>(t_end - t_start).count();
}

// CPU baseline of the ensemble workload: the reference's thread-pool (or,
// with threads <= 1, single-threaded) DOPRI5 integration of one system of n
// bodies to tf, uniform masses m by scaling the unit-mass RHS.  The reference
// sizes states at compile time, so n must be one of the instantiated sizes
// below: those of 12 systems spread over the size order of the 2048-system
// ensemble (N log-uniform in [16, 4096], seed 0).  Returns wall seconds
// (-1: size not instantiated); *n_saved = accepted steps + 1.
double ref_solve_dopri5_n(int n, int threads, const double* x0, double tol,
                          double softening, double m, double tf, double* x_out,
                          long* n_saved) {
  auto run = [&]<std::size_t N>() {
    constexpr std::size_t NV = N * 6;
    auto t_start = std::chrono::steady_clock::now();
    long k;
    if (threads <= 1) {
      SingleThreadedExecutor exe;
      k = run_nbody<BTDOPRI5, NV>(exe, x0, 0.0, tf, tol, softening, m, 0, x_out, nullptr, 0);
    } else {
      ThreadPoolExecutor exe{threads};
      k = run_nbody<BTDOPRI5, NV>(exe, x0, 0.0, tf, tol, softening, m, 0, x_out, nullptr, 0);
    }
    auto t_end = std::chrono::steady_clock::now();
    if (n_saved) *n_saved = k;
    return std::chrono::duration<double>(t_end - t_start).count();
  };
  switch (n) {
    case 16: return run.template operator()<16>();
    case 26: return run.template operator()<26>();
    case 43: return run.template operator()<43>();
    case 71: return run.template operator()<71>();
    case 116: return run.template operator()<116>();
    case 192: return run.template operator()<192>();
    case 332: return run.template operator()<332>();
    case 542: return run.template operator()<542>();
    case 912: return run.template operator()<912>();
    case 1528: return run.template operator()<1528>();
    case 2474: return run.template operator()<2474>();
    case 4085: return run.template operator()<4085>();
    default: return -1.0;
  }
}

}  // extern "C"
