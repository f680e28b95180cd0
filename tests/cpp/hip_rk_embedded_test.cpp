// The C++ host mirror (ode_prototyping_amd/cpp/odehip.hpp) driven the way the
// reference's own suites drive the CUDA policy: src/test/CudaRKEmbeddedTest.cu,
// CudaStateTest.cu and CudaExecutorTest.cu, with the golden values of
// src/test/SingleThreadedRKEmbeddedTest.cpp (copied as data; file:line cited
// per case).  GTest is not available, so a minimal EXPECT harness stands in;
// the program prints one line per check and exits non-zero on any failure.
//
// Expected agreement: the host-driven integrate uses the host's std::pow and
// the library's canonical reduction order, which for a 2-component state is
// the single-threaded order -> Van der Pol goldens hold to EXPECT_DOUBLE_EQ;
// larger states get the thread-pool suite's tolerances
// (ThreadPoolRKEmbeddedTest.cpp:314-351, :362-400).
#include <cmath>
#include <cstdio>
#include <memory>
#include <numbers>
#include <numeric>
#include <random>
#include <string>

#include "odehip.hpp"

using namespace odehip;

static int g_failures = 0;
static int g_checks = 0;

static bool ulp_eq(double a, double b) {
  if (a == b) return true;
  const double d = std::abs(a - b);
  const double u = std::nextafter(std::max(std::abs(a), std::abs(b)), INFINITY) -
                   std::max(std::abs(a), std::abs(b));
  return d <= 4 * u;
}

#define EXPECT_EQ(a, b) check((a) == (b), #a " == " #b, __LINE__, double(a), double(b))
#define EXPECT_DOUBLE_EQ(a, b) check(ulp_eq((a), (b)), "DOUBLE_EQ " #a ", " #b, __LINE__, a, b)
#define EXPECT_NEAR(a, b, tol) \
  check(std::abs((a) - (b)) <= (tol), "NEAR " #a ", " #b, __LINE__, a, b)

static void check(bool ok, const char* what, int line, double a, double b) {
  ++g_checks;
  if (!ok) {
    ++g_failures;
    std::printf("FAIL line %d: %s (%.17g vs %.17g)\n", line, what, a, b);
  }
}

// --------------------------------------------------------------- fixtures
// VanDerPolTest.hpp:10-24
struct VanDerPolCase {
  static constexpr std::size_t n_var = 2;
  HipState<double, 2> x0{std::array{2.0, 0.0}};
  double t0 = 0.0, tf = 2.0, tol = 1.0e-10;
  HipState<double, 2> atol{std::array{1.0e-10, 1.0e-10}};
  HipState<double, 2> rtol = atol;
  VanDerPolODE<double> ode{};
  RawOutput<2> output{};
};

// ExponentialTest.hpp:12-36
struct ExponentialCase {
  static constexpr std::size_t n_var = 10;
  static std::array<double, 10> iota() {
    std::array<double, 10> a{};
    std::iota(a.begin(), a.end(), 0.0);
    return a;
  }
  static std::array<double, 10> tols() {
    std::array<double, 10> a{};
    a.fill(1.0e-6);
    return a;
  }
  HipState<double, 10> x0{iota()};
  double t0 = 0.0, tf = 10.0;
  HipState<double, 10> atol{tols()};
  HipState<double, 10> rtol = atol;
  ExponentialODE<double, 10> ode{};
  RawOutput<10> output{};
};

// NBodyTest.hpp:13-37
struct NBodyCase {
  static constexpr std::size_t n_var = 30;
  static constexpr std::array<double, 30> x0_data{
      1.657666,  0.0,       0.0, 0.439775,  -0.169717, 0.0,
      -1.268608, -0.267651, 0.0, -1.268608, 0.267651,  0.0,
      0.439775,  0.169717,  0.0, 0.0,       -0.593786, 0.0,
      1.822785,  0.128248,  0.0, 1.271564,  0.168645,  0.0,
      -1.271564, 0.168645,  0.0, -1.822785, 0.128248,  0.0};
  static std::array<double, 30> tols() {
    std::array<double, 30> a{};
    a.fill(1.0e-10);
    return a;
  }
  HipState<double, 30> x0{x0_data};
  double t0 = 0.0, tf = 6.3, tol = 1.0e-10;
  HipState<double, 30> atol{tols()};
  HipState<double, 30> rtol = atol;
  NBodyODE<double, 30> ode{};
  RawOutput<30> output{};
};

// SpinningCubeNBodyTest.hpp:53-113 (N = 1024)
struct SpinningCubeCase {
  static constexpr int N = 1024;
  static constexpr std::size_t n_var = 6 * N;
  static constexpr double L = 1.0;
  double t0 = 0.0;
  double tf = std::sqrt(L * L * L / N);
  double omega = 2.0 * std::sqrt(std::numbers::pi / 3.0) / std::sqrt(L * L * L / N);
  double softening = 4.0 * L / std::sqrt(N);
  std::unique_ptr<HipState<double, n_var>> x0, atol, rtol;
  NBodyODE<double, n_var> ode{};
  RawOutput<n_var> output{};
  SpinningCubeCase() {
    auto gen = std::mt19937{0};
    auto dist = std::uniform_real_distribution<double>(0.0, L);
    auto data = std::make_unique<std::array<double, n_var>>();
    for (std::size_t i = 0; i < n_var / 2; ++i) (*data)[i] = dist(gen);
    for (int i = 0; i < N; ++i) {
      const double x = (*data)[3 * i] - L / 2.0;
      const double y = (*data)[3 * i + 1] - L / 2.0;
      (*data)[n_var / 2 + 3 * i] = -omega * y;
      (*data)[n_var / 2 + 3 * i + 1] = omega * x;
      (*data)[n_var / 2 + 3 * i + 2] = 0.0;
    }
    x0 = std::make_unique<HipState<double, n_var>>(*data);
    auto tol = std::make_unique<std::array<double, n_var>>();
    tol->fill(softening);
    atol = std::make_unique<HipState<double, n_var>>(*tol);
    rtol = std::make_unique<HipState<double, n_var>>(*tol);
    ode.softening = softening;
  }
};

// ------------------------------------------------------------------- tests
struct VdpGolden {
  const char* name;
  std::size_t count;
  double t_mid, mid0, mid1, last0, last1;
};

template <typename BT>
static void van_der_pol_consistency(const VdpGolden& g) {
  HipExecutor executor{};
  VanDerPolCase test;
  RKEmbeddedParallel<BT>::integrate(test.x0, test.t0, test.tf, test.atol, test.rtol, test.ode,
                                    test.output, executor);
  const auto& times = test.output.times();
  const auto& states = test.output.states();
  EXPECT_EQ(g.count, times.size());
  EXPECT_DOUBLE_EQ(0.0, times.front());
  EXPECT_DOUBLE_EQ(g.t_mid, times[times.size() / 2]);
  EXPECT_DOUBLE_EQ(2.0, times.back());
  EXPECT_DOUBLE_EQ(2.0, states.front()[0]);
  EXPECT_DOUBLE_EQ(0.0, states.front()[1]);
  EXPECT_DOUBLE_EQ(g.mid0, states[states.size() / 2][0]);
  EXPECT_DOUBLE_EQ(g.mid1, states[states.size() / 2][1]);
  EXPECT_DOUBLE_EQ(g.last0, states.back()[0]);
  EXPECT_DOUBLE_EQ(g.last1, states.back()[1]);
  std::printf("ok %s\n", g.name);
}

// A user's tableau in the reference's struct format (the layout of
// butcher_tableaus/BT*.hpp: a, b, bt, p, pt, n_stages) for a method the
// library does not ship: Cash & Karp's embedded 5(4) pair (ACM TOMS 16, 1990).
struct BTCashKarp {
  static constexpr auto a = std::array{
      std::array{1.0 / 5.0, 0.0, 0.0, 0.0, 0.0},
      std::array{3.0 / 40.0, 9.0 / 40.0, 0.0, 0.0, 0.0},
      std::array{3.0 / 10.0, -9.0 / 10.0, 6.0 / 5.0, 0.0, 0.0},
      std::array{-11.0 / 54.0, 5.0 / 2.0, -70.0 / 27.0, 35.0 / 27.0, 0.0},
      std::array{1631.0 / 55296.0, 175.0 / 512.0, 575.0 / 13824.0, 44275.0 / 110592.0,
                 253.0 / 4096.0}};
  static constexpr auto b =
      std::array{37.0 / 378.0, 0.0, 250.0 / 621.0, 125.0 / 594.0, 0.0, 512.0 / 1771.0};
  static constexpr auto bt = std::array{2825.0 / 27648.0, 0.0, 18575.0 / 48384.0,
                                        13525.0 / 55296.0, 277.0 / 14336.0, 1.0 / 4.0};
  static constexpr auto p = 5;
  static constexpr auto pt = 4;
  static constexpr auto n_stages = static_cast<int>(b.size());
};
static_assert(ReferenceTableau<BTCashKarp> && !LibraryTableau<BTCashKarp>);

// The Cash-Karp struct crosses the C ABI (registered on first use, an id past
// the built-in ones) and integrates Van der Pol on the host-driven and the
// device-resident paths; both end within 1e-8 of the reference's DOPRI5
// golden end state (SingleThreadedRKEmbeddedTest.cpp:138-160, tol 1e-10).
static void user_tableau_struct() {
  HipExecutor executor{};
  const int id = tableau_id<BTCashKarp>(executor);
  EXPECT_EQ(id >= ODEHIP_NUM_TABLEAUS, true);
  EXPECT_EQ(tableau_id<BTCashKarp>(executor), id);
  int S = 0, p = 0, pt = 0;
  EXPECT_EQ(odehip_tableau_info(id, &S, &p, &pt), ODEHIP_OK);
  EXPECT_EQ(S, 6);
  EXPECT_EQ(p, 5);
  VanDerPolCase test;
  RKEmbeddedParallel<BTCashKarp>::integrate(test.x0, test.t0, test.tf, test.atol, test.rtol,
                                            test.ode, test.output, executor);
  const auto& states = test.output.states();
  EXPECT_DOUBLE_EQ(2.0, test.output.times().back());
  EXPECT_NEAR(0.32331666704309109, states.back()[0], 1e-8);
  EXPECT_NEAR(-1.8329745679718159, states.back()[1], 1e-8);
  VanDerPolCase dev;
  const auto st = RKEmbeddedParallel<BTCashKarp>::integrate_on_device(
      dev.x0, dev.t0, dev.tf, dev.atol, dev.rtol, dev.ode, executor);
  std::array<double, 2> last{};
  dev.x0.copy_to_span(std::span<double, 2>{last});
  EXPECT_DOUBLE_EQ(2.0, st.t);
  EXPECT_NEAR(0.32331666704309109, last[0], 1e-8);
  EXPECT_NEAR(-1.8329745679718159, last[1], 1e-8);
  std::printf("ok user tableau struct (Cash-Karp, id %d, %zu accepted host-driven, %ld "
              "device)\n", id, states.size() - 1, st.accepted);
}

static void dopri5_exponential() {
  // SingleThreadedRKEmbeddedTest.cpp:162-192, thread-pool-style tolerance
  HipExecutor executor{};
  ExponentialCase test;
  RKEmbeddedParallel<BTDOPRI5>::integrate(test.x0, test.t0, test.tf, test.atol, test.rtol,
                                          test.ode, test.output, executor);
  const auto& t = test.output.times();
  const auto& s = test.output.states();
  EXPECT_EQ(45u, t.size());
  EXPECT_DOUBLE_EQ(10.0, t.back());
  EXPECT_NEAR(4.9896555947535841, t[t.size() / 2], 1e-10);
  EXPECT_NEAR(110132.46804595254, s.back()[5], 110132.0 * 1e-12);
  EXPECT_NEAR(198238.44248271457, s.back()[9], 198238.0 * 1e-12);
  std::printf("ok DOPRI5_Exponential\n");
}

static void rkf78_nbody() {
  // SingleThreadedRKEmbeddedTest.cpp:306-340 goldens with the tolerances of
  // ThreadPoolRKEmbeddedTest.cpp:314-351
  HipExecutor executor{};
  NBodyCase test;
  RKEmbeddedParallel<BTRKF78>::integrate(test.x0, test.t0, test.tf, test.atol, test.rtol,
                                         test.ode, test.output, executor);
  const auto& t = test.output.times();
  const auto& s = test.output.states();
  const double midpoint_tol = 1.657666 * test.tol * t.size() / 2;
  EXPECT_EQ(172u, t.size());
  EXPECT_NEAR(3.1699354066499978, t[t.size() / 2], midpoint_tol);
  EXPECT_DOUBLE_EQ(6.3, t.back());
  EXPECT_DOUBLE_EQ(1.657666, s.front()[0]);
  EXPECT_NEAR(-1.657049151812777, s[s.size() / 2][0], midpoint_tol);
  EXPECT_NEAR(1.6573213121279662, s.back()[0], test.tol);
  EXPECT_NEAR(0.46802604979077689, s.back()[3], test.tol);
  EXPECT_NEAR(-1.2481480178176445, s.back()[6], test.tol);
  EXPECT_NEAR(-1.2883624086182086, s.back()[9], test.tol);
  EXPECT_NEAR(0.41116306451711232, s.back()[12], test.tol);
  std::printf("ok RKF78_NBody\n");
}

static void rkf78_spinning_cube() {
  // SingleThreadedRKEmbeddedTest.cpp:342-382 goldens; tolerance 10 eps as
  // ThreadPoolRKEmbeddedTest.cpp:362
  HipExecutor executor{};
  SpinningCubeCase test;
  RKEmbeddedParallel<BTRKF78>::integrate(*test.x0, test.t0, test.tf, *test.atol, *test.rtol,
                                         test.ode, test.output, executor);
  const auto& t = test.output.times();
  const auto& s = test.output.states();
  const double rounding_error = 10.0 * 2.220446049250313e-16;
  constexpr int n = SpinningCubeCase::N;
  EXPECT_EQ(4u, t.size());
  EXPECT_NEAR(0.0099564200160493655, t[t.size() / 2], rounding_error);
  EXPECT_DOUBLE_EQ(0.03125, t.back());
  EXPECT_DOUBLE_EQ(0.59284461651668263, s.front()[0]);
  EXPECT_DOUBLE_EQ(0.73169375852499918, s.front()[n / 4]);
  EXPECT_NEAR(0.16121502269618729, s.back()[0], rounding_error);
  EXPECT_NEAR(0.71404096847054921, s.back()[n / 4], rounding_error);
  EXPECT_NEAR(0.59500181799963803, s.back()[n / 2], rounding_error);
  EXPECT_NEAR(0.42875743462583493, s.back()[3 * n / 4], rounding_error);
  std::printf("ok RKF78_SpinningCube1024\n");
}

static void device_resident_van_der_pol() {
  // HE21 has 183011 accepted steps (SingleThreadedRKEmbeddedTest.cpp:28-49):
  // run it, and DOPRI5, with the controller on the device.
  HipExecutor executor{};
  {
    VanDerPolCase test;
    auto x = test.x0;
    std::vector<double> times;
    auto st = RKEmbeddedParallel<BTHE21>::integrate_on_device(x, test.t0, test.tf, test.atol,
                                                              test.rtol, test.ode, executor,
                                                              &times);
    std::array<double, 2> last{};
    x.copy_to_span(last);
    EXPECT_EQ(183012u, times.size());
    EXPECT_EQ(183011, st.accepted);
    EXPECT_NEAR(0.9274958393895415, times[times.size() / 2], 1e-12);
    EXPECT_NEAR(0.32331666707211976, last[0], 1e-12);
    EXPECT_NEAR(-1.8329745679629066, last[1], 1e-12);
  }
  {
    VanDerPolCase test;
    auto x = test.x0;
    std::vector<double> times;
    RKEmbeddedParallel<BTDOPRI5>::integrate_on_device(x, test.t0, test.tf, test.atol,
                                                      test.rtol, test.ode, executor, &times);
    std::array<double, 2> last{};
    x.copy_to_span(last);
    EXPECT_EQ(76u, times.size());
    EXPECT_DOUBLE_EQ(0.67149697578171041, times[times.size() / 2]);
    EXPECT_DOUBLE_EQ(0.32331666704309109, last[0]);
    EXPECT_DOUBLE_EQ(-1.8329745679718159, last[1]);
  }
  std::printf("ok device_resident_VanDerPol\n");
}

// The DenseOutput policy on the device DOP853 solve: the reference's 20
// accumulated x values (StepperDopr853_test.cpp:181-196, all but the last,
// which lies at the end and is never written) and, on x' = x, values that
// track the closed form between steps.
static void dense_output_policy() {
  HipExecutor executor{};
  constexpr std::size_t n = 3;
  const std::array<double, n> init{1.0, 2.0, -0.5};
  HipState<double, n> x{std::span<const double, n>{init}};
  HipState<double, n> tol{};
  fill(executor, tol, 1e-10);
  DenseOutput<n> out{};
  auto st = RKEmbeddedParallel<BTDOP853>::integrate_on_device(
      x, 0.0, 2.0, tol, tol, ExponentialODE<double, n>{}, out, executor);
  EXPECT_EQ(ODEHIP_STATUS_REACHED_END, st.status);
  EXPECT_EQ(20u, out.n_steps());
  const double golden_x[] = {0.0, 0.1, 0.2, 0.30000000000000004, 0.4, 0.5, 0.6, 0.7,
                             0.7999999999999999, 0.8999999999999999, 0.9999999999999999,
                             1.0999999999999999, 1.2, 1.3, 1.4000000000000001,
                             1.5000000000000002, 1.6000000000000003, 1.7000000000000004,
                             1.8000000000000005, 1.9000000000000006};
  for (std::size_t k = 0; k < out.n_steps() && k < 20; ++k) {
    EXPECT_EQ(golden_x[k], out.x_values()[k]);
    EXPECT_NEAR(2.0 * std::exp(golden_x[k]), out.y_values()[1][k], 1e-8);
  }
  EXPECT_EQ(-0.5, out.y_values()[2][0]);
  std::printf("ok dense_output_policy\n");
}

static void state_semantics() {
  // CudaStateTest.cu:18-92
  constexpr std::size_t N = 10;
  std::array<double, N> data{};
  std::iota(data.begin(), data.end(), 0.0);
  HipState<double, N> zero{};
  std::array<double, N> out{};
  zero.copy_to_span(out);
  for (auto v : out) EXPECT_DOUBLE_EQ(0.0, v);
  HipState<double, N> a{data};
  a.copy_to_span(out);
  for (std::size_t i = 0; i < N; ++i) EXPECT_DOUBLE_EQ(data[i], out[i]);
  auto b = a;  // deep copy
  HipExecutor executor{};
  fill(executor, a, 3.0);
  b.copy_to_span(out);
  for (std::size_t i = 0; i < N; ++i) EXPECT_DOUBLE_EQ(data[i], out[i]);
  a.copy_to_span(out);
  for (auto v : out) EXPECT_DOUBLE_EQ(3.0, v);
  // CudaExecutorTest.cu:52-61 inner product of iota and iota+1 = 330, as the
  // square of a scaled RMS norm
  std::array<double, N> v{}, one{};
  for (std::size_t i = 0; i < N; ++i) {
    v[i] = std::sqrt(static_cast<double>(i) * (i + 1));
    one[i] = 1.0;
  }
  HipState<double, N> vs{v}, ones{one};
  double norm = 0.0;
  executor.check(odehip_rk_norm(executor.context(), vs.data(), ones.data(), N, &norm), "norm");
  EXPECT_NEAR(330.0, norm * norm * N, 1e-10);
  std::printf("ok state_semantics\n");
}

// Leapfrog through the mirror: one saved state per step (t0 included), the
// last equal to a run without output, energy kept (two-body orbit).
static void leapfrog_two_body() {
  HipExecutor exe{};
  constexpr std::size_t N = 12;
  // masses 1 and 1e-3 (device), circular-ish orbit
  HipState<double, 2> m{std::array{1.0, 1.0e-3}};
  const double v = std::sqrt(1.001);
  HipState<double, N> x0{std::array{0.0, 0.0, 0.0, 1.0, 0.0, 0.0,
                                    0.0, 0.0, 0.0, 0.0, v, 0.0}};
  NBodyODE<double, N> ode{0.0, m.data()};
  RawOutput<N> out{};
  ParallelLeapfrogIntegrator::integrate(x0, 0.0, 1.0, 0.01, ode, out, exe);
  EXPECT_EQ(out.times().size(), 101u);
  EXPECT_EQ(out.times()[0], 0.0);
  const auto& last = out.states().back();
  auto energy = [](const std::array<double, N>& s) {
    const double dx = s[3] - s[0], dy = s[4] - s[1], dz = s[5] - s[2];
    const double ke = 0.5 * (s[6] * s[6] + s[7] * s[7] + s[8] * s[8]) +
                      0.5e-3 * (s[9] * s[9] + s[10] * s[10] + s[11] * s[11]);
    return ke - 1.0e-3 / std::sqrt(dx * dx + dy * dy + dz * dz);
  };
  EXPECT_NEAR(energy(last), energy(out.states()[0]), 1e-7);
}

// The ensemble mirror: two systems, each equal to integrate_on_device alone.
static void ensemble_matches_single() {
  HipExecutor exe{};
  constexpr std::size_t N1 = 6 * 7, N2 = 6 * 40;
  std::mt19937_64 gen(5);
  std::uniform_real_distribution<double> u(-1.0, 1.0);
  std::array<double, N1> a{};
  std::array<double, N2> b{};
  for (auto& v : a) v = 0.1 * u(gen);
  for (auto& v : b) v = 0.1 * u(gen);
  for (std::size_t i = 0; i < N1 / 2; ++i) a[i] *= 10.0;
  for (std::size_t i = 0; i < N2 / 2; ++i) b[i] *= 10.0;
  std::array<double, N1 + N2> packed{};
  std::copy(a.begin(), a.end(), packed.begin());
  std::copy(b.begin(), b.end(), packed.begin() + N1);
  HipState<double, N1 + N2> x0{packed}, xo{};
  odehip_ensemble_options opt{};
  opt.t0 = 0.0;
  opt.tf = 0.2;
  opt.atol = opt.rtol = 1e-10;
  opt.tile_config = ODEHIP_ENS_TILE_AUTO;
  auto r = solve_ensemble<BTDOPRI5>(exe, {7, 40}, x0.data(), xo.data(), opt, 1e-3);
  EXPECT_EQ(r.totals.n_unfinished, 0);
  std::array<double, N1 + N2> got{};
  xo.copy_to_span(std::span<double, N1 + N2>{got});
  HipState<double, N1> s1{a}, t1{std::array<double, N1>{}};
  fill(exe, t1, 1e-10);
  HipState<double, N2> s2{b}, t2{std::array<double, N2>{}};
  fill(exe, t2, 1e-10);
  auto st1 = RKEmbeddedParallel<BTDOPRI5>::integrate_on_device(
      s1, 0.0, 0.2, t1, t1, NBodyODE<double, N1>{1e-3}, exe);
  auto st2 = RKEmbeddedParallel<BTDOPRI5>::integrate_on_device(
      s2, 0.0, 0.2, t2, t2, NBodyODE<double, N2>{1e-3}, exe);
  EXPECT_EQ(r.systems[0].accepted, st1.accepted);
  EXPECT_EQ(r.systems[1].rejected, st2.rejected);
  std::array<double, N1> o1{};
  std::array<double, N2> o2{};
  s1.copy_to_span(std::span<double, N1>{o1});
  s2.copy_to_span(std::span<double, N2>{o2});
  bool same = std::equal(o1.begin(), o1.end(), got.begin()) &&
              std::equal(o2.begin(), o2.end(), got.begin() + N1);
  EXPECT_EQ(same, true);
}

int main() {
  state_semantics();
  van_der_pol_consistency<BTDOPRI5>({"DOPRI5_VanDerPol", 76, 0.67149697578171041,
                                     1.7382670271053748, -0.62280413467104634,
                                     0.32331666704309109, -1.8329745679718159});
  van_der_pol_consistency<BTRKF45>({"RKF45_VanDerPol", 83, 0.65364467640860291,
                                    1.7493089062893852, -0.61420762201954171,
                                    0.32331666497810646, -1.8329745707118268});
  van_der_pol_consistency<BTDVERK>({"DVERK_VanDerPol", 42, 0.66331345368999683,
                                    1.7433476890385384, -0.61887401617955773,
                                    0.32331666651502594, -1.8329745684699728});
  van_der_pol_consistency<BTRKF78>({"RKF78_VanDerPol", 15, 0.72814921744061389,
                                    1.7022210141241134, -0.64964031512382325,
                                    0.3233166666554449, -1.8329745686289092});
  dopri5_exponential();
  user_tableau_struct();
  rkf78_nbody();
  rkf78_spinning_cube();
  device_resident_van_der_pol();
  leapfrog_two_body();
  ensemble_matches_single();
  dense_output_policy();
  std::printf("%d checks, %d failures\n", g_checks, g_failures);
  return g_failures == 0 ? 0 : 1;
}
