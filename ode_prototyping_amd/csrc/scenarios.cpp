// Host-side generators of the synthetic N-body systems the benchmarks and
// tests integrate.  The reference builds its scenarios with libstdc++'s
// std::uniform_real_distribution over std::mt19937
// (src/demo/n_body_scenarios/SpinningParticlesInBox.hpp:38-58); here the bit
// stream is std::mt19937_64 (fixed by the C++ standard) mapped to doubles
// explicitly (top 53 bits), and normals come from Box-Muller, so the
// generators do not depend on a library's distribution algorithms.
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <random>
#include <vector>

#include "odehip.h"

namespace {

constexpr double kTwoPi = 6.283185307179586476925286766559;

struct Rng {
  std::mt19937_64 g;
  explicit Rng(uint64_t seed) : g(seed) {}
  double u01() { return static_cast<double>(g() >> 11) * 0x1p-53; }  // [0, 1)
  double u01_open() { return 1.0 - u01(); }                          // (0, 1]
  bool have = false;
  double spare = 0.0;
  double normal() {
    if (have) {
      have = false;
      return spare;
    }
    double u1 = u01_open(), u2 = u01();
    double r = std::sqrt(-2.0 * std::log(u1));
    spare = r * std::sin(kTwoPi * u2);
    have = true;
    return r * std::cos(kTwoPi * u2);
  }
};

void remove_mean(double* x, size_t n, const double* m, int offset) {
  double tot = 0.0, s[3] = {0.0, 0.0, 0.0};
  for (size_t i = 0; i < n; ++i) {
    tot += m[i];
    for (int c = 0; c < 3; ++c) s[c] += m[i] * x[offset + 3 * i + c];
  }
  for (int c = 0; c < 3; ++c) s[c] /= tot;
  for (size_t i = 0; i < n; ++i)
    for (int c = 0; c < 3; ++c) x[offset + 3 * i + c] -= s[c];
}

// One Plummer sphere of n bodies, G = M = a = 1 (Aarseth, Henon & Wielen
// 1974): radius from the cumulative mass fraction (fractions above 0.999
// rejected), isotropic direction, speed q * v_esc with q drawn by rejection
// from g(q) = q^2 (1 - q^2)^(7/2).  Writes pos (3n) and vel (3n) strided into
// x at particle offset `first` of a system of `total` particles.
void plummer_unit(Rng& r, size_t n, double* x, size_t first, size_t total) {
  double* pos = x;
  double* vel = x + 3 * total;
  for (size_t i = first; i < first + n; ++i) {
    double X1;
    do {
      X1 = r.u01();
    } while (X1 <= 0.0 || X1 > 0.999);
    const double rad = 1.0 / std::sqrt(std::pow(X1, -2.0 / 3.0) - 1.0);
    const double X2 = r.u01(), X3 = r.u01();
    const double z = (1.0 - 2.0 * X2) * rad;
    const double rxy = std::sqrt(rad * rad - z * z);
    pos[3 * i] = rxy * std::cos(kTwoPi * X3);
    pos[3 * i + 1] = rxy * std::sin(kTwoPi * X3);
    pos[3 * i + 2] = z;
    double q;
    for (;;) {
      const double X4 = r.u01(), X5 = r.u01();
      if (0.1 * X5 < X4 * X4 * std::pow(1.0 - X4 * X4, 3.5)) {
        q = X4;
        break;
      }
    }
    const double ve = std::sqrt(2.0) * std::pow(1.0 + rad * rad, -0.25);
    const double v = q * ve;
    const double X6 = r.u01(), X7 = r.u01();
    const double w = (1.0 - 2.0 * X6) * v;
    const double vxy = std::sqrt(v * v - w * w);
    vel[3 * i] = vxy * std::cos(kTwoPi * X7);
    vel[3 * i + 1] = vxy * std::sin(kTwoPi * X7);
    vel[3 * i + 2] = w;
  }
}

}  // namespace

extern "C" {

int odehip_ic_uniform_cube(size_t n, uint64_t seed, double half_width,
                           double vel_sigma, double total_mass, double* x,
                           double* masses) {
  if (n == 0 || !x) return ODEHIP_ERR_INVALID;
  Rng r(seed);
  for (size_t i = 0; i < 3 * n; ++i) x[i] = -half_width + 2.0 * half_width * r.u01();
  for (size_t i = 0; i < 3 * n; ++i) x[3 * n + i] = vel_sigma * r.normal();
  std::vector<double> m(n, total_mass / static_cast<double>(n));
  remove_mean(x, n, m.data(), static_cast<int>(3 * n));
  if (masses)
    for (size_t i = 0; i < n; ++i) masses[i] = m[i];
  return ODEHIP_OK;
}

int odehip_ic_plummer(size_t n, uint64_t seed, double a, double total_mass,
                      double* x, double* masses) {
  if (n == 0 || !x || !(a > 0.0) || !(total_mass > 0.0)) return ODEHIP_ERR_INVALID;
  Rng r(seed);
  plummer_unit(r, n, x, 0, n);
  const double vs = std::sqrt(total_mass / a);
  for (size_t i = 0; i < 3 * n; ++i) {
    x[i] *= a;
    x[3 * n + i] *= vs;
  }
  std::vector<double> m(n, total_mass / static_cast<double>(n));
  remove_mean(x, n, m.data(), 0);
  remove_mean(x, n, m.data(), static_cast<int>(3 * n));
  if (masses)
    for (size_t i = 0; i < n; ++i) masses[i] = m[i];
  return ODEHIP_OK;
}

int odehip_ic_colliding_plummer(size_t n, uint64_t seed, double a,
                                double total_mass, double separation,
                                double rel_speed, double* x, double* masses) {
  if (n < 2 || n % 2 || !x || !(a > 0.0) || !(total_mass > 0.0))
    return ODEHIP_ERR_INVALID;
  const size_t h = n / 2;
  Rng r(seed);
  plummer_unit(r, h, x, 0, n);
  plummer_unit(r, h, x, h, n);
  const double msub = 0.5 * total_mass;
  const double vs = std::sqrt(msub / a);
  std::vector<double> m(n, total_mass / static_cast<double>(n));
  // centre each sphere, scale, then displace
  for (int half = 0; half < 2; ++half) {
    const size_t b = half * h;
    double cp[3] = {0, 0, 0}, cv[3] = {0, 0, 0};
    for (size_t i = b; i < b + h; ++i)
      for (int c = 0; c < 3; ++c) {
        cp[c] += x[3 * i + c];
        cv[c] += x[3 * n + 3 * i + c];
      }
    for (int c = 0; c < 3; ++c) {
      cp[c] /= static_cast<double>(h);
      cv[c] /= static_cast<double>(h);
    }
    const double sx = half == 0 ? 0.5 * separation : -0.5 * separation;
    const double svx = half == 0 ? -0.5 * rel_speed : 0.5 * rel_speed;
    for (size_t i = b; i < b + h; ++i) {
      for (int c = 0; c < 3; ++c) {
        x[3 * i + c] = (x[3 * i + c] - cp[c]) * a;
        x[3 * n + 3 * i + c] = (x[3 * n + 3 * i + c] - cv[c]) * vs;
      }
      x[3 * i] += sx;
      x[3 * n + 3 * i] += svx;
    }
  }
  if (masses)
    for (size_t i = 0; i < n; ++i) masses[i] = m[i];
  return ODEHIP_OK;
}

}  // extern "C"
