// Oracle for the learned step-size controller's network (test
// infrastructure only: tests/ call it through oracle_capi.cpp as the checker
// of the HIP training and inference kernels; nothing in the product links it).
//
// Network: 10 -> 64 -> 64 -> 1, tanh hidden layers, linear output; inputs
// are the raw controller features normalised by a stored mean / std.  The
// reference has no learned controller (its controller is
// RKEmbeddedParallel.hpp:64-79); this restates the network the library
// defines, in double precision, with a backward pass written out by hand.
// Its gradient is pinned against central finite differences in the tests.
//
// Parameter layout (one flat array of kParams):
//   W1[64][10] | b1[64] | W2[64][64] | b2[64] | w3[64] | b3
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace oracle {

constexpr int kIn = 10;
constexpr int kHid = 64;
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kHid * kIn;
constexpr int kOffW2 = kOffB1 + kHid;
constexpr int kOffB2 = kOffW2 + kHid * kHid;
constexpr int kOffW3 = kOffB2 + kHid;
constexpr int kOffB3 = kOffW3 + kHid;
constexpr int kParams = kOffB3 + 1;

// Round a double to the nearest bfloat16 value (ties to even), returned as
// a double.  (float first: the bf16 value is the float's top 16 bits.)
inline double to_bf16(double v) {
  float f = static_cast<float>(v);
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) != 0x7f800000u) {  // finite
    const uint32_t lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;
  }
  u &= 0xffff0000u;
  std::memcpy(&f, &u, 4);
  return static_cast<double>(f);
}

struct MlpForward {
  std::vector<double> x, z1, h1, z2, h2;  // per sample
  double y = 0.0;
};

// Forward pass of one sample.  bf16: weights and every layer's input are
// rounded to bfloat16 (products then summed exactly in double), the way the
// inference kernel feeds its matrix cores.
inline double mlp_forward_one(const double* p, const double* mean,
                              const double* stdv, const double* raw, bool bf16,
                              MlpForward* keep = nullptr) {
  auto R = [&](double v) { return bf16 ? to_bf16(v) : v; };
  double x[kIn], h1[kHid], h2[kHid], z1[kHid], z2[kHid];
  for (int i = 0; i < kIn; ++i) x[i] = R((raw[i] - mean[i]) / stdv[i]);
  for (int j = 0; j < kHid; ++j) {
    double s = 0.0;
    for (int i = 0; i < kIn; ++i) s += R(p[kOffW1 + j * kIn + i]) * x[i];
    z1[j] = s + p[kOffB1 + j];
    h1[j] = std::tanh(z1[j]);
  }
  for (int j = 0; j < kHid; ++j) {
    double s = 0.0;
    for (int i = 0; i < kHid; ++i) s += R(p[kOffW2 + j * kHid + i]) * R(h1[i]);
    z2[j] = s + p[kOffB2 + j];
    h2[j] = std::tanh(z2[j]);
  }
  double y = 0.0;
  for (int j = 0; j < kHid; ++j) y += R(p[kOffW3 + j]) * R(h2[j]);
  y += p[kOffB3];
  if (keep) {
    keep->x.assign(x, x + kIn);
    keep->z1.assign(z1, z1 + kHid);
    keep->h1.assign(h1, h1 + kHid);
    keep->z2.assign(z2, z2 + kHid);
    keep->h2.assign(h2, h2 + kHid);
    keep->y = y;
  }
  return y;
}

// Loss L = (1/n) sum_s 0.5 (y_s - t_s)^2 and dL/dp (double, no rounding).
inline double mlp_loss_grad(const double* p, const double* mean,
                            const double* stdv, const double* raw,
                            const double* target, long n, double* grad) {
  std::memset(grad, 0, sizeof(double) * kParams);
  double loss = 0.0;
  MlpForward f;
  for (long s = 0; s < n; ++s) {
    const double y = mlp_forward_one(p, mean, stdv, raw + s * kIn, false, &f);
    const double r = y - target[s];
    loss += 0.5 * r * r;
    const double dy = r / static_cast<double>(n);
    grad[kOffB3] += dy;
    double dz2[kHid];
    for (int j = 0; j < kHid; ++j) {
      grad[kOffW3 + j] += dy * f.h2[j];
      dz2[j] = dy * p[kOffW3 + j] * (1.0 - f.h2[j] * f.h2[j]);
      grad[kOffB2 + j] += dz2[j];
      for (int i = 0; i < kHid; ++i) grad[kOffW2 + j * kHid + i] += dz2[j] * f.h1[i];
    }
    for (int i = 0; i < kHid; ++i) {
      double dh1 = 0.0;
      for (int j = 0; j < kHid; ++j) dh1 += dz2[j] * p[kOffW2 + j * kHid + i];
      const double dz1 = dh1 * (1.0 - f.h1[i] * f.h1[i]);
      grad[kOffB1 + i] += dz1;
      for (int k = 0; k < kIn; ++k) grad[kOffW1 + i * kIn + k] += dz1 * f.x[k];
    }
  }
  return loss / static_cast<double>(n);
}

}  // namespace oracle
