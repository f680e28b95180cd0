// pow(x, y) for finite x > 0 evaluated in double-double arithmetic (about
// 100 significant bits) and rounded once, so the result is the correctly
// rounded x^y except when x^y lies within ~2^-100 (relative) of a rounding
// boundary.  The step-size controller needs pow(err, 1/(1+q)) each attempt
// (RKEmbeddedParallel.hpp:24, :65, :198); a correctly rounded value is
// platform independent, so the device controller and any host controller
// choose identical step sizes.
//
// Method: x = 2^e m with m in [sqrt(1/2), sqrt(2)); log m = 2 atanh(u),
// u = (m-1)/(m+1), by its odd power series (|u| <= 0.1716, 22 terms);
// z = y log x; exp z = 2^k exp(r) with r = z - k ln2, exp(r) from the Taylor
// series of exp(r/256) - 1 followed by eight squarings in the expm1 form
// s <- 2s + s^2.  Only +, -, *, fma and exact scaling are used, so the host
// and device builds produce the same bits.  Callers must compile with FMA
// contraction off (explicit fma is still fused).
#pragma once

#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#define ODEHIP_HD __host__ __device__
#else
#define ODEHIP_HD
#endif

namespace odehip {
namespace crpow_detail {

struct dd {
  double hi, lo;
};

ODEHIP_HD inline uint64_t to_bits(double x) { return __builtin_bit_cast(uint64_t, x); }
ODEHIP_HD inline double from_bits(uint64_t b) { return __builtin_bit_cast(double, b); }

ODEHIP_HD inline dd two_sum(double a, double b) {
  double s = a + b;
  double bb = s - a;
  double e = (a - (s - bb)) + (b - bb);
  return {s, e};
}
ODEHIP_HD inline dd quick_two_sum(double a, double b) {
  double s = a + b;
  double e = b - (s - a);
  return {s, e};
}
ODEHIP_HD inline dd two_prod(double a, double b) {
  double p = a * b;
  double e = std::fma(a, b, -p);
  return {p, e};
}
ODEHIP_HD inline dd add(dd a, dd b) {
  dd s = two_sum(a.hi, b.hi);
  dd t = two_sum(a.lo, b.lo);
  s.lo += t.hi;
  s = quick_two_sum(s.hi, s.lo);
  s.lo += t.lo;
  return quick_two_sum(s.hi, s.lo);
}
ODEHIP_HD inline dd mul(dd a, dd b) {
  dd p = two_prod(a.hi, b.hi);
  p.lo += a.hi * b.lo + a.lo * b.hi;
  return quick_two_sum(p.hi, p.lo);
}
ODEHIP_HD inline dd mul_d(dd a, double b) {
  dd p = two_prod(a.hi, b);
  p.lo += a.lo * b;
  return quick_two_sum(p.hi, p.lo);
}
ODEHIP_HD inline dd div(dd a, dd b) {
  double q1 = a.hi / b.hi;
  dd r = add(a, mul_d(b, -q1));
  double q2 = r.hi / b.hi;
  r = add(r, mul_d(b, -q2));
  double q3 = r.hi / b.hi;
  dd q = quick_two_sum(q1, q2);
  return add(q, dd{q3, 0.0});
}
ODEHIP_HD inline dd from_ratio(double num, double den) {
  // num/den to double-double precision (num, den exact small integers).
  double hi = num / den;
  double lo = std::fma(-hi, den, num) / den;
  return quick_two_sum(hi, lo);
}

// Exact power-of-two scaling through the exponent field.
ODEHIP_HD inline double scale2(double v, int k) {
  // k is small (|k| < 1000 here); split to stay in the normal range.
  while (k > 1000) {
    v *= 0x1p1000;
    k -= 1000;
  }
  while (k < -1000) {
    v *= 0x1p-1000;
    k += 1000;
  }
  return v * from_bits(static_cast<uint64_t>(k + 1023) << 52);
}

// ln 2 to ~107 bits.
constexpr double kLn2Hi = 6.93147180559945286227e-01;
constexpr double kLn2Lo = 2.31904681384629955842e-17;

ODEHIP_HD inline dd log_dd(double x) {
  // x finite, > 0, normal.
  const uint64_t bits = to_bits(x);
  int e = static_cast<int>((bits >> 52) & 0x7ff) - 1023;
  uint64_t mbits = (bits & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL;
  double m = from_bits(mbits);  // m in [1, 2)
  if (m > 1.4142135623730951) {
    m *= 0.5;
    e += 1;
  }
  // u = (m - 1) / (m + 1); m - 1 is exact (Sterbenz), m + 1 as double-double.
  dd num{m - 1.0, 0.0};
  dd den = two_sum(m, 1.0);
  dd u = div(num, den);
  dd w = mul(u, u);
  // sum_{k=0}^{21} w^k / (2k+1), Horner from the top.
  dd s = from_ratio(1.0, 43.0);
  for (int k = 20; k >= 0; --k) {
    s = add(mul(s, w), from_ratio(1.0, 2.0 * k + 1.0));
  }
  dd lm = mul(mul_d(u, 2.0), s);
  dd le = add(two_prod(static_cast<double>(e), kLn2Hi),
              dd{static_cast<double>(e) * kLn2Lo, 0.0});
  return add(le, lm);
}

ODEHIP_HD inline double exp_dd_round(dd z) {
  double kd = std::nearbyint(z.hi / kLn2Hi);
  int k = static_cast<int>(kd);
  // r = z - k ln2
  dd kl = add(two_prod(kd, kLn2Hi), two_prod(kd, kLn2Lo));
  dd r = add(z, dd{-kl.hi, -kl.lo});
  // r' = r / 256 (exact)
  dd rp{r.hi * 0.00390625, r.lo * 0.00390625};
  // s = exp(r') - 1 = r' + r'^2/2! + ... + r'^11/11!
  dd s = from_ratio(1.0, 39916800.0);  // 1/11!
  double fact[11] = {1.0,      2.0,       6.0,        24.0,        120.0,
                     720.0,    5040.0,    40320.0,    362880.0,    3628800.0,
                     39916800.0};
  for (int n = 10; n >= 1; --n) {
    s = add(mul(s, rp), from_ratio(1.0, fact[n - 1]));
  }
  s = mul(s, rp);
  for (int i = 0; i < 8; ++i) {
    // (1 + s)^2 - 1 = 2 s + s^2
    s = add(dd{2.0 * s.hi, 2.0 * s.lo}, mul(s, s));
  }
  dd v = add(dd{1.0, 0.0}, s);
  double hi = v.hi + v.lo;  // round the double-double once
  return scale2(hi, k);
}

}  // namespace crpow_detail

// Correctly rounded (see above) x^y for x > 0 finite normal, y finite, and a
// result in the normal range.  pow(0, y>0) = 0; other special inputs fall back
// to the IEEE conventions of the C library's pow.
ODEHIP_HD inline double cr_pow(double x, double y) {
  using namespace crpow_detail;
  if (x == 1.0 || y == 0.0) return 1.0;
  if (!(x > 0.0) || !std::isfinite(x) || !std::isfinite(y) ||
      x < 2.2250738585072014e-308) {
    return std::pow(x, y);
  }
  dd l = log_dd(x);
  dd z = mul_d(l, y);
  if (z.hi > 709.0 || z.hi < -708.0) return std::pow(x, y);
  return exp_dd_round(z);
}

}  // namespace odehip
