// Pair terms of the N-body right-hand side (device functions shared by the
// single-system and the ensemble kernels).  See nbody_kernels.hip for the
// three arithmetic variants.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace odehip {

// One pair term exactly as the reference writes it (NBodyODE.hpp:89-98):
// IEEE correctly rounded sqrt and divisions via the compiler's expansions.
template <bool kMass>
__device__ inline void exact_term_ieee(double dx, double dy, double dz,
                                       double mj, double softening_sq,
                                       double& ax, double& ay, double& az) {
  const double dist_sq = dx * dx + dy * dy + dz * dz + softening_sq;
  const double dist = __builtin_sqrt(dist_sq);
  const double denominator = dist * dist_sq;
  ax = dx / denominator;
  ay = dy / denominator;
  az = dz / denominator;
  if (kMass) {
    ax *= mj;
    ay *= mj;
    az *= mj;
  }
}

// The same term, same bits, cheaper: valid when the three displacement
// magnitudes lie in [2^-300, 2^100] and softening_sq in [2^-200, 2^100]
// (then dist_sq in [2^-200, 2^202] and the denominator in [2^-300, 2^304]);
// a zero displacement component gives a +-0 term (see coord_shared_ok).
//  * sqrt: the compiler's correctly rounded sqrt is a rescale around a
//    Goldschmidt refinement of v_rsq_f64 (y = rsq(s); g = s y; h = y/2;
//    r = 1/2 - h g; g += g r; h += h r; twice g += (s - g^2) h); the rescale
//    is the identity for s >= 2^-767, so the refinement alone gives the same
//    bits.
//  * division: the hardware sequence for n/d is d' = div_scale(d), r =
//    rcp(d') refined by two Newton steps, q0 = n' r, rem = n' - d' q0,
//    q = div_fmas(rem, r, q0), div_fixup.  With n and d in the range above
//    div_scale leaves both operands unscaled (no VCC), div_fmas is a plain
//    fma and div_fixup returns q unchanged -- so the refined reciprocal
//    depends on d only and the three divisions by the same denominator share
//    it: 1 rcp + 4 fma, then 3 ops per numerator instead of 11.
template <bool kMass>
__device__ inline void exact_term_shared_rcp(double dx, double dy, double dz,
                                             double mj, double softening_sq,
                                             double& ax, double& ay,
                                             double& az) {
  const double dist_sq = dx * dx + dy * dy + dz * dz + softening_sq;
  const double y = __builtin_amdgcn_rsq(dist_sq);
  double g = dist_sq * y;
  double h = y * 0.5;
  const double r0 = __builtin_fma(-h, g, 0.5);
  g = __builtin_fma(g, r0, g);
  double d = __builtin_fma(-g, g, dist_sq);
  h = __builtin_fma(h, r0, h);
  g = __builtin_fma(d, h, g);
  d = __builtin_fma(-g, g, dist_sq);
  const double dist = __builtin_fma(d, h, g);
  const double den = dist * dist_sq;
  double r = __builtin_amdgcn_rcp(den);
  double e = __builtin_fma(-den, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-den, r, 1.0);
  r = __builtin_fma(r, e, r);
  const double qx = dx * r, qy = dy * r, qz = dz * r;
  ax = __builtin_fma(__builtin_fma(-den, qx, dx), r, qx);
  ay = __builtin_fma(__builtin_fma(-den, qy, dy), r, qy);
  az = __builtin_fma(__builtin_fma(-den, qz, dz), r, qz);
  if (kMass) {
    ax *= mj;
    ay *= mj;
    az *= mj;
  }
}

// Coordinates for which the shared-reciprocal term, summed into a target's
// acceleration, gives the same bits as the IEEE expressions with no per-pair
// test: the coordinate is 0 or its magnitude is in [2^-246, 2^98].  If every
// partner and target coordinate satisfies this, |dx| <= 2^99 and a nonzero
// displacement component has magnitude >= 2^-298 (a multiple of the smaller
// operand's ulp), inside the shared form's range; a zero component gives a
// +-0 term in both forms, and adding +-0 to a sum that starts at +0 (and so
// is never -0) leaves it unchanged.  Pure integer tests on the bit pattern:
// on wave-uniform partner coordinates they run on the scalar unit.
__device__ inline bool coord_shared_ok(double c) {
  constexpr uint64_t kLo = uint64_t{1023 - 246} << 52;  // bits of 2^-246
  constexpr uint64_t kHi = uint64_t{1023 + 98} << 52;   // bits of 2^98
  const uint64_t b = __builtin_bit_cast(uint64_t, c) & 0x7fffffffffffffffull;
  return (b == 0) | (b - kLo <= kHi - kLo);
}

__device__ inline bool point_shared_ok(double x, double y, double z) {
  return coord_shared_ok(x) & coord_shared_ok(y) & coord_shared_ok(z);
}

template <bool kMass, bool kGuard>
__device__ inline void fast_term(double xj, double yj, double zj, double mj,
                                 int j, int i, double xi, double yi, double zi,
                                 double eps2, double& ax, double& ay,
                                 double& az) {
  const double dx = xj - xi;
  const double dy = yj - yi;
  const double dz = zj - zi;
  const double r2 = __builtin_fma(dx, dx, __builtin_fma(dy, dy, __builtin_fma(dz, dz, eps2)));
  const double y0 = __builtin_amdgcn_rsq(r2);
  const double h = r2 * y0;
  const double e = __builtin_fma(-h, y0, 1.0);
  const double y1 = __builtin_fma(y0 * e, __builtin_fma(e, 0.375, 0.5), y0);
  double w = (y1 * y1) * y1;
  if (kMass) w *= mj;
  if (kGuard) w = (j == i) ? 0.0 : w;
  ax = __builtin_fma(dx, w, ax);
  ay = __builtin_fma(dy, w, ay);
  az = __builtin_fma(dz, w, az);
}

template <bool kMass, bool kGuard>
__device__ inline void mixed_term(double xj, double yj, double zj, float mj,
                                  int j, int i, double xi, double yi,
                                  double zi, float eps2f, float& ax, float& ay,
                                  float& az) {
  const float dx = static_cast<float>(xj - xi);
  const float dy = static_cast<float>(yj - yi);
  const float dz = static_cast<float>(zj - zi);
  const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, eps2f)));
  const float rinv = __builtin_amdgcn_rsqf(r2);
  float w = (rinv * rinv) * rinv;
  if (kMass) w *= mj;
  if (kGuard) w = (j == i) ? 0.0f : w;
  ax = __builtin_fmaf(dx, w, ax);
  ay = __builtin_fmaf(dy, w, ay);
  az = __builtin_fmaf(dz, w, az);
}

}  // namespace odehip
