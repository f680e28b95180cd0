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

// Whole-array tests over a thread's strided share (elements first, first +
// stride, ... < lim), with L loads issued per round trip: a loop that waits
// for every load would take one memory latency per element.  Each load in
// flight holds VGPRs for its value and address, and the kernel's VGPR count
// is its largest anywhere, so kernels whose occupancy is set by VGPRs (the
// reorderable variants) scan kScanLoadsLean at a time; kernels held to a few
// waves per SIMD by their LDS or main loop anyway scan kScanLoads.
constexpr int kScanLoads = 16;
constexpr int kScanLoadsLean = 4;

// Indices past lim load element e instead (in range: the loop runs while
// e < lim) and are left out of the test, so every load is unconditional.

// Every element passes coord_shared_ok.
// I: the index type (int where every index fits, which saves VGPRs); the
// arguments convert to it rather than choose it.
template <typename T>
struct ScanIndex {
  using type = T;
};
template <int L = kScanLoads, typename I = long>
__device__ inline bool strided_all_shared_ok(const double* __restrict__ v,
                                             typename ScanIndex<I>::type lim,
                                             typename ScanIndex<I>::type first,
                                             typename ScanIndex<I>::type stride) {
  bool ok = true;
  for (I e = first; e < lim; e += L * stride) {
    double c[L];
#pragma unroll
    for (int u = 0; u < L; ++u) {
      const I x = e + u * stride;
      c[u] = v[x < lim ? x : e];
    }
#pragma unroll
    for (int u = 0; u < L; ++u)
      ok = ok & ((e + u * stride >= lim) | coord_shared_ok(c[u]));
  }
  return ok;
}

// Every element has the bit pattern b.
template <int L = kScanLoads>
__device__ inline bool strided_all_bits(const double* __restrict__ v, long lim, long first,
                                        long stride, uint64_t b) {
  bool ok = true;
  for (long e = first; e < lim; e += L * stride) {
    uint64_t c[L];
#pragma unroll
    for (int u = 0; u < L; ++u) {
      const long x = e + u * stride;
      c[u] = __builtin_bit_cast(uint64_t, v[x < lim ? x : e]);
    }
#pragma unroll
    for (int u = 0; u < L; ++u) ok = ok & ((e + u * stride >= lim) | (c[u] == b));
  }
  return ok;
}

// The reorderable variants (fast, mixed) with all n masses equal: unit-mass
// terms, each finished sum scaled by the common mass once (one multiply per
// pair less; not the per-term rounding, which these variants do not promise).
// Whether this lane's strided share of the masses all have mass[0]'s bits
// and mass[0] is finite and nonzero; callers combine the lanes (__all or a
// workgroup vote).
template <int L = kScanLoads>
__device__ inline bool lane_masses_uniform(const double* __restrict__ mass, int n, int first,
                                           int stride) {
  const double m0 = mass[0];
  const bool ok = __builtin_isfinite(m0) && m0 != 0.0;
  return ok & strided_all_bits<L>(mass, n, first, stride, __builtin_bit_cast(uint64_t, m0));
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
  // m r^-3 from the hardware rsq y0 with e = 1 - r2 y0^2: r2^(-3/2) =
  // y0^3 (1 - e)^(-3/2) = y0^3 (1 + 3e/2 (1 + 5e/4) + O(e^3)), the correction
  // applied to m y0^3 directly.  Each instruction takes at most one constant
  // that is not an inline one (the scalar-operand limit), and the longest
  // dependency chain from y0 is five instructions.
  const double y0 = __builtin_amdgcn_rsq(r2);
  const double y2 = y0 * y0;
  const double e = __builtin_fma(-r2, y2, 1.0);
  const double y3 = y2 * (kMass ? y0 * mj : y0);
  const double c = e * __builtin_fma(e, 1.25, 1.0);
  double w = __builtin_fma(y3 * c, 1.5, y3);
  if (kGuard) w = (j == i) ? 0.0 : w;
  ax = __builtin_fma(dx, w, ax);
  ay = __builtin_fma(dy, w, ay);
  az = __builtin_fma(dz, w, az);
}

// All-FP32 mixed-precision pair term (the ensemble kernels' form): the
// displacement is formed in FP64 and rounded to FP32; r^-3 (hardware rsqrt)
// and the term are FP32, summed into FP32 partials.
template <bool kMass, bool kGuard>
__device__ inline void mixed_term_fp32(double xj, double yj, double zj, float mj,
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

// Mixed-precision pair term.  The displacement is formed in FP64 and rounded
// to FP32; r^-3 (hardware rsqrt) and the term are FP32 and go into the FP32
// partial sums (fx, fy, fz).  A near pair -- FP32 r^2 below near2 -- is left
// out here and the function returns false (true: a far pair, done); the
// caller then adds its FP64 term with mixed_near_term.  In a close encounter one partner dominates a body's
// acceleration, and FP32's ~1e-7 relative error on that one term would set
// the accuracy of the whole solve.  near2 = 0 gives the all-FP32 form.
template <bool kMass, bool kGuard>
__device__ inline bool mixed_term(double xj, double yj, double zj, double mj,
                                  int j, int i, double xi, double yi,
                                  double zi, float eps2f, float near2,
                                  float& fx, float& fy, float& fz) {
  const float dx = static_cast<float>(xj - xi);
  const float dy = static_cast<float>(yj - yi);
  const float dz = static_cast<float>(zj - zi);
  const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, eps2f)));
  const float rinv = __builtin_amdgcn_rsqf(r2);
  float w = (rinv * rinv) * rinv;
  if (kMass) w *= static_cast<float>(mj);
  if (kGuard) w = (j == i) ? 0.0f : w;
  const bool far = !(r2 < near2);
  w = far ? w : 0.0f;
  fx = __builtin_fmaf(dx, w, fx);
  fy = __builtin_fmaf(dy, w, fy);
  fz = __builtin_fmaf(dz, w, fz);
  return far;
}

// Mask of the active lanes of the wave for which c is false (from the
// compare result kept in a scalar register pair; OR-ing such masks runs on
// the scalar unit).
__device__ inline uint64_t ballot_not(bool c) {
  return ~__builtin_amdgcn_ballot_w64(c) & __builtin_amdgcn_ballot_w64(true);
}

// Hides a value's origin from the compiler, so that expressions of it are
// not merged with the same expressions elsewhere (no code is emitted).  The
// near-pair pass recomputes displacements instead of keeping every partner's
// FP64 displacement of the far-pair pass alive across it.
__device__ inline double opaque(double v) {
  asm volatile("" : "+v"(v));
  return v;
}
// The same for the FP32 partial sums just before the near-pair branch: they
// are complete there, so the far-pair pass is not sunk past the branch (which
// would keep every partner's displacement and r^2 live across it).
__device__ inline float opaque(float v) {
  asm volatile("" : "+v"(v));
  return v;
}

// The FP64 term of a pair that mixed_term left out (the same FP32 test picks
// it), added into the FP64 sums.  Callers batch partners and enter this only
// when some lane has a near pair among them, so the loop over far pairs stays
// one straight-line block; they pass the target through opaque().
template <bool kMass, bool kGuard>
__device__ inline void mixed_near_term(double xj, double yj, double zj,
                                       double mj, int j, int i, double xi,
                                       double yi, double zi, float eps2f,
                                       double eps2, float near2, double& ax,
                                       double& ay, double& az) {
  const float dx = static_cast<float>(xj - xi);
  const float dy = static_cast<float>(yj - yi);
  const float dz = static_cast<float>(zj - zi);
  const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, eps2f)));
  if (r2 < near2)
    fast_term<kMass, kGuard>(xj, yj, zj, mj, j, i, xi, yi, zi, eps2, ax, ay, az);
}

}  // namespace odehip
