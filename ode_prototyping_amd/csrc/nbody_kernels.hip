// N-body gravity right-hand side on gfx950 (CDNA4): dxdt = [vel | acc(pos)]
// for the state layout of src/ode/NBodyODE.hpp:12-19 (positions then
// velocities, xyz interleaved per particle), G = 1, Plummer softening.
//
// The reference evaluates the N(N-1)/2 pairs in parallel and scatters six
// FP64 atomic adds per pair (NBodyODE.hpp:68-133).  Here every kernel is an
// atomic-free gather: a workgroup owns a tile of 64 target particles (one per
// lane) and streams the partner particles j; partner coordinates are
// wave-uniform, so they arrive through scalar loads into SGPRs and cost no
// VGPR bandwidth or LDS (the 32-target exact kernel, whose half-waves take
// different partners, uses vector loads instead).
//
//  * exact:  the reference's FP64 expression (((dx^2 + dy^2) + dz^2) + eps^2,
//            correctly rounded sqrt, den = dist * dist_sq, three correctly
//            rounded divisions) and, per target, the partners summed in
//            increasing j -- the order the single-threaded pair loop adds into
//            acc[i].  Terms are computed by all sixteen waves of a workgroup in
//            parallel and written to LDS; three waves then fold them in order
//            (one coordinate each), double-buffered so folding tile t-1
//            overlaps computing tile t.  Result: bit-identical to NBodyODE
//            with SingleThreadedExecutor.  Must be compiled with
//            -ffp-contract=off.
//  * fast:   FP64, r^-3 from the hardware reciprocal square root with a
//            third-order correction (a few ulp), FMA accumulation, the j range
//            split over 16 waves per workgroup and the partial sums combined
//            in a fixed order (deterministic, not bit-identical).
//  * mixed:  displacement computed in FP64 and rounded to FP32, r^-3 and the
//            pair force in FP32 (hardware rsqrt), per-wave partial sums in
//            FP32 flushed into FP64 accumulators every 32 partners; near pairs
//            (r^2 below a per-launch threshold) in FP64, summed in FP64.
#include "internal.hpp"
#include "nbody_terms.hpp"

namespace odehip {
namespace {

// Partner data is read at wave-uniform addresses and never written during a
// launch; reading it through the constant address space lets the compiler
// use scalar loads (SGPR results broadcast to all lanes for free).
typedef const __attribute__((address_space(4))) double* uniform_ptr;
__device__ inline uniform_ptr as_uniform(const double* p) {
  return (uniform_ptr)(p);
}

__device__ inline const double* select_input(const NBodyArgs& a) {
  if (a.y) return a.y;
  return a.st->cur ? a.xb1 : a.xb0;
}

// Returns true if the launch has nothing to compute (solve finished, or k1
// of this attempt is reused, or a dense-output evaluation whose pass does not
// run).  On FSAL reuse the launch copies k1 instead.
__device__ inline bool prologue_skip(const NBodyArgs& a) {
  if (!a.st) return false;
  if (a.stage0 == 2) return a.st->dense_go == 0;
  if (a.st->done) return true;
  if (a.stage0 && a.st->k1_valid) {
    if (a.st->k1_valid == 1) {
      const size_t n_var = static_cast<size_t>(a.n) * 6;
      for (size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
           e < n_var; e += static_cast<size_t>(gridDim.x) * blockDim.x)
        a.k[e] = a.kfsal[e];
    }
    return true;
  }
  return false;
}

// dxdt[0:3N] = x[3N:6N] for this tile's particles (nbody_init_dxdt_kernel,
// NBodyODE.hpp:51-56).
__device__ inline void copy_velocities(const double* __restrict__ y,
                                       double* __restrict__ k, int n, int i0) {
  const int e = threadIdx.x;
  const long base = 3L * i0;
  const long lim = 3L * n;
  if (e < 192 && base + e < lim) k[base + e] = y[lim + base + e];
}

// Whether every position coordinate of the state passes coord_shared_ok:
// all threads of the workgroup test a strided share of the 3N coordinates.
__device__ inline bool all_positions_shared_ok(const double* __restrict__ y, int n) {
  const bool ok = strided_all_shared_ok(y, 3L * n, threadIdx.x, blockDim.x);
  return __syncthreads_and(ok) != 0;
}

// The common mass when all n masses are one power of two 2^k with |k| <= 300,
// else 0 (every thread tests a strided share).  Then the per-term mass factor
// can move out of the sum: 2^k t is exact, and scaling by 2^k commutes with
// every rounding of the in-order sum while the values stay normal, so
// sum_j (2^k t_j) equals 2^k (sum_j t_j) bit for bit.  On the shared-
// reciprocal path every nonzero term is >= 2^-602 in magnitude and every
// nonzero partial sum >= 2^-654 (a multiple of its operands' ulp), at most
// 2^430, so |k| <= 300 keeps all of them normal and finite.
__device__ inline double uniform_pow2_mass(const double* __restrict__ mass, int n) {
  const uint64_t b0 = __builtin_bit_cast(uint64_t, mass[0]);
  const uint64_t ex = b0 >> 52;  // sign 0 and the biased exponent
  bool ok = (b0 & 0x000fffffffffffffull) == 0 && ex >= 1023 - 300 && ex <= 1023 + 300;
  ok = ok & strided_all_bits(mass, n, threadIdx.x, blockDim.x, b0);
  return __syncthreads_and(ok) != 0 ? mass[0] : 0.0;
}

// ------------------------------------------------------------------ exact --
constexpr int kExactWaves = 16;
constexpr int kExactJW = 3;  // partners per wave per tile (48-partner tiles, 144 KB LDS)
constexpr int kExactTJ = kExactWaves * kExactJW;

// The kExactJW partners one wave evaluates in one tile (wave-uniform, held
// in SGPRs).
struct PartnerSet {
  double x[kExactJW], y[kExactJW], z[kExactJW], m[kExactJW];
};

// Partners jb .. jb + kExactJW - 1, all < n.
template <bool kMass>
__device__ __forceinline__ void load_partner_set(PartnerSet& p, uniform_ptr yu,
                                                 uniform_ptr massu, int jb) {
  const uniform_ptr pp = yu + 3 * jb;
#pragma unroll
  for (int jj = 0; jj < kExactJW; ++jj) {
    p.x[jj] = pp[3 * jj];
    p.y[jj] = pp[3 * jj + 1];
    p.z[jj] = pp[3 * jj + 2];
    p.m[jj] = kMass ? massu[jb + jj] : 1.0;
  }
}

// The same with indices past the end clamped to n - 1: their terms are
// computed but never folded.
template <bool kMass>
__device__ __forceinline__ void load_partner_set_clamped(PartnerSet& p, uniform_ptr yu,
                                                         uniform_ptr massu, int jb, int n) {
  if (jb + kExactJW <= n) {
    load_partner_set<kMass>(p, yu, massu, jb);
    return;
  }
#pragma unroll
  for (int jj = 0; jj < kExactJW; ++jj) {
    const int j = min(jb + jj, n - 1);
    p.x[jj] = yu[3 * j];
    p.y[jj] = yu[3 * j + 1];
    p.z[jj] = yu[3 * j + 2];
    p.m[jj] = kMass ? massu[j] : 1.0;
  }
}

// This wave's kExactJW terms of one tile for its 64 targets, stored to
// dst[jj][coordinate][lane].
template <bool kMass, bool kShared>
__device__ __forceinline__ void exact_tile_terms(const PartnerSet& p, double xi, double yi,
                                                 double zi, double softening_sq,
                                                 double (*dst)[3][64], int lane) {
  double ax[kExactJW], ay[kExactJW], az[kExactJW];
#pragma unroll
  for (int jj = 0; jj < kExactJW; ++jj) {
    const double dx = p.x[jj] - xi, dy = p.y[jj] - yi, dz = p.z[jj] - zi;
    if (kShared)
      exact_term_shared_rcp<kMass>(dx, dy, dz, p.m[jj], softening_sq, ax[jj], ay[jj], az[jj]);
    else
      exact_term_ieee<kMass>(dx, dy, dz, p.m[jj], softening_sq, ax[jj], ay[jj], az[jj]);
  }
#pragma unroll
  for (int jj = 0; jj < kExactJW; ++jj) {
    dst[jj][0][lane] = ax[jj];
    dst[jj][1][lane] = ay[jj];
    dst[jj][2][lane] = az[jj];
  }
}

// Adds the terms of partners base .. base + cnt - 1 (one coordinate; src
// points at partner 0's entry for this lane) to acc in partner order,
// skipping the target's own index.  The skip adds +0 instead of the term:
// acc starts at +0 and a round-to-nearest sum is -0 only if both operands
// are, so acc is never -0 and acc + +0 is acc bit for bit -- no per-lane
// branch.
__device__ __forceinline__ double exact_fold(double acc, const double* src, int base, int cnt,
                                             int i0, int i) {
  if (cnt == kExactTJ && (base + kExactTJ <= i0 || base >= i0 + 64)) {
#pragma unroll
    for (int m = 0; m < kExactTJ; ++m) acc += src[m * 192];
  } else {
    for (int m = 0; m < cnt; ++m) {
      const double v = src[m * 192];
      acc += base + m != i ? v : 0.0;
    }
  }
  return acc;
}

// All tiles of one workgroup; returns the folded sum of the calling lane
// (waves 0-2: coordinate = wave).  Partner data is loaded one tile ahead
// (scalar loads hidden behind the current tile's arithmetic).  Tiles whose
// partners and next partners are all in range run in a loop unrolled by two
// over two partner sets, with no bounds tests and no register copies: the
// scalar instructions per tile stay few, since the scalar unit is shared by
// the 16 waves that all leave the same barrier together.
template <bool kMass, bool kShared>
__device__ __forceinline__ double exact_tiles(double (*terms)[kExactTJ][3][64], uniform_ptr yu,
                                              uniform_ptr massu, int n, double softening_sq,
                                              double xi, double yi, double zi, int wave, int lane,
                                              int i0, int i) {
  const int n_tiles = (n + kExactTJ - 1) / kExactTJ;
  const int n_full = n / kExactTJ;
  const int jw = wave * kExactJW;
  double acc = 0.0;
  PartnerSet pa, pb;
  // tile 0: nothing to fold yet
  load_partner_set_clamped<kMass>(pa, yu, massu, jw, n);
  if (n_tiles > 1) load_partner_set_clamped<kMass>(pb, yu, massu, kExactTJ + jw, n);
  exact_tile_terms<kMass, kShared>(pa, xi, yi, zi, softening_sq, terms[0] + jw, lane);
  __syncthreads();
  int t = 1;
  // Steady state: tiles t - 1 .. t + 2 are all full; tile t's partners are
  // in pb.  Tile t is odd, so its terms go to buffer 1 and tile t + 1's to
  // buffer 0.
  for (; t + 2 < n_full; t += 2) {
    if (wave < 3) acc = exact_fold(acc, &terms[0][0][wave][lane], (t - 1) * kExactTJ,
                                   kExactTJ, i0, i);
    load_partner_set<kMass>(pa, yu, massu, (t + 1) * kExactTJ + jw);
    exact_tile_terms<kMass, kShared>(pb, xi, yi, zi, softening_sq, terms[1] + jw, lane);
    __syncthreads();
    if (wave < 3) acc = exact_fold(acc, &terms[1][0][wave][lane], t * kExactTJ, kExactTJ,
                                   i0, i);
    load_partner_set<kMass>(pb, yu, massu, (t + 2) * kExactTJ + jw);
    exact_tile_terms<kMass, kShared>(pa, xi, yi, zi, softening_sq, terms[0] + jw, lane);
    __syncthreads();
  }
  // The remaining tiles (partial last tile included), then the last fold.
  for (; t <= n_tiles; ++t) {
    if (wave < 3) {
      const int base = (t - 1) * kExactTJ;
      acc = exact_fold(acc, &terms[(t - 1) & 1][0][wave][lane], base, min(kExactTJ, n - base),
                       i0, i);
    }
    if (t < n_tiles) {
      pa = pb;
      if (t + 1 < n_tiles)
        load_partner_set_clamped<kMass>(pb, yu, massu, (t + 1) * kExactTJ + jw, n);
      exact_tile_terms<kMass, kShared>(pa, xi, yi, zi, softening_sq, terms[t & 1] + jw, lane);
    }
    __syncthreads();
  }
  return acc;
}

template <bool kMass, bool kSharedRcp>
__global__ __launch_bounds__(64 * kExactWaves) void k_nbody_exact(NBodyArgs a) {
  if (prologue_skip(a)) return;
  __shared__ double terms[2][kExactTJ][3][64];
  const double* __restrict__ y = select_input(a);
  const uniform_ptr yu = as_uniform(y);
  double* __restrict__ k = a.k;
  const uniform_ptr massu = as_uniform(a.mass);
  const int n = a.n;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int i0 = blockIdx.x * 64;
  const int i = i0 + lane;
  const bool valid_i = i < n;
  const double xi = valid_i ? y[3 * i] : 0.0;
  const double yi = valid_i ? y[3 * i + 1] : 0.0;
  const double zi = valid_i ? y[3 * i + 2] : 0.0;
  // the shared-reciprocal form is used when every position coordinate passes
  // coord_shared_ok (one pass over the 3N coordinates per workgroup); with
  // masses that are all one power of two, unit-mass terms and one scale of the
  // sum (uniform_pow2_mass)
  const bool shared_ok = kSharedRcp && all_positions_shared_ok(y, n);
  const double m_uni = kMass && kSharedRcp && shared_ok ? uniform_pow2_mass(a.mass, n) : 0.0;
  double acc;
  if (kSharedRcp && shared_ok && kMass && m_uni != 0.0)
    acc = exact_tiles<false, true>(terms, yu, massu, n, a.eps2, xi, yi, zi, wave, lane, i0, i) *
          m_uni;
  else if (kSharedRcp && shared_ok)
    acc = exact_tiles<kMass, true>(terms, yu, massu, n, a.eps2, xi, yi, zi, wave, lane, i0, i);
  else
    acc = exact_tiles<kMass, false>(terms, yu, massu, n, a.eps2, xi, yi, zi, wave, lane, i0, i);
  if (wave < 3 && valid_i) k[3L * n + 3 * i + wave] = acc;
  copy_velocities(y, k, n, i0);
}

// The same exact sum with 32-target tiles: twice the workgroups, for N too
// small to fill the chip with 64-target tiles.  Lane l evaluates target l % 32
// against the kE32JW partners of slot l / 32 of its wave; each lane fetches
// its slot's partners with vector loads (the two half-waves read different
// addresses, one tile ahead), so no per-lane selection between wave-uniform
// values is needed.  Two waves fold a tile's terms in partner order: wave 0
// the x (lanes 0-31) and y (lanes 32-63) chains, wave 1 the z chains.  W waves
// per workgroup: a tile has 6 W partners and the two term buffers take
// 2 x 6W x 3 x 32 doubles (144 KB at W = 16: one workgroup per CU; 72 KB at
// W = 8: two).
constexpr int kE32JW = 3;  // partners per slot per tile
// The launcher takes the instance with the power-of-two mass path from this
// many bodies on.  Below, the second tile loop (and the register spill and
// scratch set-up it brings) costs more per launch than the 3 products per
// interaction it saves: 64-body solves ran 3-11 % slower with it.
constexpr int kE32UniformMassMinN = 4096;
template <int W>
constexpr int e32_tj = 2 * kE32JW * W;  // partners per tile

// This lane's kE32JW partners (x, y, z, m), starting at index jb.
typedef double E32Partners[kE32JW][4];

// Partners jb .. jb + kE32JW - 1, all < n.
template <bool kMass>
__device__ __forceinline__ void e32_load(E32Partners& q, const double* __restrict__ y,
                                         const double* __restrict__ mass, int jb) {
#pragma unroll
  for (int jj = 0; jj < kE32JW; ++jj) {
    q[jj][0] = y[3 * (jb + jj)];
    q[jj][1] = y[3 * (jb + jj) + 1];
    q[jj][2] = y[3 * (jb + jj) + 2];
    q[jj][3] = kMass ? mass[jb + jj] : 1.0;
  }
}

// The same with indices past the end clamped to n - 1 (their terms are
// computed but never folded).
template <bool kMass>
__device__ __forceinline__ void e32_load_clamped(E32Partners& q, const double* __restrict__ y,
                                                 const double* __restrict__ mass, int jb,
                                                 int n) {
#pragma unroll
  for (int jj = 0; jj < kE32JW; ++jj) {
    const int j = min(jb + jj, n - 1);
    q[jj][0] = y[3 * j];
    q[jj][1] = y[3 * j + 1];
    q[jj][2] = y[3 * j + 2];
    q[jj][3] = kMass ? mass[j] : 1.0;
  }
}

// This lane's terms of one tile, stored to dst[m0 + jj][coordinate][tgt].
template <bool kMass, bool kShared>
__device__ __forceinline__ void e32_terms(const E32Partners& q, double xi, double yi, double zi,
                                          double softening_sq, double (*dst)[3][32], int m0,
                                          int tgt) {
  double ax[kE32JW], ay[kE32JW], az[kE32JW];
#pragma unroll
  for (int jj = 0; jj < kE32JW; ++jj) {
    const double dx = q[jj][0] - xi, dy = q[jj][1] - yi, dz = q[jj][2] - zi;
    if (kShared)
      exact_term_shared_rcp<kMass>(dx, dy, dz, q[jj][3], softening_sq, ax[jj], ay[jj], az[jj]);
    else
      exact_term_ieee<kMass>(dx, dy, dz, q[jj][3], softening_sq, ax[jj], ay[jj], az[jj]);
  }
#pragma unroll
  for (int jj = 0; jj < kE32JW; ++jj) {
    dst[m0 + jj][0][tgt] = ax[jj];
    dst[m0 + jj][1][tgt] = ay[jj];
    dst[m0 + jj][2][tgt] = az[jj];
  }
}

// exact_fold for the 32-target layout (partner stride 96 doubles).
template <int W>
__device__ __forceinline__ double e32_fold(double acc, const double* src, int base, int cnt,
                                           int i0, int i) {
  constexpr int TJ = e32_tj<W>;
  if (cnt == TJ && (base + TJ <= i0 || base >= i0 + 32)) {
#pragma unroll
    for (int m = 0; m < TJ; ++m) acc += src[m * 96];
  } else {
    for (int m = 0; m < cnt; ++m) {
      const double v = src[m * 96];
      acc += base + m != i ? v : 0.0;
    }
  }
  return acc;
}

// All tiles of one workgroup (the structure of exact_tiles: peeled first
// tile, a two-tile steady-state loop over two register partner sets with no
// bounds tests, then the remaining tiles).  Returns the folded sum of the
// calling lane (meaningful on the folding lanes).
template <bool kMass, bool kShared, int W>
__device__ __forceinline__ double e32_tiles(double (*terms)[e32_tj<W>][3][32],
                                            const double* __restrict__ y,
                                            const double* __restrict__ mass, int n,
                                            double softening_sq, double xi, double yi, double zi,
                                            int wave, int slot, int tgt, int i0, int i, int fc) {
  constexpr int kE32TJ = e32_tj<W>;
  const int n_tiles = (n + kE32TJ - 1) / kE32TJ;
  const int n_full = n / kE32TJ;
  const int m0 = 2 * kE32JW * wave + slot * kE32JW;  // this lane's first partner in a tile
  double acc = 0.0;
  E32Partners qa, qb;
  e32_load_clamped<kMass>(qa, y, mass, m0, n);
  if (n_tiles > 1) e32_load_clamped<kMass>(qb, y, mass, kE32TJ + m0, n);
  e32_terms<kMass, kShared>(qa, xi, yi, zi, softening_sq, terms[0], m0, tgt);
  __syncthreads();
  int t = 1;
  for (; t + 2 < n_full; t += 2) {
    if (wave < 2)
      acc = e32_fold<W>(acc, &terms[0][0][fc][tgt], (t - 1) * kE32TJ, kE32TJ, i0, i);
    e32_load<kMass>(qa, y, mass, (t + 1) * kE32TJ + m0);
    e32_terms<kMass, kShared>(qb, xi, yi, zi, softening_sq, terms[1], m0, tgt);
    __syncthreads();
    if (wave < 2) acc = e32_fold<W>(acc, &terms[1][0][fc][tgt], t * kE32TJ, kE32TJ, i0, i);
    e32_load<kMass>(qb, y, mass, (t + 2) * kE32TJ + m0);
    e32_terms<kMass, kShared>(qa, xi, yi, zi, softening_sq, terms[0], m0, tgt);
    __syncthreads();
  }
  for (; t <= n_tiles; ++t) {
    if (wave < 2) {
      const int base = (t - 1) * kE32TJ;
      acc = e32_fold<W>(acc, &terms[(t - 1) & 1][0][fc][tgt], base, min(kE32TJ, n - base), i0,
                        i);
    }
    if (t < n_tiles) {
      E32Partners qc;
#pragma unroll
      for (int jj = 0; jj < kE32JW; ++jj)
#pragma unroll
        for (int c = 0; c < 4; ++c) qc[jj][c] = qb[jj][c];
      if (t + 1 < n_tiles) e32_load_clamped<kMass>(qb, y, mass, (t + 1) * kE32TJ + m0, n);
      e32_terms<kMass, kShared>(qc, xi, yi, zi, softening_sq, terms[t & 1], m0, tgt);
    }
    __syncthreads();
  }
  return acc;
}

// At W = 8 two workgroups share a CU: 4 waves per SIMD, <= 128 VGPRs each.
// kUniMass: the power-of-two mass path (with kMass and kSharedRcp).
template <bool kMass, bool kSharedRcp, int W, bool kUniMass>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(4)))
void k_nbody_exact32(NBodyArgs a) {
  if (prologue_skip(a)) return;
  __shared__ double terms[2][e32_tj<W>][3][32];
  const double* __restrict__ y = select_input(a);
  double* __restrict__ k = a.k;
  const double* __restrict__ mass = a.mass;
  const int n = a.n;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int tgt = lane & 31;
  const int slot = lane >> 5;
  const int i0 = blockIdx.x * 32;
  const int i = i0 + tgt;
  const int ic = i < n ? i : n - 1;
  const double xi = y[3 * ic], yi = y[3 * ic + 1], zi = y[3 * ic + 2];
  // fold role: wave 0 lanes 0-31 -> x, 32-63 -> y; wave 1 lanes 0-31 -> z
  // (wave 1 lanes 32-63 repeat the z fold and discard it)
  const int fc = wave == 0 ? slot : 2;
  const bool folds = (wave == 0) || (wave == 1 && slot == 0);
  const bool shared_ok = kSharedRcp && all_positions_shared_ok(y, n);
  // masses that are all one power of two: unit-mass terms and one exact scale
  // of each finished sum (uniform_pow2_mass), as in k_nbody_exact
  double m_uni = 0.0;
  if constexpr (kUniMass && kMass && kSharedRcp)
    m_uni = shared_ok ? uniform_pow2_mass(mass, n) : 0.0;
  double acc;
  if (kUniMass && kMass && kSharedRcp && shared_ok && m_uni != 0.0)
    acc = e32_tiles<false, true, W>(terms, y, mass, n, a.eps2, xi, yi, zi, wave, slot, tgt, i0,
                                    i, fc) *
          m_uni;
  else if (kSharedRcp && shared_ok)
    acc = e32_tiles<kMass, true, W>(terms, y, mass, n, a.eps2, xi, yi, zi, wave, slot, tgt, i0,
                                    i, fc);
  else
    acc = e32_tiles<kMass, false, W>(terms, y, mass, n, a.eps2, xi, yi, zi, wave, slot, tgt, i0,
                                     i, fc);
  if (folds && i < n) k[3L * n + 3 * i + fc] = acc;
  // dxdt[0:3n] = x[3n:6n] for this tile's 32 particles
  const int e = threadIdx.x;
  const long base = 3L * i0, lim = 3L * n;
  if (e < 96 && base + e < lim) k[base + e] = y[lim + base + e];
}

// ------------------------------------------------------------------- fast --
constexpr int kFastWaves = 16;

// kFastSet partners' coordinates (and masses), wave-uniform in SGPRs.
constexpr int kFastSet = 4;
template <bool kMass>
struct FastSet {
  double c[3 * kFastSet];
  double m[kMass ? kFastSet : 1];
};

template <bool kMass>
__device__ __forceinline__ void fast_load(FastSet<kMass>& s, uniform_ptr yu, uniform_ptr mass,
                                          int j) {
  const uniform_ptr pp = yu + 3 * j;
#pragma unroll
  for (int e = 0; e < 3 * kFastSet; ++e) s.c[e] = pp[e];
  if constexpr (kMass) {
#pragma unroll
    for (int u = 0; u < kFastSet; ++u) s.m[u] = mass[j + u];
  }
}

template <bool kMass, bool kGuard>
__device__ __forceinline__ void fast_set_terms(const FastSet<kMass>& s, int j, int i, double xi,
                                               double yi, double zi, double eps2, double& ax,
                                               double& ay, double& az) {
#pragma unroll
  for (int u = 0; u < kFastSet; ++u)
    fast_term<kMass, kGuard>(s.c[3 * u], s.c[3 * u + 1], s.c[3 * u + 2], kMass ? s.m[u] : 1.0,
                             j + u, i, xi, yi, zi, eps2, ax, ay, az);
}

// Waits for every scalar load issued so far.  Scalar loads complete out of
// order, so the wait counter cannot single out an older set while a newer one
// is in flight: the partner loop therefore drains the set it is about to use
// (issued one set earlier) before it issues the next.  The scheduling
// barriers keep the wait where it is written: after the previous set's terms
// and before the next set's loads.
__device__ __forceinline__ void scalar_loads_drain() {
  __builtin_amdgcn_sched_barrier(0);
  __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0); vector and export counters untouched
  __builtin_amdgcn_sched_barrier(0);
}

// Partners jbeg .. jend - 1 of target i in increasing j (wave-uniform scalar
// loads); kMass: per-term masses.  kPrefetch: full sets of kFastSet partners
// alternate between two SGPR sets, each loaded while the other's terms are
// computed -- worth it with 4 waves per SIMD (one workgroup per CU), whose
// scalar-load waits the other waves cannot cover; with 8 waves per SIMD the
// extra waits cost more than they hide, and steps of 8 partners that load and
// then wait are faster.  Both add the terms in the same order.
template <bool kMass, bool kGuard, bool kPrefetch>
__device__ __forceinline__ void fast_partners(uniform_ptr yu, uniform_ptr mass, int jbeg,
                                              int jend, int i, double xi, double yi,
                                              double zi, double eps2, double& ax, double& ay,
                                              double& az) {
  constexpr int U = kFastSet;
  int j = jbeg;
  if constexpr (!kPrefetch) {
    for (; j + 8 <= jend; j += 8) {
      const uniform_ptr pp = yu + 3 * j;
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const double mj = kMass ? mass[j + u] : 1.0;
        fast_term<kMass, kGuard>(pp[3 * u], pp[3 * u + 1], pp[3 * u + 2], mj, j + u, i, xi,
                                 yi, zi, eps2, ax, ay, az);
      }
    }
  } else if (j + U <= jend) {
    FastSet<kMass> a, b;
    fast_load<kMass>(a, yu, mass, j);
    // a holds partners j .. j + U - 1 (j + U <= jend); the next two sets are in range
    for (; j + 3 * U <= jend; j += 2 * U) {
      scalar_loads_drain();
      fast_load<kMass>(b, yu, mass, j + U);
      __builtin_amdgcn_sched_barrier(0);  // issue the loads before the terms
      fast_set_terms<kMass, kGuard>(a, j, i, xi, yi, zi, eps2, ax, ay, az);
      scalar_loads_drain();
      fast_load<kMass>(a, yu, mass, j + 2 * U);
      __builtin_amdgcn_sched_barrier(0);
      fast_set_terms<kMass, kGuard>(b, j + U, i, xi, yi, zi, eps2, ax, ay, az);
    }
    scalar_loads_drain();
    fast_set_terms<kMass, kGuard>(a, j, i, xi, yi, zi, eps2, ax, ay, az);
    j += U;
  }
  for (; j < jend; ++j) {
    const double mj = kMass ? mass[j] : 1.0;
    fast_term<kMass, kGuard>(yu[3 * j], yu[3 * j + 1], yu[3 * j + 2], mj, j, i, xi,
                             yi, zi, eps2, ax, ay, az);
  }
}

// The common mass when all n masses are equal (finite, nonzero), else 0
// (a workgroup vote over strided shares).
__device__ inline double workgroup_uniform_mass(const double* __restrict__ mass, int n) {
  const bool ok = lane_masses_uniform<kScanLoadsLean>(mass, n, threadIdx.x, blockDim.x);
  return __syncthreads_and(ok) != 0 ? mass[0] : 0.0;
}

template <bool kMass, bool kGuard, bool kPrefetch>
__global__ __launch_bounds__(64 * kFastWaves) void k_nbody_fast(NBodyArgs a) {
  if (prologue_skip(a)) return;
  __shared__ double red[kFastWaves][3][64];
  const double* __restrict__ y = select_input(a);
  const uniform_ptr yu = as_uniform(y);
  double* __restrict__ k = a.k;
  const uniform_ptr mass = as_uniform(a.mass);
  const int n = a.n;
  const double eps2 = a.eps2;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int i0 = blockIdx.x * 64;
  const int i = i0 + lane;
  const bool valid_i = i < n;
  const int ic = valid_i ? i : n - 1;
  const double xi = y[3 * ic], yi = y[3 * ic + 1], zi = y[3 * ic + 2];
  const int seg = (n + kFastWaves - 1) / kFastWaves;
  const int jbeg = min(n, wave * seg);
  const int jend = min(n, jbeg + seg);
  // equal masses: unit-mass terms and one scale of each finished sum
  const double m_uni = kMass ? workgroup_uniform_mass(a.mass, n) : 0.0;
  double ax = 0.0, ay = 0.0, az = 0.0;
  if (kMass && m_uni == 0.0)
    fast_partners<true, kGuard, kPrefetch>(yu, mass, jbeg, jend, i, xi, yi, zi, eps2, ax, ay,
                                           az);
  else
    fast_partners<false, kGuard, kPrefetch>(yu, mass, jbeg, jend, i, xi, yi, zi, eps2, ax, ay,
                                            az);
  red[wave][0][lane] = ax;
  red[wave][1][lane] = ay;
  red[wave][2][lane] = az;
  __syncthreads();
  if (wave < 3 && valid_i) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < kFastWaves; ++w) s += red[w][wave][lane];
    k[3L * n + 3 * i + wave] = kMass && m_uni != 0.0 ? s * m_uni : s;
  }
  copy_velocities(y, k, n, i0);
}

// ------------------------------------------------------------------ mixed --
constexpr int kMixedWaves = 16;
constexpr int kMixedFlush = 32;

// Partners jbeg .. jend - 1 of target i into the FP64 sums a*: FP32 terms
// flushed every kMixedFlush partners; kNear: near pairs (FP32 r^2 < near2) in
// FP64; kMass: per-term masses.
template <bool kMass, bool kGuard, bool kNear>
__device__ __forceinline__ void mixed_partners(uniform_ptr yu, uniform_ptr mass, int jbeg,
                                               int jend, int i, double xi, double yi,
                                               double zi, float eps2f, double eps2,
                                               float near2, double& ax, double& ay,
                                               double& az) {
  int j = jbeg;
  while (j < jend) {
    const int jstop = min(jend, j + kMixedFlush);
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    if constexpr (!kNear) {
      for (; j + 8 <= jstop; j += 8) {
        const uniform_ptr pp = yu + 3 * j;
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const float mj = kMass ? static_cast<float>(mass[j + u]) : 1.0f;
          mixed_term_fp32<kMass, kGuard>(pp[3 * u], pp[3 * u + 1], pp[3 * u + 2], mj, j + u,
                                         i, xi, yi, zi, eps2f, fx, fy, fz);
        }
      }
      for (; j < jstop; ++j) {
        const float mj = kMass ? static_cast<float>(mass[j]) : 1.0f;
        mixed_term_fp32<kMass, kGuard>(yu[3 * j], yu[3 * j + 1], yu[3 * j + 2], mj, j, i, xi,
                                       yi, zi, eps2f, fx, fy, fz);
      }
    }
    for (; kNear && j + 8 <= jstop; j += 8) {
      const uniform_ptr pp = yu + 3 * j;
      // bit u: some lane has partner j + u in its near field
      uint32_t near = 0;
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const double mj = kMass ? mass[j + u] : 1.0;
        const bool far = mixed_term<kMass, kGuard>(pp[3 * u], pp[3 * u + 1], pp[3 * u + 2], mj,
                                                   j + u, i, xi, yi, zi, eps2f, near2, fx, fy,
                                                   fz);
        near |= ballot_not(far) ? (1u << u) : 0u;
      }
      fx = opaque(fx);
      fy = opaque(fy);
      fz = opaque(fz);
      if (near) {
        const double xo = opaque(xi), yo = opaque(yi), zo = opaque(zi);
        do {
          const int u = __builtin_ctz(near);
          near &= near - 1;
          const double mj = kMass ? mass[j + u] : 1.0;
          mixed_near_term<kMass, kGuard>(pp[3 * u], pp[3 * u + 1], pp[3 * u + 2], mj, j + u,
                                         i, xo, yo, zo, eps2f, eps2, near2, ax, ay, az);
        } while (near);
      }
    }
    for (; kNear && j < jstop; ++j) {
      const double mj = kMass ? mass[j] : 1.0;
      if (!mixed_term<kMass, kGuard>(yu[3 * j], yu[3 * j + 1], yu[3 * j + 2], mj, j, i, xi,
                                     yi, zi, eps2f, near2, fx, fy, fz))
        mixed_near_term<kMass, kGuard>(yu[3 * j], yu[3 * j + 1], yu[3 * j + 2], mj, j, i,
                                       xi, yi, zi, eps2f, eps2, near2, ax, ay, az);
    }
    ax += static_cast<double>(fx);
    ay += static_cast<double>(fy);
    az += static_cast<double>(fz);
  }
}

// kNear: near pairs (a.near2 > 0) in FP64; without it every pair is FP32.
template <bool kMass, bool kGuard, bool kNear>
__global__ __launch_bounds__(64 * kMixedWaves) void k_nbody_mixed(NBodyArgs a) {
  if (prologue_skip(a)) return;
  __shared__ double red[kMixedWaves][3][64];
  const double* __restrict__ y = select_input(a);
  const uniform_ptr yu = as_uniform(y);
  double* __restrict__ k = a.k;
  const uniform_ptr mass = as_uniform(a.mass);
  const int n = a.n;
  const float eps2f = static_cast<float>(a.eps2);
  const float near2 = a.near2;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int i0 = blockIdx.x * 64;
  const int i = i0 + lane;
  const bool valid_i = i < n;
  const int ic = valid_i ? i : n - 1;
  const double xi = y[3 * ic], yi = y[3 * ic + 1], zi = y[3 * ic + 2];
  const int seg = (n + kMixedWaves - 1) / kMixedWaves;
  const int jbeg = min(n, wave * seg);
  const int jend = min(n, jbeg + seg);
  // equal masses: unit-mass terms and one FP64 scale of each finished sum
  const double m_uni = kMass ? workgroup_uniform_mass(a.mass, n) : 0.0;
  double ax = 0.0, ay = 0.0, az = 0.0;
  if (kMass && m_uni == 0.0)
    mixed_partners<true, kGuard, kNear>(yu, mass, jbeg, jend, i, xi, yi, zi, eps2f, a.eps2,
                                        near2, ax, ay, az);
  else
    mixed_partners<false, kGuard, kNear>(yu, mass, jbeg, jend, i, xi, yi, zi, eps2f, a.eps2,
                                         near2, ax, ay, az);
  red[wave][0][lane] = ax;
  red[wave][1][lane] = ay;
  red[wave][2][lane] = az;
  __syncthreads();
  if (wave < 3 && valid_i) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < kMixedWaves; ++w) s += red[w][wave][lane];
    k[3L * n + 3 * i + wave] = kMass && m_uni != 0.0 ? s * m_uni : s;
  }
  copy_velocities(y, k, n, i0);
}

// ----------------------------------------------------------------- energy --
// Per particle i: 0.5 m_i |v_i|^2 - 0.5 m_i sum_{j != i} m_j / sqrt(r^2+eps^2)
// (each pair counted from both ends), FP64 with correctly rounded sqrt and
// division; reduced afterwards in a fixed tree order.
template <bool kMass>
__global__ __launch_bounds__(256) void k_energy(const double* __restrict__ y,
                                                int n, double eps2,
                                                const double* __restrict__ mass,
                                                double* __restrict__ out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double xi = y[3 * i], yi = y[3 * i + 1], zi = y[3 * i + 2];
  double pe = 0.0;
  for (int j = 0; j < n; ++j) {
    if (j == i) continue;
    const double dx = y[3 * j] - xi, dy = y[3 * j + 1] - yi, dz = y[3 * j + 2] - zi;
    const double inv = 1.0 / __builtin_sqrt(dx * dx + dy * dy + dz * dz + eps2);
    pe += kMass ? mass[j] * inv : inv;
  }
  const double mi = kMass ? mass[i] : 1.0;
  const double* v = y + 3L * n;
  const double v2 = v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2];
  out[i] = 0.5 * mi * v2;            // kinetic
  out[n + i] = -0.5 * mi * pe;       // potential (half of each pair)
}

// m_i v_i per component: out[c * n + i].
__global__ void k_momentum_terms(const double* __restrict__ y, int n,
                                 const double* __restrict__ mass,
                                 double* __restrict__ out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double mi = mass ? mass[i] : 1.0;
  const double* v = y + 3L * n;
  for (int c = 0; c < 3; ++c) out[static_cast<long>(c) * n + i] = mi * v[3 * i + c];
}

// Eight independent FMA chains per lane; lane 0 of block 0 records the
// shader-clock and 100 MHz real-time counters around the loop.
__global__ __launch_bounds__(1024) void k_probe_fp64(double* out, long long* clk,
                                                     int iters) {
  double a[8];
  for (int u = 0; u < 8; ++u) a[u] = threadIdx.x * 1e-9 + u;
  const double b = 0.999999, c = 1e-7;
  long long t0 = 0, r0 = 0;
  const bool rec = threadIdx.x == 0 && blockIdx.x == 0;
  if (rec) {
    t0 = __builtin_amdgcn_s_memtime();
    r0 = __builtin_amdgcn_s_memrealtime();
  }
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int u = 0; u < 8; ++u) a[u] = __builtin_fma(a[u], b, c);
  }
  if (rec) {
    clk[0] = __builtin_amdgcn_s_memtime() - t0;
    clk[1] = __builtin_amdgcn_s_memrealtime() - r0;
  }
  double s = 0.0;
  for (int u = 0; u < 8; ++u) s += a[u];
  out[static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x] = s;
}

}  // namespace

hipError_t launch_probe_fp64(double* out, long long* clk, int blocks, int iters,
                             hipStream_t s) {
  hipLaunchKernelGGL(k_probe_fp64, dim3(blocks), dim3(1024), 0, s, out, clk, iters);
  return hipGetLastError();
}

hipError_t launch_momentum_terms(const double* x, int n, const double* mass,
                                 double* out, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_momentum_terms, dim3((n + 255) / 256), dim3(256), 0, s,
                     x, n, mass, out);
  return hipGetLastError();
}

namespace {
// Compute units of the current device (cached per device id).
int device_cus() {
  static int cache[64] = {};
  int d = 0;
  if (hipGetDevice(&d) != hipSuccess || d < 0 || d >= 64) return 256;
  if (cache[d] <= 0) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, d) != hipSuccess ||
        v <= 0)
      v = 256;
    cache[d] = v;
  }
  return cache[d];
}
}  // namespace

hipError_t launch_nbody_rhs(const NBodyArgs& a, int variant, hipStream_t s) {
  if (a.n <= 0) return hipSuccess;
  const unsigned blocks = static_cast<unsigned>((a.n + 63) / 64);
  const bool mass = a.mass != nullptr;
  const bool guard = !(a.eps2 >= 1e-150);
  switch (variant) {
    case ODEHIP_NBODY_EXACT: {
      // the shared-reciprocal form needs softening^2 in [2^-200, 2^100]
      const bool shared = a.eps2 >= 0x1p-200 && a.eps2 <= 0x1p100;
      // Both tile shapes sum every target's partners in the same order and
      // hold one workgroup per CU (144 KB LDS each), so a launch takes
      // ceil(workgroups / CUs) rounds.  A round of the 32-target kernel costs
      // 0.62 of a 64-target one (MI355X sweep, N = 16384 ... 65536), so it
      // wins while it needs at most as many rounds (N <= 8192 on 256 CUs) and
      // where the 64-target grid's last round is half empty (e.g. N = 24576).
      const long cus = device_cus();
      const long r64 = ((a.n + 63) / 64 + cus - 1) / cus;
      const long r32 = ((a.n + 31) / 32 + cus - 1) / cus;
      if (0.62 * static_cast<double>(r32) < static_cast<double>(r64)) {
        const unsigned b32 = static_cast<unsigned>((a.n + 31) / 32);
        const dim3 th(64 * 16);
        const bool uni = a.n >= kE32UniformMassMinN;
        if (mass) {
          if (shared && uni)
            hipLaunchKernelGGL((k_nbody_exact32<true, true, 16, true>), dim3(b32), th, 0, s, a);
          else if (shared)
            hipLaunchKernelGGL((k_nbody_exact32<true, true, 16, false>), dim3(b32), th, 0, s, a);
          else
            hipLaunchKernelGGL((k_nbody_exact32<true, false, 16, false>), dim3(b32), th, 0, s, a);
        } else {
          if (shared)
            hipLaunchKernelGGL((k_nbody_exact32<false, true, 16, false>), dim3(b32), th, 0, s, a);
          else
            hipLaunchKernelGGL((k_nbody_exact32<false, false, 16, false>), dim3(b32), th, 0, s, a);
        }
        break;
      }
      if (mass) {
        if (shared)
          hipLaunchKernelGGL((k_nbody_exact<true, true>), dim3(blocks), dim3(64 * kExactWaves), 0, s, a);
        else
          hipLaunchKernelGGL((k_nbody_exact<true, false>), dim3(blocks), dim3(64 * kExactWaves), 0, s, a);
      } else {
        if (shared)
          hipLaunchKernelGGL((k_nbody_exact<false, true>), dim3(blocks), dim3(64 * kExactWaves), 0, s, a);
        else
          hipLaunchKernelGGL((k_nbody_exact<false, false>), dim3(blocks), dim3(64 * kExactWaves), 0, s, a);
      }
      break;
    }
    case ODEHIP_NBODY_FAST: {
      // one workgroup per CU at most: 4 waves per SIMD, prefetching partner loop
      const bool prefetch = static_cast<long>(blocks) <= device_cus();
      auto* kern = mass ? (guard ? (prefetch ? k_nbody_fast<true, true, true>
                                             : k_nbody_fast<true, true, false>)
                                 : (prefetch ? k_nbody_fast<true, false, true>
                                             : k_nbody_fast<true, false, false>))
                        : (guard ? (prefetch ? k_nbody_fast<false, true, true>
                                             : k_nbody_fast<false, true, false>)
                                 : (prefetch ? k_nbody_fast<false, false, true>
                                             : k_nbody_fast<false, false, false>));
      hipLaunchKernelGGL(kern, dim3(blocks), dim3(64 * kFastWaves), 0, s, a);
      break;
    }
    case ODEHIP_NBODY_MIXED: {
      const bool guardf = !(static_cast<float>(a.eps2) >= 1e-12f);
      const bool near = a.near2 > 0.0f;
      const dim3 g(blocks), th(64 * kMixedWaves);
      auto* kern = mass ? (guardf ? (near ? k_nbody_mixed<true, true, true>
                                          : k_nbody_mixed<true, true, false>)
                                  : (near ? k_nbody_mixed<true, false, true>
                                          : k_nbody_mixed<true, false, false>))
                        : (guardf ? (near ? k_nbody_mixed<false, true, true>
                                          : k_nbody_mixed<false, true, false>)
                                  : (near ? k_nbody_mixed<false, false, true>
                                          : k_nbody_mixed<false, false, false>));
      hipLaunchKernelGGL(kern, g, th, 0, s, a);
      break;
    }
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_nbody_energy(const double* x, int n, double eps2,
                               const double* mass, double* per_particle,
                               hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const unsigned blocks = static_cast<unsigned>((n + 255) / 256);
  if (mass)
    hipLaunchKernelGGL(k_energy<true>, dim3(blocks), dim3(256), 0, s, x, n, eps2, mass, per_particle);
  else
    hipLaunchKernelGGL(k_energy<false>, dim3(blocks), dim3(256), 0, s, x, n, eps2, mass, per_particle);
  return hipGetLastError();
}

}  // namespace odehip
