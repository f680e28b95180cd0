// Ensemble kernels: the embedded RK attempt of many independent N-body
// systems advanced together, one launch per stage for all of them.
//
// Layout: the systems' states are packed (system k at element offset
// eoff_k = 6 (n_0 + ... + n_{k-1})).  The element-wise kernels run one
// 256-thread workgroup per reduction group; a group never straddles two
// systems (groups are counted from each system's first element), so a
// system's error norm is reduced in exactly the order of a single-system
// solve.  Per-element arithmetic, the group reduction and the controller are
// the functions the single-system kernels call (rk_device.hpp).
//
// The right-hand side runs one wavefront per tile of 64 (or 128) targets of
// one system: the wave walks that system's partners j = 0 .. n-1 in order,
// partner coordinates arriving through scalar loads (wave-uniform), and
// accumulates every target's acceleration in registers -- for the exact
// variant exactly the increasing-j order of the reference's single-threaded
// pair loop, so no cross-wave fold is needed (the fast and mixed variants add
// partner chunks of 128 in chunk order).  Tiles are listed largest system
// first, so the long tiles start early and the short ones fill the tail; the
// list is built on the host per tile configuration.  When few systems still
// run, a tile can instead be shared by the 2, 4 or 8 waves of a workgroup
// (k_ens_rhs_split, k_ens_rhs_exact_split), with the same bits.
#include "ensemble.hpp"
#include "mlp.hpp"
#include "nbody_terms.hpp"
#include "rk_device.hpp"

namespace odehip {
namespace {

// built-in tableaux, then the slots of registered ones
__constant__ Tableau e_tab[ODEHIP_MAX_TABLEAUS] = {
    kTableaus[0], kTableaus[1], kTableaus[2],
    kTableaus[3], kTableaus[4], kTableaus[5]};

__constant__ Dopri5Dense e_dense5 = kDopri5Dense;
__constant__ Dop853Dense e_dense8 = kDop853Dense;

constexpr int kBlock = ODEHIP_REDUCE_GROUP;

typedef const __attribute__((address_space(4))) double* uniform_ptr;
__device__ inline uniform_ptr as_uniform(const double* p) {
  return (uniform_ptr)(p);
}

// ------------------------------------------------------------- RHS tiles --

// dxdt[0:3n] = x[3n:6n] for the tile's targets (NBodyODE.hpp:51-56).
template <int TPL>
__device__ inline void tile_velocities(const double* __restrict__ y,
                                       double* __restrict__ k, int n,
                                       int i0) {
  const int lane = threadIdx.x & 63;
  const long lim = 3L * n;
  const long base = 3L * i0;
#pragma unroll
  for (int c = 0; c < 3 * TPL; ++c) {
    const long e = base + 64 * c + lane;
    if (e < lim) k[e] = y[lim + e];
  }
}

template <int TPL>
__device__ inline void tile_store(double* __restrict__ k, int n, int i0,
                                  const double* ax, const double* ay,
                                  const double* az) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int t = 0; t < TPL; ++t) {
    const int i = i0 + 64 * t + lane;
    if (i < n) {
      k[3L * n + 3 * i] = ax[t];
      k[3L * n + 3 * i + 1] = ay[t];
      k[3L * n + 3 * i + 2] = az[t];
    }
  }
}

// exact: every target sums its partners' terms in increasing j, skipping
// j == i (NBodyODE.hpp:89-98 with the order of the single-threaded loop).
// Partners j0 .. j1 - 1 in groups of U computed together; kSelf: the range
// may hold this wave's own targets, whose term is then not added.
template <bool kMass, bool kShared, bool kSelf, int TPL>
__device__ __forceinline__ void exact_sweep(uniform_ptr yu, uniform_ptr mu, int j0, int j1,
                                            const int* ic, const double* xi, const double* yi,
                                            const double* zi, double eps2, double* ax,
                                            double* ay, double* az) {
  constexpr int U = TPL == 1 ? 4 : 2;
  double px[U], py[U], pz[U], pm[U];
  auto load = [&](int j) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      px[u] = yu[3 * (j + u)];
      py[u] = yu[3 * (j + u) + 1];
      pz[u] = yu[3 * (j + u) + 2];
      pm[u] = kMass ? mu[j + u] : 1.0;
    }
  };
  int j = j0;
  if (j + U <= j1) load(j);
  for (; j + U <= j1; j += U) {
    double cx[U], cy[U], cz[U], cm[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      cx[u] = px[u];
      cy[u] = py[u];
      cz[u] = pz[u];
      cm[u] = pm[u];
    }
    if (j + 2 * U <= j1) load(j + U);  // next partners, one group ahead
    double tx[U][TPL], ty[U][TPL], tz[U][TPL];
#pragma unroll
    for (int u = 0; u < U; ++u) {
#pragma unroll
      for (int t = 0; t < TPL; ++t) {
        const double dx = cx[u] - xi[t], dy = cy[u] - yi[t], dz = cz[u] - zi[t];
        if (kShared)
          exact_term_shared_rcp<kMass>(dx, dy, dz, cm[u], eps2, tx[u][t], ty[u][t], tz[u][t]);
        else
          exact_term_ieee<kMass>(dx, dy, dz, cm[u], eps2, tx[u][t], ty[u][t], tz[u][t]);
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
#pragma unroll
      for (int t = 0; t < TPL; ++t) {
        if (kSelf && j + u == ic[t]) continue;
        ax[t] += tx[u][t];
        ay[t] += ty[u][t];
        az[t] += tz[u][t];
      }
    }
  }
  for (; j < j1; ++j) {
    const double cx = yu[3 * j], cy = yu[3 * j + 1], cz = yu[3 * j + 2];
    const double cm = kMass ? mu[j] : 1.0;
#pragma unroll
    for (int t = 0; t < TPL; ++t) {
      double tx, ty, tz;
      exact_term_ieee<kMass>(cx - xi[t], cy - yi[t], cz - zi[t], cm, eps2, tx, ty, tz);
      if (kSelf && j == ic[t]) continue;
      ax[t] += tx;
      ay[t] += ty;
      az[t] += tz;
    }
  }
}

// The shared-reciprocal form is used when all of the system's coordinates
// pass coord_shared_ok.  Partners are swept in three ranges in order: before
// the wave's targets, the targets' own range (the only one that needs the
// j == i test) and after it.
template <bool kMass, bool kSharedRcp, int TPL>
__device__ __forceinline__ void exact_tile(const double* __restrict__ y,
                                           const double* __restrict__ mass,
                                           double* __restrict__ k, int n, int i0,
                                           double eps2) {
  const int lane = threadIdx.x & 63;
  const uniform_ptr yu = as_uniform(y);
  const uniform_ptr mu = as_uniform(mass);
  int ic[TPL];
  double xi[TPL], yi[TPL], zi[TPL], ax[TPL], ay[TPL], az[TPL];
#pragma unroll
  for (int t = 0; t < TPL; ++t) {
    const int i = i0 + 64 * t + lane;
    ic[t] = i < n ? i : n - 1;
    xi[t] = y[3 * ic[t]];
    yi[t] = y[3 * ic[t] + 1];
    zi[t] = y[3 * ic[t] + 2];
    ax[t] = ay[t] = az[t] = 0.0;
  }
  // the shared-reciprocal form when every position coordinate of the system
  // passes coord_shared_ok (one pass of the wave over its 3n coordinates; int
  // indices: a system holds at most 2^24 bodies)
  bool shared_ok = kSharedRcp;
  if (kSharedRcp)
    shared_ok = __all(strided_all_shared_ok<kScanLoadsLean, int>(y, 3 * n, lane, 64));
  // targets of this wave: i0 .. s1 - 1 (lanes past n are clamped to n - 1)
  const int s1 = min(n, i0 + 64 * TPL);
  if (kSharedRcp && shared_ok) {
    exact_sweep<kMass, true, false, TPL>(yu, mu, 0, i0, ic, xi, yi, zi, eps2, ax, ay, az);
    exact_sweep<kMass, true, true, TPL>(yu, mu, i0, s1, ic, xi, yi, zi, eps2, ax, ay, az);
    exact_sweep<kMass, true, false, TPL>(yu, mu, s1, n, ic, xi, yi, zi, eps2, ax, ay, az);
  } else {
    exact_sweep<kMass, false, false, TPL>(yu, mu, 0, i0, ic, xi, yi, zi, eps2, ax, ay, az);
    exact_sweep<kMass, false, true, TPL>(yu, mu, i0, s1, ic, xi, yi, zi, eps2, ax, ay, az);
    exact_sweep<kMass, false, false, TPL>(yu, mu, s1, n, ic, xi, yi, zi, eps2, ax, ay, az);
  }
  tile_store<TPL>(k, n, i0, ax, ay, az);
  tile_velocities<TPL>(y, k, n, i0);
}

// Partner chunks of the reorderable variants (fast, mixed): each target's
// sum is the sum, in chunk order, of the sums over partners
// [c kEnsChunk, (c + 1) kEnsChunk) taken in increasing j.  The order does not
// depend on how many waves share a tile, so the split kernel below (chunks
// of one tile on several waves, folded in order) gives the same bits as one
// wave walking every chunk.

// fast: partners j0 .. j1 - 1 summed into p* (which the caller zeroes).
template <bool kMass, bool kGuard, int TPL>
__device__ inline void fast_chunk(uniform_ptr yu, uniform_ptr mu, int j0, int j1,
                                  const int* ic, const double* xi, const double* yi,
                                  const double* zi, double eps2, double* px, double* py,
                                  double* pz) {
  constexpr int U = TPL == 1 ? 8 : 4;
  int j = j0;
  for (; j + U <= j1; j += U) {
    const uniform_ptr pp = yu + 3 * j;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const double mj = kMass ? mu[j + u] : 1.0;
#pragma unroll
      for (int t = 0; t < TPL; ++t)
        fast_term<kMass, kGuard>(pp[3 * u], pp[3 * u + 1], pp[3 * u + 2], mj, j + u,
                                 ic[t], xi[t], yi[t], zi[t], eps2, px[t], py[t], pz[t]);
    }
  }
  for (; j < j1; ++j) {
    const double mj = kMass ? mu[j] : 1.0;
#pragma unroll
    for (int t = 0; t < TPL; ++t)
      fast_term<kMass, kGuard>(yu[3 * j], yu[3 * j + 1], yu[3 * j + 2], mj, j, ic[t],
                               xi[t], yi[t], zi[t], eps2, px[t], py[t], pz[t]);
  }
}

// The targets of a tile: coordinates of i0 + 64 t + lane (lanes past n
// clamped to n - 1, their results not stored).
template <int TPL>
__device__ inline void tile_targets(const double* __restrict__ y, int n, int i0, int* ic,
                                    double* xi, double* yi, double* zi) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int t = 0; t < TPL; ++t) {
    const int i = i0 + 64 * t + lane;
    ic[t] = i < n ? i : n - 1;
    xi[t] = y[3 * ic[t]];
    yi[t] = y[3 * ic[t] + 1];
    zi[t] = y[3 * ic[t] + 2];
  }
}

// scale: 1, or the common mass when the terms were summed with unit masses.
template <bool kMass, bool kGuard, int TPL>
__device__ inline void fast_tile(const double* __restrict__ y,
                                 const double* __restrict__ mass,
                                 double* __restrict__ k, int n, int i0,
                                 double eps2, double scale) {
  const uniform_ptr yu = as_uniform(y);
  const uniform_ptr mu = as_uniform(mass);
  int ic[TPL];
  double xi[TPL], yi[TPL], zi[TPL], ax[TPL], ay[TPL], az[TPL];
  tile_targets<TPL>(y, n, i0, ic, xi, yi, zi);
#pragma unroll
  for (int t = 0; t < TPL; ++t) ax[t] = ay[t] = az[t] = 0.0;
  for (int c0 = 0; c0 < n; c0 += kEnsChunk) {
    double px[TPL], py[TPL], pz[TPL];
#pragma unroll
    for (int t = 0; t < TPL; ++t) px[t] = py[t] = pz[t] = 0.0;
    fast_chunk<kMass, kGuard, TPL>(yu, mu, c0, min(n, c0 + kEnsChunk), ic, xi, yi, zi, eps2,
                                   px, py, pz);
#pragma unroll
    for (int t = 0; t < TPL; ++t) {
      ax[t] += px[t];
      ay[t] += py[t];
      az[t] += pz[t];
    }
  }
#pragma unroll
  for (int t = 0; t < TPL; ++t) {
    ax[t] *= scale;
    ay[t] *= scale;
    az[t] *= scale;
  }
  tile_store<TPL>(k, n, i0, ax, ay, az);
  tile_velocities<TPL>(y, k, n, i0);
}

// mixed: FP64 displacement rounded to FP32, FP32 pair force, partial sums
// in FP32 flushed into FP64 every 32 partners.  kNear: near pairs (FP32
// r^2 < near2) are left out of the FP32 sums and their FP64 term is added
// into the FP64 sums; without it every pair is FP32.
// Partners j0 .. j1 - 1 (j0 a multiple of kFlush) summed into the FP64
// sums a* (which the caller zeroes).
template <bool kMass, bool kGuard, bool kNear, int TPL>
__device__ inline void mixed_chunk(uniform_ptr yu, uniform_ptr mu, int j0, int j1,
                                   const int* ic, const double* xi, const double* yi,
                                   const double* zi, double eps2, float near2, double* ax,
                                   double* ay, double* az) {
  constexpr int kFlush = 32;
  constexpr int U = 8;
  static_assert(kEnsChunk % kFlush == 0, "chunks start on a flush boundary");
  const float eps2f = static_cast<float>(eps2);
  int j = j0;
  while (j < j1) {
    const int jstop = min(j1, j + kFlush);
    float fx[TPL], fy[TPL], fz[TPL];
#pragma unroll
    for (int t = 0; t < TPL; ++t) fx[t] = fy[t] = fz[t] = 0.0f;
    if constexpr (!kNear) {
      for (; j + U <= jstop; j += U) {
        const uniform_ptr pp = yu + 3 * j;
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const float mj = kMass ? static_cast<float>(mu[j + u]) : 1.0f;
#pragma unroll
          for (int t = 0; t < TPL; ++t)
            mixed_term_fp32<kMass, kGuard>(pp[3 * u], pp[3 * u + 1], pp[3 * u + 2], mj,
                                           j + u, ic[t], xi[t], yi[t], zi[t], eps2f, fx[t],
                                           fy[t], fz[t]);
        }
      }
      for (; j < jstop; ++j) {
        const float mj = kMass ? static_cast<float>(mu[j]) : 1.0f;
#pragma unroll
        for (int t = 0; t < TPL; ++t)
          mixed_term_fp32<kMass, kGuard>(yu[3 * j], yu[3 * j + 1], yu[3 * j + 2], mj, j,
                                         ic[t], xi[t], yi[t], zi[t], eps2f, fx[t], fy[t],
                                         fz[t]);
      }
    } else {
      for (; j + U <= jstop; j += U) {
        const uniform_ptr pp = yu + 3 * j;
        // bit u: some lane has partner j + u in the near field of one of its targets
        uint32_t near = 0;
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const double mj = kMass ? mu[j + u] : 1.0;
          uint64_t lanes = 0;
#pragma unroll
          for (int t = 0; t < TPL; ++t)
            lanes |= ballot_not(mixed_term<kMass, kGuard>(pp[3 * u], pp[3 * u + 1],
                                                          pp[3 * u + 2], mj, j + u, ic[t],
                                                          xi[t], yi[t], zi[t], eps2f, near2,
                                                          fx[t], fy[t], fz[t]));
          near |= lanes ? (1u << u) : 0u;
        }
#pragma unroll
        for (int t = 0; t < TPL; ++t) {
          fx[t] = opaque(fx[t]);
          fy[t] = opaque(fy[t]);
          fz[t] = opaque(fz[t]);
        }
        if (near) {
          double xo[TPL], yo[TPL], zo[TPL];
#pragma unroll
          for (int t = 0; t < TPL; ++t) {
            xo[t] = opaque(xi[t]);
            yo[t] = opaque(yi[t]);
            zo[t] = opaque(zi[t]);
          }
          do {
            const int u = __builtin_ctz(near);
            near &= near - 1;
            const double mj = kMass ? mu[j + u] : 1.0;
#pragma unroll
            for (int t = 0; t < TPL; ++t)
              mixed_near_term<kMass, kGuard>(pp[3 * u], pp[3 * u + 1], pp[3 * u + 2], mj,
                                             j + u, ic[t], xo[t], yo[t], zo[t], eps2f, eps2,
                                             near2, ax[t], ay[t], az[t]);
          } while (near);
        }
      }
      for (; j < jstop; ++j) {
        const double mj = kMass ? mu[j] : 1.0;
#pragma unroll
        for (int t = 0; t < TPL; ++t)
          if (!mixed_term<kMass, kGuard>(yu[3 * j], yu[3 * j + 1], yu[3 * j + 2], mj, j,
                                         ic[t], xi[t], yi[t], zi[t], eps2f, near2, fx[t],
                                         fy[t], fz[t]))
            mixed_near_term<kMass, kGuard>(yu[3 * j], yu[3 * j + 1], yu[3 * j + 2], mj, j,
                                           ic[t], xi[t], yi[t], zi[t], eps2f, eps2, near2,
                                           ax[t], ay[t], az[t]);
      }
    }
#pragma unroll
    for (int t = 0; t < TPL; ++t) {
      ax[t] += static_cast<double>(fx[t]);
      ay[t] += static_cast<double>(fy[t]);
      az[t] += static_cast<double>(fz[t]);
    }
  }
}

template <bool kMass, bool kGuard, bool kNear, int TPL>
__device__ inline void mixed_tile(const double* __restrict__ y,
                                  const double* __restrict__ mass,
                                  double* __restrict__ k, int n, int i0,
                                  double eps2, float near2, double scale) {
  const uniform_ptr yu = as_uniform(y);
  const uniform_ptr mu = as_uniform(mass);
  int ic[TPL];
  double xi[TPL], yi[TPL], zi[TPL], ax[TPL], ay[TPL], az[TPL];
  tile_targets<TPL>(y, n, i0, ic, xi, yi, zi);
#pragma unroll
  for (int t = 0; t < TPL; ++t) ax[t] = ay[t] = az[t] = 0.0;
  for (int c0 = 0; c0 < n; c0 += kEnsChunk) {
    double px[TPL], py[TPL], pz[TPL];
#pragma unroll
    for (int t = 0; t < TPL; ++t) px[t] = py[t] = pz[t] = 0.0;
    mixed_chunk<kMass, kGuard, kNear, TPL>(yu, mu, c0, min(n, c0 + kEnsChunk), ic, xi, yi,
                                           zi, eps2, near2, px, py, pz);
#pragma unroll
    for (int t = 0; t < TPL; ++t) {
      ax[t] += px[t];
      ay[t] += py[t];
      az[t] += pz[t];
    }
  }
#pragma unroll
  for (int t = 0; t < TPL; ++t) {
    ax[t] *= scale;
    ay[t] *= scale;
    az[t] *= scale;
  }
  tile_store<TPL>(k, n, i0, ax, ay, az);
  tile_velocities<TPL>(y, k, n, i0);
}

// The system's common mass when all its masses are equal (finite, nonzero),
// else 0: from the solver state a solve fills at its start, or, for a plain
// evaluation, a vote of the calling wave (every lane must call).
__device__ inline double system_uniform_mass(const SolverState* st, const double* m, int n) {
  if (st) return st->mass_uniform;
  return __all(lane_masses_uniform<kScanLoadsLean>(m, n, threadIdx.x & 63, 64)) ? m[0] : 0.0;
}

// One wave per tile slot.  kSpecial: shared reciprocal (exact) or the
// diagonal guard (fast, mixed); kNear: the mixed variant's near field.
template <int kVariant, bool kMass, bool kSpecial, bool kNear, int WPB>
__global__ __launch_bounds__(64 * WPB) void k_ens_rhs(EnsBufs b,
                                                      const double* y_in,
                                                      double* k_out,
                                                      int stage0) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int slot = blockIdx.x * WPB + wave;
  if (slot >= b.n_tiles) return;
  const EnsTile tile = b.tiles[slot];
  if (tile.sys < 0 || tile.sys >= b.n_sys) return;
  const SolverState* st = b.st ? b.st + tile.sys : nullptr;
  // stage0 == 2: an evaluation of the dense-output pass, for the systems
  // whose pass runs this round (finished or not)
  if (st && (stage0 == 2 ? st->dense_go == 0 : st->done)) return;
  const EnsSys sy = b.sys[tile.sys];
  const int n = sy.n;
  const int tpl = n >= b.tpl2_min_n ? 2 : 1;
  double* k = k_out + sy.eoff;
  if (st && stage0 == 1 && st->k1_valid) {
    // k1 of this attempt is the FSAL copy (or kept after a rejection)
    if (st->k1_valid == 1) {
      const int lane = threadIdx.x & 63;
      const long lim = 3L * n;
      const double* src = b.kfsal + sy.eoff;
      for (int c = 0; c < 3 * tpl; ++c) {
        const long e = 3L * tile.i0 + 64 * c + lane;
        if (e < lim) {
          k[e] = src[e];
          k[lim + e] = src[lim + e];
        }
      }
    }
    return;
  }
  const double* y = (y_in ? y_in : (st->cur ? b.xb1 : b.xb0)) + sy.eoff;
  const double* m = kMass ? b.mass + sy.eoff / 6 : nullptr;
  if (kVariant == ODEHIP_NBODY_EXACT) {
    if (tpl == 2)
      exact_tile<kMass, kSpecial, 2>(y, m, k, n, tile.i0, b.eps2);
    else
      exact_tile<kMass, kSpecial, 1>(y, m, k, n, tile.i0, b.eps2);
  } else {
    // equal masses: unit-mass terms, each finished sum scaled once
    const double m_uni = kMass ? system_uniform_mass(st, m, n) : 0.0;
    const bool per_term = kMass && m_uni == 0.0;
    const double scale = per_term || !kMass ? 1.0 : m_uni;
    if (kVariant == ODEHIP_NBODY_FAST) {
      if (per_term) {
        if (tpl == 2)
          fast_tile<true, kSpecial, 2>(y, m, k, n, tile.i0, b.eps2, 1.0);
        else
          fast_tile<true, kSpecial, 1>(y, m, k, n, tile.i0, b.eps2, 1.0);
      } else {
        if (tpl == 2)
          fast_tile<false, kSpecial, 2>(y, m, k, n, tile.i0, b.eps2, scale);
        else
          fast_tile<false, kSpecial, 1>(y, m, k, n, tile.i0, b.eps2, scale);
      }
    } else {
      const float near2 = kNear ? mixed_near2(b.near_c, n) : 0.0f;
      if (per_term) {
        if (tpl == 2)
          mixed_tile<true, kSpecial, kNear, 2>(y, m, k, n, tile.i0, b.eps2, near2, 1.0);
        else
          mixed_tile<true, kSpecial, kNear, 1>(y, m, k, n, tile.i0, b.eps2, near2, 1.0);
      } else {
        if (tpl == 2)
          mixed_tile<false, kSpecial, kNear, 2>(y, m, k, n, tile.i0, b.eps2, near2, scale);
        else
          mixed_tile<false, kSpecial, kNear, 1>(y, m, k, n, tile.i0, b.eps2, near2, scale);
      }
    }
  }
}

// The exact variant with one 64-target tile per workgroup of C waves, for the
// tail of a solve: the C waves compute the terms of a block of C kExactP
// partners (kExactP consecutive ones each) into LDS, then waves 0, 1, 2 add
// the block's x, y, z terms to their component's sum in increasing j,
// skipping j == i -- the order of exact_tile, so the same bits.
constexpr int kExactP = 4;

template <bool kMass, bool kSharedRcp, int C>
__global__ __launch_bounds__(64 * C) void k_ens_rhs_exact_split(EnsBufs b,
                                                                const double* y_in,
                                                                double* k_out,
                                                                int stage0) {
  constexpr int B = C * kExactP;
  __shared__ double term[3][B][64];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  // every return up to the block loop is uniform over the workgroup
  if (static_cast<int>(blockIdx.x) >= b.n_tiles) return;
  const EnsTile tile = b.tiles[blockIdx.x];
  if (tile.sys < 0 || tile.sys >= b.n_sys) return;
  const SolverState* st = b.st ? b.st + tile.sys : nullptr;
  if (st && (stage0 == 2 ? st->dense_go == 0 : st->done)) return;
  const EnsSys sy = b.sys[tile.sys];
  const int n = sy.n;
  double* k = k_out + sy.eoff;
  if (st && stage0 == 1 && st->k1_valid) {
    if (st->k1_valid == 1 && wave == 0) {
      const long lim = 3L * n;
      const double* src = b.kfsal + sy.eoff;
      for (int c = 0; c < 3; ++c) {
        const long e = 3L * tile.i0 + 64 * c + lane;
        if (e < lim) {
          k[e] = src[e];
          k[lim + e] = src[lim + e];
        }
      }
    }
    return;
  }
  const double* y = (y_in ? y_in : (st->cur ? b.xb1 : b.xb0)) + sy.eoff;
  const double* m = kMass ? b.mass + sy.eoff / 6 : nullptr;
  const uniform_ptr yu = as_uniform(y);
  const uniform_ptr mu = as_uniform(m);
  int ic[1];
  double xi[1], yi[1], zi[1];
  tile_targets<1>(y, n, tile.i0, ic, xi, yi, zi);
  // the shared-reciprocal form when every position coordinate of the system
  // passes coord_shared_ok (the waves test strided shares, then vote; int
  // indices: a system holds at most 2^24 bodies)
  bool shared_ok = kSharedRcp;
  if (kSharedRcp)
    shared_ok =
        __syncthreads_and(strided_all_shared_ok<kScanLoads, int>(y, 3 * n, threadIdx.x, 64 * C)) != 0;
  const int i0 = tile.i0;
  const int s1 = min(n, i0 + 64);
  // acc0: component `wave` (waves 0..2); acc1: component 2 on wave 0 when C == 2
  double acc0 = 0.0, acc1 = 0.0;
  auto fold = [&](int comp, int jb, int cnt, bool self, double& acc) {
    if (self) {
      for (int q = 0; q < cnt; ++q) {
        const double t = term[comp][q][lane];
        acc = (jb + q == ic[0]) ? acc : acc + t;
      }
    } else if (cnt == B) {
#pragma unroll
      for (int q = 0; q < B; ++q) acc += term[comp][q][lane];
    } else {
      for (int q = 0; q < cnt; ++q) acc += term[comp][q][lane];
    }
  };
  for (int jb = 0; jb < n; jb += B) {
#pragma unroll
    for (int p = 0; p < kExactP; ++p) {
      const int j = jb + wave * kExactP + p;
      double tx = 0.0, ty = 0.0, tz = 0.0;
      if (j < n) {
        const double dx = yu[3 * j] - xi[0];
        const double dy = yu[3 * j + 1] - yi[0];
        const double dz = yu[3 * j + 2] - zi[0];
        const double mj = kMass ? mu[j] : 1.0;
        if (kSharedRcp && shared_ok)
          exact_term_shared_rcp<kMass>(dx, dy, dz, mj, b.eps2, tx, ty, tz);
        else
          exact_term_ieee<kMass>(dx, dy, dz, mj, b.eps2, tx, ty, tz);
      }
      term[0][wave * kExactP + p][lane] = tx;
      term[1][wave * kExactP + p][lane] = ty;
      term[2][wave * kExactP + p][lane] = tz;
    }
    __syncthreads();
    const int cnt = min(B, n - jb);
    const bool self = jb < s1 && jb + B > i0;  // the block holds targets of this tile
    if (wave < 3) fold(wave, jb, cnt, self, acc0);
    if constexpr (C == 2) {
      if (wave == 0) fold(2, jb, cnt, self, acc1);
    }
    __syncthreads();  // term is written again by the next block
  }
  const int i = i0 + lane;
  if (i < n) {
    if (wave < 3) k[3L * n + 3 * i + wave] = acc0;
    if constexpr (C == 2) {
      if (wave == 0) k[3L * n + 3 * i + 2] = acc1;
    }
  }
  if (wave == C - 1) tile_velocities<1>(y, k, n, i0);
}

// The reorderable variants with one tile per workgroup of C waves: wave w
// sums chunks w, w + C, ... of the tile's partners, and after every C chunks
// wave 0 adds their sums, in chunk order, from LDS -- the order fast_tile and
// mixed_tile use, so the result is the same bits.  For the tail of a solve,
// when few systems still run and one wave per tile would leave SIMDs idle.
// Tiles hold 64 targets.
template <int kVariant, bool kMass, bool kSpecial, bool kNear, int C>
__global__ __launch_bounds__(64 * C) void k_ens_rhs_split(EnsBufs b,
                                                          const double* y_in,
                                                          double* k_out,
                                                          int stage0) {
  static_assert(kVariant != ODEHIP_NBODY_EXACT, "the exact order is sequential");
  __shared__ double part[C][3][64];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  // every return up to the chunk loop is uniform over the workgroup
  if (static_cast<int>(blockIdx.x) >= b.n_tiles) return;
  const EnsTile tile = b.tiles[blockIdx.x];
  if (tile.sys < 0 || tile.sys >= b.n_sys) return;
  const SolverState* st = b.st ? b.st + tile.sys : nullptr;
  if (st && (stage0 == 2 ? st->dense_go == 0 : st->done)) return;
  const EnsSys sy = b.sys[tile.sys];
  const int n = sy.n;
  double* k = k_out + sy.eoff;
  if (st && stage0 == 1 && st->k1_valid) {
    if (st->k1_valid == 1 && wave == 0) {
      const long lim = 3L * n;
      const double* src = b.kfsal + sy.eoff;
      for (int c = 0; c < 3; ++c) {
        const long e = 3L * tile.i0 + 64 * c + lane;
        if (e < lim) {
          k[e] = src[e];
          k[lim + e] = src[lim + e];
        }
      }
    }
    return;
  }
  const double* y = (y_in ? y_in : (st->cur ? b.xb1 : b.xb0)) + sy.eoff;
  const double* m = kMass ? b.mass + sy.eoff / 6 : nullptr;
  const uniform_ptr yu = as_uniform(y);
  const uniform_ptr mu = as_uniform(m);
  const float near2 = kNear ? mixed_near2(b.near_c, n) : 0.0f;
  // equal masses: unit-mass terms, each finished sum scaled once (as k_ens_rhs)
  const double m_uni = kMass ? system_uniform_mass(st, m, n) : 0.0;
  const bool per_term = kMass && m_uni == 0.0;
  const double scale = per_term || !kMass ? 1.0 : m_uni;
  int ic[1];
  double xi[1], yi[1], zi[1];
  tile_targets<1>(y, n, tile.i0, ic, xi, yi, zi);
  double ax = 0.0, ay = 0.0, az = 0.0;
  const int n_chunks = (n + kEnsChunk - 1) / kEnsChunk;
  for (int r = 0; r < n_chunks; r += C) {
    const int c = r + wave;
    double px[1] = {0.0}, py[1] = {0.0}, pz[1] = {0.0};
    if (c < n_chunks) {
      const int j0 = c * kEnsChunk;
      const int j1 = min(n, j0 + kEnsChunk);
      if (kVariant == ODEHIP_NBODY_FAST) {
        if (per_term)
          fast_chunk<true, kSpecial, 1>(yu, mu, j0, j1, ic, xi, yi, zi, b.eps2, px, py, pz);
        else
          fast_chunk<false, kSpecial, 1>(yu, mu, j0, j1, ic, xi, yi, zi, b.eps2, px, py, pz);
      } else {
        if (per_term)
          mixed_chunk<true, kSpecial, kNear, 1>(yu, mu, j0, j1, ic, xi, yi, zi, b.eps2, near2,
                                                px, py, pz);
        else
          mixed_chunk<false, kSpecial, kNear, 1>(yu, mu, j0, j1, ic, xi, yi, zi, b.eps2,
                                                 near2, px, py, pz);
      }
    }
    if (wave > 0) {
      part[wave][0][lane] = px[0];
      part[wave][1][lane] = py[0];
      part[wave][2][lane] = pz[0];
    }
    __syncthreads();
    if (wave == 0) {
      ax += px[0];
      ay += py[0];
      az += pz[0];
      for (int w = 1; w < C && r + w < n_chunks; ++w) {
        ax += part[w][0][lane];
        ay += part[w][1][lane];
        az += part[w][2][lane];
      }
    }
    __syncthreads();  // part is written again by the next C chunks
  }
  if (wave == 0) {
    ax *= scale;
    ay *= scale;
    az *= scale;
    tile_store<1>(k, n, tile.i0, &ax, &ay, &az);
  }
  if (wave == C - 1) tile_velocities<1>(y, k, n, tile.i0);
}

// ------------------------------------------------------ element kernels --

struct GroupPos {
  int sys;
  int g;  // reduction group id (index of its partial)
  bool valid;
  size_t e;  // packed element index
};

// Workgroup blockIdx.x of an element launch: its entry of the group list.
__device__ inline GroupPos group_pos(const EnsBufs& b) {
  const int2 ge = b.groups[blockIdx.x];
  GroupPos p;
  p.sys = ge.x;
  p.g = ge.y;
  const EnsSys& sy = b.sys[ge.x];
  const long l = static_cast<long>(ge.y - sy.gbase) * kBlock + threadIdx.x;
  p.valid = l < 6L * sy.n;
  p.e = static_cast<size_t>(sy.eoff + l);
  return p;
}

// rk_stage_kernel (RKEmbeddedParallel.hpp:91-101) for every active system.
__global__ __launch_bounds__(kBlock) void k_ens_stage(EnsBufs b, int stage) {
  const GroupPos p = group_pos(b);
  const SolverState* st = b.st + p.sys;
  if (st->done) return;
  if (!p.valid) return;
  const double* x0 = st->cur ? b.xb1 : b.xb0;
  b.ytmp[p.e] = stage_value(e_tab[b.tableau], stage, st->dt, b.ks, x0, p.e, b.nv);
}

// The fused final pass of an attempt (RKEmbeddedParallel.hpp:47-55,
// 103-137): new state, embedded error, error target, the scaled squared
// error reduced into this group's partial, and the FSAL copy.
__global__ __launch_bounds__(kBlock) void k_ens_final(EnsBufs b) {
  const GroupPos p = group_pos(b);
  const SolverState* st = b.st + p.sys;
  if (st->done) return;  // uniform over the workgroup
  const Tableau& T = e_tab[b.tableau];
  const int cur = st->cur;
  const double* x0 = cur ? b.xb1 : b.xb0;
  double* x = cur ? b.xb0 : b.xb1;
  double term = 0.0;
  if (p.valid) {
    double xv, ev;
    update_values(T, st->dt, b.ks, x0, p.e, b.nv, xv, ev);
    x[p.e] = xv;
    const double sc = b.atol + b.rtol * max_abs(xv, x0[p.e]);
    const double scaled = ev / sc;
    term = scaled * scaled;
    if (st->fsal && T.fsal) b.kfsal[p.e] = b.ks[(T.S - 1) * b.nv + p.e];
  }
  const double g = group_reduce(term);
  if (threadIdx.x == 0) b.partA[p.g] = g;
}

// Accept/reject and step size, one thread per system.
__global__ void k_ens_controller(EnsBufs b) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= b.n_sys) return;
  SolverState* st = b.st + s;
  if (st->done) return;
  const EnsSys sy = b.sys[s];
  const double sum = fold_partials(b.partA + sy.gbase, sy.n_groups);
  double* trace = b.trace ? b.trace + static_cast<size_t>(s) * b.trace_cap * 3 : nullptr;
  ctrl_attempt(st, e_tab[st->tableau], sum, 6 * static_cast<size_t>(sy.n), nullptr,
               trace);
  if (st->done) atomicSub(b.n_active, 1);
}

// The same with the learned controller: one wave per 16 systems; their
// features go through the network on the matrix cores together.
__global__ __launch_bounds__(64) void k_ens_controller_learned(EnsBufs b) {
  __shared__ MlpTileLds tile;
  const MlpInfer* M = static_cast<const MlpInfer*>(b.mlp);
  const int lane = threadIdx.x;
  const int s = blockIdx.x * 16 + lane;
  SolverState* st = (lane < 16 && s < b.n_sys) ? b.st + s : nullptr;
  const bool active = st && !st->done;
  bool accept = false, judged = false;
  float raw[kMlpIn];
  for (int i = 0; i < kMlpIn; ++i) raw[i] = 0.0f;
  if (active) {
    const EnsSys sy = b.sys[s];
    const double sum = fold_partials(b.partA + sy.gbase, sy.n_groups);
    double* trace =
        b.trace ? b.trace + static_cast<size_t>(s) * b.trace_cap * 3 : nullptr;
    judged = ctrl_judge(st, e_tab[st->tableau], sum, 6 * static_cast<size_t>(sy.n),
                        nullptr, trace, &accept);
    if (judged) mlp_features(st, sy.n, raw);
  }
  if (lane < 16) mlp_tile_set_row(M, tile, lane, raw);
  __syncthreads();
  mlp_tile_forward(M, tile);
  if (active) {
    if (judged) {
      double y = tile.y[lane];
      y = y < kMlpLogRatioMin ? kMlpLogRatioMin : (y > kMlpLogRatioMax ? kMlpLogRatioMax : y);
      ctrl_step(st, accept, std::isfinite(y) ? std::exp(y) : NAN);
    }
    if (st->done) atomicSub(b.n_active, 1);
  }
}

// Dense output of every system whose attempt of this round was accepted and
// covered an output time (StepperDopr5 / StepperDopr853 prepare_dense,
// dense_out and DenseOutput::save, as the single-system pass): the gate, one
// thread per system; for DOP853 f at the new state and the three extra
// stages, evaluated for those systems only; then the rows, one workgroup per
// reduction group.  A system that finished in this round still writes its
// last step's rows.
__global__ void k_ens_dense_begin(EnsBufs b, int extra_rhs) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= b.n_sys) return;
  dense_gate(b.st + s, b.dense_t, b.n_dense, extra_rhs);
}

__global__ __launch_bounds__(kBlock) void k_ens_dense_stage8(EnsBufs b, int x) {
  const GroupPos p = group_pos(b);
  const SolverState* st = b.st + p.sys;
  if (!st->dense_go || !p.valid) return;
  b.ytmp[p.e] = dop853_dense_stage(e_dense8, x, b.ks, b.kfsal, b.nv, p.e,
                                   st->cur ? b.xb0 : b.xb1, st->dt_acc);
}

__global__ __launch_bounds__(kBlock) void k_ens_dense_out8(EnsBufs b) {
  const GroupPos p = group_pos(b);
  const SolverState* st = b.st + p.sys;
  if (!st->dense_go || !p.valid) return;
  dop853_dense_rows(e_dense8, st, b.ks, b.kfsal, b.nv, p.e, st->cur ? b.xb0 : b.xb1,
                    st->cur ? b.xb1 : b.xb0, b.dense_t, b.dense_out, b.nv);
}

__global__ __launch_bounds__(kBlock) void k_ens_dense_out5(EnsBufs b) {
  const GroupPos p = group_pos(b);
  const SolverState* st = b.st + p.sys;
  if (!st->dense_go || !p.valid) return;
  dopri5_dense_rows(e_dense5, st, b.ks, b.nv, p.e, st->cur ? b.xb0 : b.xb1,
                    st->cur ? b.xb1 : b.xb0, b.dense_t, b.dense_out, b.nv);
}

__global__ void k_ens_init(EnsBufs b, double t0, double tf, long max_acc,
                           long max_att, int fsal, int controller) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s == 0) *b.n_active = (t0 < tf) ? b.n_sys : 0;
  if (s >= b.n_sys) return;
  ctrl_init(b.st + s, e_tab[b.tableau], t0, tf, max_acc, max_att, fsal, b.tableau,
            b.sys[s].n_groups, 0, b.trace ? b.trace_cap : 0, controller);
}

// Each system's common mass (SolverState::mass_uniform), one wave per system.
__global__ __launch_bounds__(64) void k_ens_mass_uniform(EnsBufs b) {
  const int s = blockIdx.x;
  if (s >= b.n_sys) return;
  double m_uni = 0.0;
  if (b.mass) {
    const EnsSys sy = b.sys[s];
    const double* m = b.mass + sy.eoff / 6;
    m_uni = __all(lane_masses_uniform(m, sy.n, threadIdx.x, 64)) ? m[0] : 0.0;
  }
  if (threadIdx.x == 0) b.st[s].mass_uniform = m_uni;
}

// compute_error_target_kernel (RKEmbeddedParallel.hpp:151-157).
__global__ void k_ens_init_target(EnsBufs b) {
  const size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e < b.nv) b.sc0[e] = std::fabs(b.xb0[e]) * b.rtol + b.atol;
}

// Group partials of sum(((v - sub) / sc)^2) (rk_norm / difference,
// RKEmbeddedParallel.hpp:132-149, 165-168).
__global__ __launch_bounds__(kBlock) void k_ens_norm_partials(
    EnsBufs b, const double* v, const double* sub, double* partials) {
  const GroupPos p = group_pos(b);
  if (b.st[p.sys].done) return;
  double term = 0.0;
  if (p.valid) {
    const double d = sub ? (v[p.e] - sub[p.e]) : v[p.e];
    const double q = d / b.sc0[p.e];
    term = q * q;
  }
  const double g = group_reduce(term);
  if (threadIdx.x == 0) partials[p.g] = g;
}

__global__ void k_ens_init_step_a(EnsBufs b) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= b.n_sys) return;
  SolverState* st = b.st + s;
  if (st->done) return;
  const EnsSys sy = b.sys[s];
  const double sa = fold_partials(b.partA + sy.gbase, sy.n_groups);
  const double sb = fold_partials(b.partB + sy.gbase, sy.n_groups);
  ctrl_init_step_a(st, sa, sb, 6 * static_cast<size_t>(sy.n));
}

// euler_step_kernel (RKEmbeddedParallel.hpp:159-163).
__global__ __launch_bounds__(kBlock) void k_ens_euler(EnsBufs b) {
  const GroupPos p = group_pos(b);
  const SolverState* st = b.st + p.sys;
  if (st->done) return;
  if (p.valid) b.ytmp[p.e] = b.xb0[p.e] + b.ks[p.e] * st->dt0;
}

__global__ void k_ens_init_step_b(EnsBufs b) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= b.n_sys) return;
  SolverState* st = b.st + s;
  if (st->done) return;
  const EnsSys sy = b.sys[s];
  const double sa = fold_partials(b.partA + sy.gbase, sy.n_groups);
  ctrl_init_step_b(st, e_tab[st->tableau], sa, 6 * static_cast<size_t>(sy.n));
}

__global__ __launch_bounds__(kBlock) void k_ens_copy_out(EnsBufs b,
                                                         double* out) {
  const GroupPos p = group_pos(b);
  if (!p.valid) return;
  out[p.e] = (b.st[p.sys].cur ? b.xb1 : b.xb0)[p.e];
}

template <int kVariant, bool kMass, bool kSpecial, bool kNear>
hipError_t launch_rhs_v(const EnsBufs& b, const double* y, double* k,
                        int stage0, hipStream_t s) {
  if (b.split > 1) {
    // one 64-target tile per workgroup of b.split waves
    if (b.wpb != 1 || b.tpl2_min_n != kEnsNoTpl2) return hipErrorInvalidValue;
    const unsigned blocks = static_cast<unsigned>(b.n_tiles);
    if (blocks == 0) return hipSuccess;
    if constexpr (kVariant == ODEHIP_NBODY_EXACT) {
      if (b.split == 2)
        hipLaunchKernelGGL((k_ens_rhs_exact_split<kMass, kSpecial, 2>), dim3(blocks),
                           dim3(128), 0, s, b, y, k, stage0);
      else if (b.split == 4)
        hipLaunchKernelGGL((k_ens_rhs_exact_split<kMass, kSpecial, 4>), dim3(blocks),
                           dim3(256), 0, s, b, y, k, stage0);
      else if (b.split == 8)
        hipLaunchKernelGGL((k_ens_rhs_exact_split<kMass, kSpecial, 8>), dim3(blocks),
                           dim3(512), 0, s, b, y, k, stage0);
      else
        return hipErrorInvalidValue;
      return hipGetLastError();
    } else {
      if (b.split == 2)
        hipLaunchKernelGGL((k_ens_rhs_split<kVariant, kMass, kSpecial, kNear, 2>),
                           dim3(blocks), dim3(128), 0, s, b, y, k, stage0);
      else if (b.split == 4)
        hipLaunchKernelGGL((k_ens_rhs_split<kVariant, kMass, kSpecial, kNear, 4>),
                           dim3(blocks), dim3(256), 0, s, b, y, k, stage0);
      else if (b.split == 8)
        hipLaunchKernelGGL((k_ens_rhs_split<kVariant, kMass, kSpecial, kNear, 8>),
                           dim3(blocks), dim3(512), 0, s, b, y, k, stage0);
      else
        return hipErrorInvalidValue;
      return hipGetLastError();
    }
  }
  const unsigned blocks = static_cast<unsigned>(b.n_tiles / b.wpb);
  if (blocks == 0) return hipSuccess;
  if (b.wpb == 4) {
    hipLaunchKernelGGL((k_ens_rhs<kVariant, kMass, kSpecial, kNear, 4>), dim3(blocks),
                       dim3(256), 0, s, b, y, k, stage0);
  } else if (b.wpb == 1) {
    hipLaunchKernelGGL((k_ens_rhs<kVariant, kMass, kSpecial, kNear, 1>), dim3(blocks),
                       dim3(64), 0, s, b, y, k, stage0);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int kVariant, bool kNear = false>
hipError_t launch_rhs_m(const EnsBufs& b, const double* y, double* k,
                        int stage0, hipStream_t s, bool special) {
  if (b.mass)
    return special ? launch_rhs_v<kVariant, true, true, kNear>(b, y, k, stage0, s)
                   : launch_rhs_v<kVariant, true, false, kNear>(b, y, k, stage0, s);
  return special ? launch_rhs_v<kVariant, false, true, kNear>(b, y, k, stage0, s)
                 : launch_rhs_v<kVariant, false, false, kNear>(b, y, k, stage0, s);
}

inline unsigned groups_grid(const EnsBufs& b) {
  return static_cast<unsigned>(b.n_groups);
}
inline unsigned sys_grid(const EnsBufs& b) {
  return static_cast<unsigned>((b.n_sys + 63) / 64);
}

}  // namespace

hipError_t ens_launch_rhs(const EnsBufs& b, const double* y, double* k,
                          int stage0, hipStream_t s) {
  switch (b.variant) {
    case ODEHIP_NBODY_EXACT: {
      const bool shared = b.eps2 >= 0x1p-200 && b.eps2 <= 0x1p100;
      return launch_rhs_m<ODEHIP_NBODY_EXACT>(b, y, k, stage0, s, shared);
    }
    case ODEHIP_NBODY_FAST:
      return launch_rhs_m<ODEHIP_NBODY_FAST>(b, y, k, stage0, s, !(b.eps2 >= 1e-150));
    case ODEHIP_NBODY_MIXED: {
      const bool guard = !(static_cast<float>(b.eps2) >= 1e-12f);
      return b.near_c > 0.0
                 ? launch_rhs_m<ODEHIP_NBODY_MIXED, true>(b, y, k, stage0, s, guard)
                 : launch_rhs_m<ODEHIP_NBODY_MIXED, false>(b, y, k, stage0, s, guard);
    }
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t ens_enqueue_init(const EnsBufs& b, const double* x_in, double t0,
                            double tf, long max_acc, long max_att, int fsal,
                            int controller, hipStream_t s) {
  hipError_t e = hipMemcpyAsync(b.xb0, x_in, b.nv * sizeof(double),
                                hipMemcpyDeviceToDevice, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_ens_init, dim3(sys_grid(b)), dim3(64), 0, s, b, t0, tf,
                     max_acc, max_att, fsal, controller);
  hipLaunchKernelGGL(k_ens_mass_uniform, dim3(b.n_sys), dim3(64), 0, s, b);
  return hipGetLastError();
}

// estimate_initial_step (RKEmbeddedParallel.hpp:170-200) for every system.
hipError_t ens_enqueue_initial_step(const EnsBufs& b, hipStream_t s) {
  double* f0 = b.ks;
  double* f1 = b.ks + b.nv;
  const unsigned eblocks = static_cast<unsigned>((b.nv + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(k_ens_init_target, dim3(eblocks), dim3(kBlock), 0, s, b);
  hipError_t e = ens_launch_rhs(b, b.xb0, f0, 0, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_ens_norm_partials, dim3(groups_grid(b)), dim3(kBlock), 0, s,
                     b, b.xb0, nullptr, b.partA);
  hipLaunchKernelGGL(k_ens_norm_partials, dim3(groups_grid(b)), dim3(kBlock), 0, s,
                     b, f0, nullptr, b.partB);
  hipLaunchKernelGGL(k_ens_init_step_a, dim3(sys_grid(b)), dim3(64), 0, s, b);
  hipLaunchKernelGGL(k_ens_euler, dim3(groups_grid(b)), dim3(kBlock), 0, s, b);
  if ((e = ens_launch_rhs(b, b.ytmp, f1, 0, s)) != hipSuccess) return e;
  hipLaunchKernelGGL(k_ens_norm_partials, dim3(groups_grid(b)), dim3(kBlock), 0, s,
                     b, f1, f0, b.partA);
  hipLaunchKernelGGL(k_ens_init_step_b, dim3(sys_grid(b)), dim3(64), 0, s, b);
  return hipGetLastError();
}

// One lockstep round: an attempt of every active system.
hipError_t ens_enqueue_round(const EnsBufs& b, hipStream_t s) {
  const Tableau& T = host_tableau(b.tableau);
  hipError_t e = ens_launch_rhs(b, nullptr, b.ks, 1, s);
  if (e != hipSuccess) return e;
  for (int stage = 1; stage < T.S; ++stage) {
    hipLaunchKernelGGL(k_ens_stage, dim3(groups_grid(b)), dim3(kBlock), 0, s, b,
                       stage);
    if ((e = ens_launch_rhs(b, b.ytmp, b.ks + stage * b.nv, 0, s)) != hipSuccess)
      return e;
  }
  hipLaunchKernelGGL(k_ens_final, dim3(groups_grid(b)), dim3(kBlock), 0, s, b);
  if (b.controller == ODEHIP_CONTROLLER_LEARNED) {
    hipLaunchKernelGGL(k_ens_controller_learned, dim3((b.n_sys + 15) / 16), dim3(64), 0, s,
                       b);
  } else {
    hipLaunchKernelGGL(k_ens_controller, dim3(sys_grid(b)), dim3(64), 0, s, b);
  }
  if (b.n_dense > 0 && b.tableau == ODEHIP_TABLEAU_DOPRI5) {
    hipLaunchKernelGGL(k_ens_dense_begin, dim3(sys_grid(b)), dim3(64), 0, s, b, 0);
    hipLaunchKernelGGL(k_ens_dense_out5, dim3(groups_grid(b)), dim3(kBlock), 0, s, b);
  } else if (b.n_dense > 0) {
    hipLaunchKernelGGL(k_ens_dense_begin, dim3(sys_grid(b)), dim3(64), 0, s, b, 4);
    if ((e = ens_launch_rhs(b, nullptr, b.kfsal, 2, s)) != hipSuccess) return e;
    for (int x = 0; x < 3; ++x) {
      hipLaunchKernelGGL(k_ens_dense_stage8, dim3(groups_grid(b)), dim3(kBlock), 0, s, b, x);
      if ((e = ens_launch_rhs(b, b.ytmp, b.ks + (x + 1) * b.nv, 2, s)) != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_ens_dense_out8, dim3(groups_grid(b)), dim3(kBlock), 0, s, b);
  }
  return hipGetLastError();
}

hipError_t ens_enqueue_copy_out(const EnsBufs& b, double* x_out,
                                hipStream_t s) {
  hipLaunchKernelGGL(k_ens_copy_out, dim3(groups_grid(b)), dim3(kBlock), 0, s, b,
                     x_out);
  return hipGetLastError();
}

hipError_t upload_ens_tableau(int id, const Tableau& T) {
  if (id < ODEHIP_NUM_TABLEAUS || id >= ODEHIP_MAX_TABLEAUS) return hipErrorInvalidValue;
  return hipMemcpyToSymbol(HIP_SYMBOL(e_tab), &T, sizeof(Tableau), id * sizeof(Tableau),
                           hipMemcpyHostToDevice);
}

}  // namespace odehip
