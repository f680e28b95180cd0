// Host-side planning of an ensemble: the packed layout of the systems, the
// right-hand-side tile lists, the capacity of the tile region, the arena plan
// and the key of the tables held on the device.  Host code only (no HIP
// calls), so the library and its CPU tests run the same functions.
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstring>
#include <vector>

#include "ensemble.hpp"

namespace odehip {

// Host-side layout of an ensemble.
struct Layout {
  std::vector<EnsSys> sys;
  std::vector<int2> groups;  // (system, group id) of every reduction group
  std::vector<int> sizes;
  size_t nv = 0;
  int n_groups = 0;
  double interactions = 0.0;
  double work = 0.0;  // sum n^2
};

// Fills L for the given system sizes; returns null or the reason for refusing.
inline const char* build_layout(int n_systems, const int* n_particles, Layout* L) {
  if (n_systems <= 0) return "n_systems must be positive";
  if (!n_particles) return "n_particles is null";
  L->sys.resize(n_systems);
  L->sizes.assign(n_particles, n_particles + n_systems);
  long eoff = 0;
  long groups = 0;
  for (int s = 0; s < n_systems; ++s) {
    const int n = n_particles[s];
    if (n < 1 || n > (1 << 24)) return "system sizes must be in [1, 2^24]";
    const int ng = n_groups_for(6 * static_cast<size_t>(n));
    L->sys[s] = EnsSys{eoff, n, static_cast<int>(groups), ng, 0};
    eoff += 6L * n;
    groups += ng;
    L->interactions += static_cast<double>(n) * (n - 1);
    L->work += static_cast<double>(n) * n;
  }
  if (groups > (1L << 30)) return "ensemble too large";
  L->nv = static_cast<size_t>(eoff);
  L->n_groups = static_cast<int>(groups);
  L->groups.resize(groups);
  for (int s = 0; s < n_systems; ++s)
    for (int g = 0; g < L->sys[s].n_groups; ++g) {
      const int id = L->sys[s].gbase + g;
      L->groups[id] = int2{s, id};
    }
  return nullptr;
}

// Right-hand-side work items for one tile configuration: tiles of the
// largest systems first.  With xcd placement the systems are dealt to eight
// queues of about equal sum n^2 and workgroup b takes its tiles from queue
// b mod 8, the XCD that the hardware dispatches workgroup b to, so all tiles
// of one system share one XCD's L2.  `running` (optional) keeps only the
// systems it marks.
inline std::vector<EnsTile> make_tiles(const Layout& L, const EnsTileConfig& cfg,
                                       const std::vector<char>* running = nullptr) {
  const int n_sys = static_cast<int>(L.sys.size());
  std::vector<int> order;
  for (int s = 0; s < n_sys; ++s)
    if (!running || (*running)[s]) order.push_back(s);
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return L.sys[a].n > L.sys[b].n; });
  auto tiles_of = [&](int s, std::vector<EnsTile>& out) {
    const int n = L.sys[s].n;
    const int step = 64 * (n >= cfg.tpl2_min_n ? 2 : 1);
    for (int i0 = 0; i0 < n; i0 += step) out.push_back(EnsTile{s, i0});
  };
  std::vector<EnsTile> tiles;
  const EnsTile empty{-1, 0};
  if (!cfg.xcd) {
    for (int s : order) tiles_of(s, tiles);
    while (tiles.size() % cfg.wpb) tiles.push_back(empty);
    return tiles;
  }
  constexpr int kXcds = 8;
  std::vector<std::vector<EnsTile>> q(kXcds);
  double load[kXcds] = {};
  for (int s : order) {
    const int x = static_cast<int>(std::min_element(load, load + kXcds) - load);
    load[x] += static_cast<double>(L.sys[s].n) * L.sys[s].n;
    tiles_of(s, q[x]);
  }
  size_t rounds = 0;
  for (auto& v : q) rounds = std::max(rounds, (v.size() + cfg.wpb - 1) / cfg.wpb);
  tiles.reserve(rounds * kXcds * cfg.wpb);
  for (size_t r = 0; r < rounds; ++r)
    for (int x = 0; x < kXcds; ++x)
      for (int w = 0; w < cfg.wpb; ++w) {
        const size_t idx = r * cfg.wpb + w;
        tiles.push_back(idx < q[x].size() ? q[x][idx] : empty);
      }
  return tiles;
}

// Slots of the tile region: an upper bound on make_tiles(L, cfg, running)
// for every configuration and every subset of running systems.  A subset
// has at most T64 = sum ceil(n/64) real tiles (64-target tiles are the
// finest).  Without xcd placement the list is those tiles padded to the
// waves per workgroup; with it, every queue holds at most T64 of them and
// the list is 8 x wpb x (rounds of the fullest queue).  The bound of the
// full ensemble's own lists is not enough: greedy dealing of a subset can
// put more tiles into one queue than the full set did.
inline size_t tiles_capacity(const Layout& L) {
  size_t t64 = 0;
  for (const auto& sy : L.sys) t64 += (static_cast<size_t>(sy.n) + 63) / 64;
  size_t cap = 0;
  for (const auto& cfg : kEnsTileConfigs) {
    const size_t w = static_cast<size_t>(cfg.wpb);
    const size_t rounds = (t64 + w - 1) / w;
    cap = std::max(cap, (cfg.xcd ? 8 : 1) * w * rounds);
  }
  return cap;
}

// Null if every entry of a tile list names a system of L (or is empty) and
// the list fits `cap` slots; else the reason.
inline const char* check_tiles(const std::vector<EnsTile>& tiles, const Layout& L,
                               size_t cap) {
  if (tiles.size() > cap) return "tile list exceeds its region";
  const int ns = static_cast<int>(L.sys.size());
  for (const EnsTile& t : tiles) {
    if (t.sys >= ns) return "tile names a system outside the ensemble";
    if (t.sys >= 0 && (t.i0 < 0 || t.i0 >= L.sys[t.sys].n))
      return "tile starts outside its system";
  }
  return nullptr;
}

// Tiles that hold targets (empty padding tiles not counted): the waves a
// right-hand-side launch under an unsplit `cfg` puts to work, over the
// systems `running` marks (null: all).
inline long real_tiles(const Layout& L, const EnsTileConfig& cfg,
                       const std::vector<char>* running = nullptr) {
  long t = 0;
  for (size_t s = 0; s < L.sys.size(); ++s) {
    if (running && !(*running)[s]) continue;
    const int n = L.sys[s].n;
    const int step = 64 * (n >= cfg.tpl2_min_n ? 2 : 1);
    t += (n + step - 1) / step;
  }
  return t;
}

// Waves with partner chunks to sum when `split` waves share each 64-target
// tile of the running systems (a system of n particles has ceil(n / 128)
// chunks, so waves past that count have nothing to do).
inline long chunk_waves(const Layout& L, const std::vector<char>* running, int split) {
  long w = 0;
  for (size_t s = 0; s < L.sys.size(); ++s) {
    if (running && !(*running)[s]) continue;
    const long n = L.sys[s].n;
    const long chunks = (n + kEnsChunk - 1) / kEnsChunk;
    w += (n + 63) / 64 * std::min<long>(split, chunks);
  }
  return w;
}

// The configuration with the same waves per workgroup, XCD placement and
// split, and 64-target tiles for every system.
inline int without_tpl2(int cfg) {
  const EnsTileConfig& c = kEnsTileConfigs[cfg];
  for (int k = 0; k < ODEHIP_ENS_TILE_CONFIGS; ++k)
    if (kEnsTileConfigs[k].wpb == c.wpb && kEnsTileConfigs[k].xcd == c.xcd &&
        kEnsTileConfigs[k].split == c.split && kEnsTileConfigs[k].tpl2_min_n == kEnsNoTpl2)
      return k;
  return cfg;
}

// The one-wave-per-workgroup configuration of 64-target tiles with this XCD
// placement and `split` waves per tile (1: unsplit).
inline int split_config(int xcd, int split) {
  for (int k = 0; k < ODEHIP_ENS_TILE_CONFIGS; ++k) {
    const EnsTileConfig& c = kEnsTileConfigs[k];
    if (c.wpb == 1 && c.tpl2_min_n == kEnsNoTpl2 && c.xcd == xcd && c.split == split)
      return k;
  }
  return -1;
}

// SIMDs of the device (256 CUs x 4): the waves one launch needs to give
// each SIMD one.
constexpr long kEnsSimds = 1024;
// Tiles from which the reorderable variants keep the tuned configuration.
constexpr long kEnsSplitAmple = 8192;

// Shape of the running systems' right-hand side: 64-target tiles, the
// longest partner walk of a tile and the ratio of that walk to the partner
// steps per SIMD of a perfectly balanced launch (above 1 the longest tiles,
// not the total work, set the time of a launch with one wave per tile).
struct TailShape {
  long tiles = 0;
  long longest = 0;
  double ratio = 0.0;
};

inline TailShape tail_shape(const Layout& L, const std::vector<char>* running) {
  TailShape t;
  double work = 0.0;
  for (size_t s = 0; s < L.sys.size(); ++s) {
    if (running && !(*running)[s]) continue;
    const long n = L.sys[s].n;
    t.tiles += (n + 63) / 64;
    t.longest = std::max(t.longest, n);
    work += static_cast<double>((n + 63) / 64) * static_cast<double>(n);
  }
  t.ratio = work > 0.0 ? static_cast<double>(t.longest) * kEnsSimds / work : 0.0;
  return t;
}

// The configuration for the right-hand sides of the running systems, from
// `cfg` (a decision table measured on MI355X over subsets of the 2048-system
// ensemble; every choice gives the same results):
//  * exact: a tile shared by 8 waves when the longest walks dominate (ratio
//    above 1.2: the terms of a block of partners on all waves, the ordered
//    fold on three, 1.3-1.7x faster there and 4-21 % slower on balanced
//    launches); otherwise one wave per tile, with 64-target tiles once fewer
//    than kEnsFillWaves remain;
//  * fast, mixed: the configuration kept while kEnsSplitAmple tiles remain;
//    below, a tile's chunks shared by 8 waves (ratio at least 1) or 4 (3-11 %
//    faster even on balanced launches of 2000-7000 tiles, up to 2.3x on
//    dominated ones);
// in both cases never more waves per tile than add working waves
// (chunk_waves).
inline int tail_config(const Layout& L, const std::vector<char>* running, int cfg,
                       bool exact) {
  const EnsTileConfig& c = kEnsTileConfigs[cfg];
  const TailShape t = tail_shape(L, running);
  const int unsplit = c.split > 1 ? split_config(c.xcd, 1) : cfg;
  auto useful = [&](int split) {
    while (split > 1 && chunk_waves(L, running, split) <= chunk_waves(L, running, split / 2))
      split /= 2;
    return split;
  };
  if (exact) {
    if (t.ratio > 1.2) {
      if (useful(8) == 8) return split_config(c.xcd, 8);
    }
    return real_tiles(L, kEnsTileConfigs[unsplit], running) < kEnsFillWaves
               ? without_tpl2(unsplit)
               : unsplit;
  }
  if (t.tiles >= kEnsSplitAmple) return cfg;
  const int split = useful(t.ratio >= 1.0 ? 8 : 4);
  return split > 1 ? split_config(c.xcd, split) : without_tpl2(unsplit);
}

// Device arena: tables, per-system state, vectors (byte offsets, 256-aligned).
struct Plan {
  size_t sys, groups, tiles, groups_act, st, n_active, xb0, xb1, ks, ytmp, kfsal,
      sc0, partA, partB, trace, dense_t, dense_out, total;
};

// The tables (sys, groups, tiles) come first and their offsets depend on the
// layout alone, so a right-hand-side call and a solve on the same systems
// find them at the same place.  groups_act (the running systems' groups) is
// a solve-only region after them.
inline Plan plan(const Layout& L, size_t tiles_cap, int S, long trace_cap,
                 bool with_state, long n_dense = 0) {
  Plan p{};
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t at = off;
    off += (bytes + 255) & ~static_cast<size_t>(255);
    return at;
  };
  const size_t ns = L.sys.size();
  p.sys = take(ns * sizeof(EnsSys));
  p.groups = take(L.groups.size() * sizeof(int2));
  p.tiles = take(tiles_cap * sizeof(EnsTile));
  p.groups_act = with_state ? take(L.groups.size() * sizeof(int2)) : p.groups;
  p.st = take(ns * sizeof(SolverState));
  p.n_active = take(sizeof(int));
  const size_t v = L.nv * sizeof(double);
  if (with_state) {
    p.xb0 = take(v);
    p.xb1 = take(v);
    p.ks = take(static_cast<size_t>(S) * v);
    p.ytmp = take(v);
    p.kfsal = take(v);
    p.sc0 = take(v);
    p.partA = take(L.groups.size() * sizeof(double));
    p.partB = take(L.groups.size() * sizeof(double));
    p.trace = take(ns * static_cast<size_t>(trace_cap) * 3 * sizeof(double));
    p.dense_t = take(static_cast<size_t>(n_dense) * sizeof(double));
    p.dense_out = take(static_cast<size_t>(n_dense) * v);
  }
  p.total = off;
  return p;
}

// What the device's table regions hold: the tables of `cfg` for `sizes`,
// written at these offsets of the arena at `base`.  An upload is skipped only
// when every field matches.
struct TableKey {
  int cfg = -2;  // -2: nothing known (fresh arena, or overwritten by a solve)
  const void* base = nullptr;
  size_t sys = 0, groups = 0, tiles = 0;
  std::vector<int> sizes;

  bool matches(int c, const void* b, const Plan& p, const Layout& L) const {
    return cfg == c && cfg >= 0 && base == b && sys == p.sys && groups == p.groups &&
           tiles == p.tiles && sizes == L.sizes;
  }
  void set(int c, const void* b, const Plan& p, const Layout& L) {
    cfg = c;
    base = b;
    sys = p.sys;
    groups = p.groups;
    tiles = p.tiles;
    sizes = L.sizes;
  }
  void invalidate() { cfg = -2; }
};

}  // namespace odehip
