// CPU test hooks over the ensemble planner (ode_prototyping_amd/csrc/
// ensemble_plan.hpp): the very functions libodehip runs on the host to lay
// out an ensemble, build its tile lists and key the tables on the device.
// Built with g++ into libensemble_plan_check.so; tests/test_ensemble_plan.py
// drives it through ctypes.  No HIP calls.
#include <cstring>
#include <vector>

#include "ensemble_plan.hpp"

using namespace odehip;

namespace {
bool layout_of(int n, const int* sizes, Layout* L) {
  return build_layout(n, sizes, L) == nullptr;
}
}  // namespace

extern "C" {

int epc_n_configs() { return ODEHIP_ENS_TILE_CONFIGS; }

// Slots of the tile region the library allocates (-1: bad sizes).
long epc_tiles_capacity(int n, const int* sizes) {
  Layout L;
  return layout_of(n, sizes, &L) ? static_cast<long>(tiles_capacity(L)) : -1;
}

// The capacity rule the library used before: the longest full-ensemble list.
long epc_full_list_max(int n, const int* sizes) {
  Layout L;
  if (!layout_of(n, sizes, &L)) return -1;
  size_t m = 0;
  for (const auto& cfg : kEnsTileConfigs) m = std::max(m, make_tiles(L, cfg).size());
  return static_cast<long>(m);
}

// make_tiles for configuration `cfg` over the systems `running` marks (null:
// all).  Writes min(count, cap) entries and returns the full count.
long epc_make_tiles(int n, const int* sizes, int cfg, const char* running, int* sys_out,
                    int* i0_out, long cap) {
  Layout L;
  if (!layout_of(n, sizes, &L) || cfg < 0 || cfg >= ODEHIP_ENS_TILE_CONFIGS) return -1;
  std::vector<char> live;
  if (running) live.assign(running, running + n);
  const std::vector<EnsTile> t = make_tiles(L, kEnsTileConfigs[cfg], running ? &live : nullptr);
  for (long k = 0; k < cap && k < static_cast<long>(t.size()); ++k) {
    sys_out[k] = t[k].sys;
    i0_out[k] = t[k].i0;
  }
  return static_cast<long>(t.size());
}

// 0 if check_tiles accepts the list against the ensemble and `cap`, else 1.
int epc_check_tiles(int n, const int* sizes, const int* sys, const int* i0, long count,
                    long cap) {
  Layout L;
  if (!layout_of(n, sizes, &L)) return -1;
  std::vector<EnsTile> t(count);
  for (long k = 0; k < count; ++k) t[k] = EnsTile{sys[k], i0[k]};
  return check_tiles(t, L, static_cast<size_t>(cap)) ? 1 : 0;
}

int epc_without_tpl2(int cfg) { return without_tpl2(cfg); }

int epc_config(int cfg, int* wpb, int* tpl2_min_n, int* xcd, int* split) {
  if (cfg < 0 || cfg >= ODEHIP_ENS_TILE_CONFIGS) return -1;
  *wpb = kEnsTileConfigs[cfg].wpb;
  *tpl2_min_n = kEnsTileConfigs[cfg].tpl2_min_n;
  *xcd = kEnsTileConfigs[cfg].xcd;
  *split = kEnsTileConfigs[cfg].split;
  return 0;
}

int epc_n_unsplit() { return ODEHIP_ENS_TILE_UNSPLIT; }
int epc_split_config(int xcd, int split) { return split_config(xcd, split); }
int epc_chunk() { return kEnsChunk; }
long epc_fill_waves() { return kEnsFillWaves; }

// tail_config over the systems `running` marks (null: all); -1: bad input.
int epc_tail_config(int n, const int* sizes, const char* running, int cfg, int exact) {
  Layout L;
  if (!layout_of(n, sizes, &L) || cfg < 0 || cfg >= ODEHIP_ENS_TILE_CONFIGS) return -1;
  std::vector<char> live;
  if (running) live.assign(running, running + n);
  return tail_config(L, running ? &live : nullptr, cfg, exact != 0);
}

// Arena plan: 18 byte offsets in Plan's field order (sys, groups, tiles,
// groups_act, st, n_active, xb0, xb1, ks, ytmp, kfsal, sc0, partA, partB,
// trace, dense_t, dense_out, total).
int epc_plan(int n, const int* sizes, int S, long trace_cap, int with_state, long n_dense,
             unsigned long* out) {
  Layout L;
  if (!layout_of(n, sizes, &L)) return -1;
  const Plan p = plan(L, tiles_capacity(L), S, trace_cap, with_state != 0, n_dense);
  static_assert(sizeof(Plan) == 18 * sizeof(size_t));
  std::memcpy(out, &p, sizeof p);
  return 0;
}

// Table cache across two calls on one arena: the key the first call (plan
// `a`) leaves behind, asked by the second (plan `b`) for configuration cfg.
// Returns 1 when the second call would skip its upload.
int epc_table_key_hit(int n, const int* sizes, int cfg, int S_a, int with_state_a, int S_b,
                      int with_state_b) {
  Layout L;
  if (!layout_of(n, sizes, &L)) return -1;
  const size_t cap = tiles_capacity(L);
  const Plan a = plan(L, cap, S_a, 0, with_state_a != 0);
  const Plan b = plan(L, cap, S_b, 0, with_state_b != 0);
  static char arena;
  TableKey key;
  key.set(cfg, &arena, a, L);
  return key.matches(cfg, &arena, b, L) ? 1 : 0;
}

// A key set with plan `a` asked with the same plan but its tile region moved
// by `shift` bytes (as a plan with other offsets would have it).
int epc_table_key_hit_shifted(int n, const int* sizes, int cfg, long shift) {
  Layout L;
  if (!layout_of(n, sizes, &L)) return -1;
  const Plan a = plan(L, tiles_capacity(L), 7, 0, true);
  Plan b = a;
  b.tiles += static_cast<size_t>(shift);
  static char arena;
  TableKey key;
  key.set(cfg, &arena, a, L);
  return key.matches(cfg, &arena, b, L) ? 1 : 0;
}

}  // extern "C"
