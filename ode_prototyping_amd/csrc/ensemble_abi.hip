// C ABI of the ensemble solve (declarations in odehip.h): layout and tile
// tables, the tile autotuner, and the lockstep-round driver.
//
// The driver mirrors odehip_rk_solve: initial step for every system, then K
// lockstep rounds (one attempt of each active system per round) captured in
// a HIP graph and replayed until the device-side count of running systems
// reaches zero; the host polls that count one replay behind.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <numeric>
#include <vector>

#include "context.hpp"
#include "ensemble.hpp"
#include "ensemble_plan.hpp"
#include "mlp.hpp"

namespace odehip {

struct EnsembleWorkspace {
  char* dev = nullptr;  // device arena
  size_t cap = 0;       // bytes
  int* h_active = nullptr;  // pinned, 2 slots
  hipEvent_t ev[2] = {nullptr, nullptr};
  hipEvent_t t0 = nullptr, t1 = nullptr;
  // cached graph of K rounds and the arguments it was captured with
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  EnsBufs graph_bufs{};
  int graph_k = 0;
  // tables currently on the device
  TableKey tables;
  // autotuner cache
  std::vector<int> tuned_sizes;
  int tuned_variant = -1;
  bool tuned_mass = false;
  double tuned_eps2 = -1.0;
  double tuned_near = -1.0;
  int tuned_cfg = -1;
  double tuned_ms[ODEHIP_ENS_TILE_CONFIGS] = {};
};

void destroy_ensemble_workspace(EnsembleWorkspace* w) {
  if (!w) return;
  if (w->exec) (void)hipGraphExecDestroy(w->exec);
  if (w->graph) (void)hipGraphDestroy(w->graph);
  if (w->dev) (void)hipFree(w->dev);
  if (w->h_active) (void)hipHostFree(w->h_active);
  for (auto& e : w->ev)
    if (e) (void)hipEventDestroy(e);
  if (w->t0) (void)hipEventDestroy(w->t0);
  if (w->t1) (void)hipEventDestroy(w->t1);
  delete w;
}

}  // namespace odehip

using namespace odehip;

namespace {

int make_layout(odehip_context* ctx, int n_systems, const int* n_particles,
                Layout* L) {
  const char* why = build_layout(n_systems, n_particles, L);
  return why ? fail(ctx, ODEHIP_ERR_INVALID, why) : ODEHIP_OK;
}

int workspace(odehip_context* ctx, size_t bytes, EnsembleWorkspace** out) {
  if (!ctx->ens) {
    auto* w = new (std::nothrow) EnsembleWorkspace();
    if (!w) return fail(ctx, ODEHIP_ERR_NOMEM, "ensemble workspace");
    ctx->ens = w;
    HIP_TRY(ctx, hipHostMalloc(reinterpret_cast<void**>(&w->h_active), 2 * sizeof(int)));
    for (auto& e : w->ev) HIP_TRY(ctx, hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HIP_TRY(ctx, hipEventCreate(&w->t0));
    HIP_TRY(ctx, hipEventCreate(&w->t1));
  }
  EnsembleWorkspace* w = ctx->ens;
  if (w->cap < bytes) {
    HIP_TRY(ctx, hipDeviceSynchronize());
    if (w->exec) (void)hipGraphExecDestroy(w->exec);
    if (w->graph) (void)hipGraphDestroy(w->graph);
    w->exec = nullptr;
    w->graph = nullptr;
    if (w->dev) HIP_TRY(ctx, hipFree(w->dev));
    w->dev = nullptr;
    w->cap = 0;
    w->tables.invalidate();
    HIP_TRY(ctx, hipMalloc(reinterpret_cast<void**>(&w->dev), bytes));
    w->cap = bytes;
  }
  *out = w;
  return ODEHIP_OK;
}

// Uploads the system/group tables and the tile list of `cfg` (skipped when
// the device already holds exactly these tables at these offsets).
int upload_tables(odehip_context* ctx, EnsembleWorkspace* w, const Plan& p,
                  const Layout& L, size_t tiles_cap, int cfg, int* n_tiles,
                  hipStream_t s) {
  const std::vector<EnsTile> tiles = make_tiles(L, kEnsTileConfigs[cfg]);
  if (const char* why = check_tiles(tiles, L, tiles_cap))
    return fail(ctx, ODEHIP_ERR_INVALID, why);
  *n_tiles = static_cast<int>(tiles.size());
  if (w->tables.matches(cfg, w->dev, p, L)) return ODEHIP_OK;
  HIP_TRY(ctx, hipStreamSynchronize(s));  // no kernel may still read them
  HIP_TRY(ctx, hipMemcpy(w->dev + p.sys, L.sys.data(), L.sys.size() * sizeof(EnsSys),
                         hipMemcpyHostToDevice));
  HIP_TRY(ctx, hipMemcpy(w->dev + p.groups, L.groups.data(),
                         L.groups.size() * sizeof(int2), hipMemcpyHostToDevice));
  HIP_TRY(ctx, hipMemcpy(w->dev + p.tiles, tiles.data(), tiles.size() * sizeof(EnsTile),
                         hipMemcpyHostToDevice));
  w->tables.set(cfg, w->dev, p, L);
  return ODEHIP_OK;
}

EnsBufs make_bufs(EnsembleWorkspace* w, const Plan& p, const Layout& L,
                  bool with_state) {
  EnsBufs b;
  std::memset(&b, 0, sizeof b);  // padding too: the graph cache compares bytes
  char* d = w->dev;
  b.sys = reinterpret_cast<const EnsSys*>(d + p.sys);
  b.groups = reinterpret_cast<const int2*>(d + p.groups);
  b.tiles = reinterpret_cast<const EnsTile*>(d + p.tiles);
  b.n_active = reinterpret_cast<int*>(d + p.n_active);
  if (with_state) {
    b.st = reinterpret_cast<SolverState*>(d + p.st);
    b.xb0 = reinterpret_cast<double*>(d + p.xb0);
    b.xb1 = reinterpret_cast<double*>(d + p.xb1);
    b.ks = reinterpret_cast<double*>(d + p.ks);
    b.ytmp = reinterpret_cast<double*>(d + p.ytmp);
    b.kfsal = reinterpret_cast<double*>(d + p.kfsal);
    b.sc0 = reinterpret_cast<double*>(d + p.sc0);
    b.partA = reinterpret_cast<double*>(d + p.partA);
    b.partB = reinterpret_cast<double*>(d + p.partB);
  }
  b.nv = L.nv;
  b.n_sys = static_cast<int>(L.sys.size());
  b.n_groups = L.n_groups;
  return b;
}

void set_tile_cfg(EnsBufs* b, int cfg, int n_tiles) {
  b->wpb = kEnsTileConfigs[cfg].wpb;
  b->tpl2_min_n = kEnsTileConfigs[cfg].tpl2_min_n;
  b->split = kEnsTileConfigs[cfg].split;
  b->n_tiles = n_tiles;
}

bool tune_cache_hit(const EnsembleWorkspace* w, const Layout& L, int variant,
                    bool mass, double eps2, double near_c) {
  return w->tuned_cfg >= 0 && w->tuned_variant == variant && w->tuned_mass == mass &&
         std::memcmp(&w->tuned_eps2, &eps2, sizeof eps2) == 0 &&
         std::memcmp(&w->tuned_near, &near_c, sizeof near_c) == 0 && w->tuned_sizes == L.sizes;
}

// Times one right-hand side of the state `y` into `k` for every candidate
// configuration (one warm-up, best of two).  Among the candidates within 3 %
// of the fastest it keeps the one with the most target tiles (then the
// fastest of those): on the full ensemble 128-target tiles often tie with
// 64-target ones, but they halve the waves once compaction leaves a few
// large systems running.
int autotune(odehip_context* ctx, EnsembleWorkspace* w, const Plan& p,
             const Layout& L, size_t tiles_cap, EnsBufs b, const double* y,
             double* k, hipStream_t s) {
  double best = 1e300;
  int best_cfg = kEnsDefaultTileConfig;
  for (int c = 0; c < ODEHIP_ENS_TILE_CONFIGS; ++c) {
    int n_tiles = 0;
    int rc = upload_tables(ctx, w, p, L, tiles_cap, c, &n_tiles, s);
    if (rc) return rc;
    set_tile_cfg(&b, c, n_tiles);
    HIP_TRY(ctx, ens_launch_rhs(b, y, k, 0, s));
    double ms_best = 1e300;
    for (int r = 0; r < 2; ++r) {
      HIP_TRY(ctx, hipEventRecord(w->t0, s));
      HIP_TRY(ctx, ens_launch_rhs(b, y, k, 0, s));
      HIP_TRY(ctx, hipEventRecord(w->t1, s));
      HIP_TRY(ctx, hipEventSynchronize(w->t1));
      float ms = 0.0f;
      HIP_TRY(ctx, hipEventElapsedTime(&ms, w->t0, w->t1));
      ms_best = std::min(ms_best, static_cast<double>(ms));
    }
    w->tuned_ms[c] = ms_best;
    if (ms_best < best) {
      best = ms_best;
      best_cfg = c;
    }
  }
  long most = -1;
  for (int c = 0; c < ODEHIP_ENS_TILE_CONFIGS; ++c) {
    if (!(w->tuned_ms[c] <= 1.03 * best)) continue;
    const long t = real_tiles(L, kEnsTileConfigs[c]);
    if (t > most || (t == most && w->tuned_ms[c] < w->tuned_ms[best_cfg])) {
      most = t;
      best_cfg = c;
    }
  }
  w->tuned_cfg = best_cfg;
  return ODEHIP_OK;
}

int check_common(odehip_context* ctx, int nbody_variant, double softening,
                 int tile_config) {
  if (nbody_variant < 0 || nbody_variant > 2)
    return fail(ctx, ODEHIP_ERR_INVALID, "bad nbody_variant");
  if (!(softening >= 0.0)) return fail(ctx, ODEHIP_ERR_INVALID, "bad softening");
  if (tile_config < ODEHIP_ENS_TILE_AUTO || tile_config >= ODEHIP_ENS_TILE_CONFIGS)
    return fail(ctx, ODEHIP_ERR_INVALID, "bad tile_config");
  return ODEHIP_OK;
}

}  // namespace

extern "C" {

int odehip_ensemble_rhs(odehip_context* ctx, int n_systems,
                        const int* n_particles, int nbody_variant,
                        double softening, const double* masses,
                        int tile_config, const double* x, double* dxdt,
                        double* ms_out) {
  ODEHIP_USE_DEVICE(ctx);
  if (!ctx) return ODEHIP_ERR_INVALID;
  if (!x || !dxdt) return fail(ctx, ODEHIP_ERR_INVALID, "null argument");
  int rc = check_common(ctx, nbody_variant, softening, tile_config);
  if (rc) return rc;
  Layout L;
  if ((rc = make_layout(ctx, n_systems, n_particles, &L))) return rc;
  const size_t tcap = tiles_capacity(L);
  const Plan p = plan(L, tcap, 0, 0, false);
  EnsembleWorkspace* w = nullptr;
  if ((rc = workspace(ctx, p.total, &w))) return rc;
  const double eps2 = softening * softening;
  const double near_c = nbody_variant == ODEHIP_NBODY_MIXED ? ctx->near_c : 0.0;
  int cfg = tile_config;
  if (cfg == ODEHIP_ENS_TILE_AUTO)
    cfg = tune_cache_hit(w, L, nbody_variant, masses != nullptr, eps2, near_c)
              ? w->tuned_cfg
              : kEnsDefaultTileConfig;
  hipStream_t s = begin_work(ctx);
  int n_tiles = 0;
  if ((rc = upload_tables(ctx, w, p, L, tcap, cfg, &n_tiles, s))) return rc;
  EnsBufs b = make_bufs(w, p, L, false);
  b.mass = masses;
  b.variant = nbody_variant;
  b.eps2 = eps2;
  b.near_c = near_c;
  set_tile_cfg(&b, cfg, n_tiles);
  HIP_TRY(ctx, hipEventRecord(w->t0, s));
  HIP_TRY(ctx, ens_launch_rhs(b, x, dxdt, 0, s));
  HIP_TRY(ctx, hipEventRecord(w->t1, s));
  if (ms_out) {
    HIP_TRY(ctx, hipEventSynchronize(w->t1));
    float ms = 0.0f;
    HIP_TRY(ctx, hipEventElapsedTime(&ms, w->t0, w->t1));
    *ms_out = ms;
  }
  end_work(ctx);
  return ODEHIP_OK;
}

int odehip_ensemble_solve(odehip_context* ctx, int tableau, int n_systems,
                          const int* n_particles, int nbody_variant,
                          double softening, const double* masses,
                          const odehip_ensemble_options* opt,
                          const double* x0, double* x_out,
                          odehip_ensemble_stats* stats,
                          odehip_system_stats* per_system) {
  ODEHIP_USE_DEVICE(ctx);
  if (!ctx) return ODEHIP_ERR_INVALID;
  if (!opt || !x0 || !x_out) return fail(ctx, ODEHIP_ERR_INVALID, "null argument");
  if (const int trc = use_tableau(ctx, tableau)) return trc;
  int rc = check_common(ctx, nbody_variant, softening, opt->tile_config);
  if (rc) return rc;
  if (opt->max_accepted < 0 || opt->max_attempts < 0 || opt->trace_cap < 0)
    return fail(ctx, ODEHIP_ERR_INVALID, "negative budget or capacity");
  if (opt->controller < ODEHIP_CONTROLLER_REFERENCE ||
      opt->controller > ODEHIP_CONTROLLER_LEARNED)
    return fail(ctx, ODEHIP_ERR_INVALID, "bad controller");
  if (opt->controller == ODEHIP_CONTROLLER_LEARNED && !opt->mlp)
    return fail(ctx, ODEHIP_ERR_INVALID, "the learned controller needs a network (mlp)");
  if (opt->controller == ODEHIP_CONTROLLER_LEARNED && mlp_device(opt->mlp) != ctx->device)
    return fail(ctx, ODEHIP_ERR_INVALID, "the network lives on another device");
  const long n_dense = opt->n_dense;
  if (n_dense < 0) return fail(ctx, ODEHIP_ERR_INVALID, "negative n_dense");
  if (n_dense > 0) {
    if (tableau != ODEHIP_TABLEAU_DOPRI5 && tableau != ODEHIP_TABLEAU_DOP853)
      return fail(ctx, ODEHIP_ERR_INVALID, "ensemble dense output needs DOPRI5 or DOP853");
    if (!opt->dense_times || !opt->dense_states)
      return fail(ctx, ODEHIP_ERR_INVALID, "dense_times / dense_states null");
    double prev = opt->t0;
    for (long r = 0; r < n_dense; ++r) {
      const double tr = opt->dense_times[r];
      if (!std::isfinite(tr) || !(tr > prev))
        return fail(ctx, ODEHIP_ERR_INVALID,
                    "dense_times must be finite and strictly increasing after t0");
      prev = tr;
    }
  }
  Layout L;
  if ((rc = make_layout(ctx, n_systems, n_particles, &L))) return rc;
  const int S = host_tableau(tableau).S;
  const long trace_cap = opt->trace ? opt->trace_cap : 0;
  const size_t tcap = tiles_capacity(L);
  const Plan p = plan(L, tcap, S, trace_cap, true, n_dense);
  EnsembleWorkspace* w = nullptr;
  if ((rc = workspace(ctx, p.total, &w))) return rc;
  const double eps2 = softening * softening;
  const double near_c = nbody_variant == ODEHIP_NBODY_MIXED ? ctx->near_c : 0.0;
  hipStream_t s = begin_work(ctx);

  EnsBufs b = make_bufs(w, p, L, true);
  b.trace = trace_cap > 0 ? reinterpret_cast<double*>(w->dev + p.trace) : nullptr;
  b.trace_cap = trace_cap;
  b.mass = masses;
  b.atol = opt->atol;
  b.rtol = opt->rtol;
  b.tableau = tableau;
  b.variant = nbody_variant;
  b.eps2 = eps2;
  b.near_c = near_c;
  b.controller = opt->controller;
  b.mlp = opt->controller == ODEHIP_CONTROLLER_LEARNED ? mlp_infer_ptr(opt->mlp) : nullptr;
  if (n_dense > 0) {
    b.dense_t = reinterpret_cast<const double*>(w->dev + p.dense_t);
    b.dense_out = reinterpret_cast<double*>(w->dev + p.dense_out);
    b.n_dense = n_dense;
    HIP_TRY(ctx, hipMemcpyAsync(w->dev + p.dense_t, opt->dense_times, n_dense * sizeof(double),
                                hipMemcpyHostToDevice, s));
    HIP_TRY(ctx, launch_fill(b.dense_out, NAN, static_cast<size_t>(n_dense) * L.nv, s));
  }

  // tile configuration
  int cfg = opt->tile_config;
  if (cfg == ODEHIP_ENS_TILE_AUTO) {
    if (!tune_cache_hit(w, L, nbody_variant, masses != nullptr, eps2, near_c)) {
      std::fill(w->tuned_ms, w->tuned_ms + ODEHIP_ENS_TILE_CONFIGS, 0.0);
      EnsBufs tb = b;
      tb.st = nullptr;  // plain evaluations of x0 into the k1 slot
      if ((rc = autotune(ctx, w, p, L, tcap, tb, x0, b.ks, s))) return rc;
      w->tuned_variant = nbody_variant;
      w->tuned_mass = masses != nullptr;
      w->tuned_eps2 = eps2;
      w->tuned_near = near_c;
      w->tuned_sizes = L.sizes;
    }
    cfg = w->tuned_cfg;
  }
  const bool tuned = tune_cache_hit(w, L, nbody_variant, masses != nullptr, eps2, near_c);
  if (opt->tune_ms) {
    for (int c = 0; c < ODEHIP_ENS_TILE_CONFIGS; ++c)
      opt->tune_ms[c] = tuned ? w->tuned_ms[c] : 0.0;
  }
  // a tile's partners are shared by several waves from the start when the
  // ensemble's tiles would leave SIMDs idle
  const bool exact = nbody_variant == ODEHIP_NBODY_EXACT;
  int run_cfg = tail_config(L, nullptr, cfg, exact);
  int n_tiles = 0;
  if ((rc = upload_tables(ctx, w, p, L, tcap, run_cfg, &n_tiles, s))) return rc;
  set_tile_cfg(&b, run_cfg, n_tiles);

  HIP_TRY(ctx, ens_enqueue_init(b, x0, opt->t0, opt->tf, opt->max_accepted,
                                opt->max_attempts, opt->fsal ? 1 : 0, opt->controller, s));
  HIP_TRY(ctx, ens_enqueue_initial_step(b, s));

  // Rounds run on lists of the running systems' workgroups: every system at
  // first, rebuilt (and the graph re-captured) each time the number of
  // running systems has fallen to 7/8 of the last list's, so the long tail of
  // a few slow systems does not launch workgroups for all the finished ones
  // (on 2048 systems 3/4 measured 1-1.5 % faster than 1/2, and 7/8 another
  // 0-1 %; same results).
  auto rounds_per_replay = [&](double sum_n2) {
    int k = opt->rounds_per_launch;
    if (k <= 0) {
      const double work = static_cast<double>(S) * sum_n2;
      k = work > 1.0e10 ? 2 : (work > 1.0e9 ? 4 : (work > 1.0e8 ? 8 : 32));
    }
    return k;
  };
  auto capture = [&](const EnsBufs& rb, int k) -> int {
    if (w->exec && w->graph_k == k && std::memcmp(&w->graph_bufs, &rb, sizeof rb) == 0)
      return ODEHIP_OK;
    // a replay of the cached graph may still run (e.g. after a failed call)
    HIP_TRY(ctx, hipStreamSynchronize(s));
    if (w->exec) (void)hipGraphExecDestroy(w->exec);
    if (w->graph) (void)hipGraphDestroy(w->graph);
    w->exec = nullptr;
    w->graph = nullptr;
    HIP_TRY(ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    hipError_t ce = hipSuccess;
    for (int r = 0; r < k && ce == hipSuccess; ++r) ce = ens_enqueue_round(rb, s);
    hipGraph_t g = nullptr;
    hipError_t ee = hipStreamEndCapture(s, &g);
    if (ce != hipSuccess || ee != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return ce != hipSuccess ? hip_fail(ctx, ce, "capture round")
                              : hip_fail(ctx, ee, "hipStreamEndCapture");
    }
    w->graph = g;
    HIP_TRY(ctx, hipGraphInstantiate(&w->exec, g, nullptr, nullptr, 0));
    w->graph_bufs = rb;
    w->graph_k = k;
    return ODEHIP_OK;
  };
  EnsBufs rb = b;  // run_cfg: the tile shape of the running systems
  int K = rounds_per_replay(L.work);
  if ((rc = capture(rb, K))) return rc;

  const long cap_rounds = opt->max_attempts > 0 ? opt->max_attempts : (1L << 40);
  long launched = 0;
  int slot = 0;
  bool have_prev = false;
  int compact_below = n_systems * 7 / 8;
  std::vector<SolverState> hst;
  for (;;) {
    HIP_TRY(ctx, hipGraphLaunch(w->exec, s));
    launched += K;
    HIP_TRY(ctx, hipMemcpyAsync(&w->h_active[slot], b.n_active, sizeof(int),
                                hipMemcpyDeviceToHost, s));
    HIP_TRY(ctx, hipEventRecord(w->ev[slot], s));
    if (launched >= cap_rounds + 2L * K) break;
    if (have_prev) {
      HIP_TRY(ctx, hipEventSynchronize(w->ev[slot ^ 1]));
      const int running = w->h_active[slot ^ 1];
      if (running == 0) break;
      if (running <= compact_below) {
        HIP_TRY(ctx, hipStreamSynchronize(s));
        hst.resize(L.sys.size());
        HIP_TRY(ctx, hipMemcpy(hst.data(), b.st, hst.size() * sizeof(SolverState),
                               hipMemcpyDeviceToHost));
        std::vector<char> live(L.sys.size());
        std::vector<int2> act;
        double sum_n2 = 0.0;
        int n_live = 0;
        for (size_t k = 0; k < L.sys.size(); ++k) {
          live[k] = hst[k].done ? 0 : 1;
          if (!live[k]) continue;
          ++n_live;
          sum_n2 += static_cast<double>(L.sys[k].n) * L.sys[k].n;
          for (int g = 0; g < L.sys[k].n_groups; ++g)
            act.push_back(L.groups[L.sys[k].gbase + g]);
        }
        if (n_live == 0) break;
        // once the running systems' tiles no longer give every SIMD four
        // waves: 64-target tiles, then partners shared by several waves per
        // tile; same results
        run_cfg = tail_config(L, &live, run_cfg, exact);
        const std::vector<EnsTile> tiles = make_tiles(L, kEnsTileConfigs[run_cfg], &live);
        if (const char* why = check_tiles(tiles, L, tcap))
          return fail(ctx, ODEHIP_ERR_INVALID, why);
        HIP_TRY(ctx, hipMemcpy(w->dev + p.groups_act, act.data(), act.size() * sizeof(int2),
                               hipMemcpyHostToDevice));
        HIP_TRY(ctx, hipMemcpy(w->dev + p.tiles, tiles.data(), tiles.size() * sizeof(EnsTile),
                               hipMemcpyHostToDevice));
        w->tables.invalidate();  // the tile region now holds the running systems' tiles
        rb.groups = reinterpret_cast<const int2*>(w->dev + p.groups_act);
        rb.n_groups = static_cast<int>(act.size());
        set_tile_cfg(&rb, run_cfg, static_cast<int>(tiles.size()));
        K = rounds_per_replay(sum_n2);
        if ((rc = capture(rb, K))) return rc;
        compact_below = n_live * 7 / 8;
        have_prev = false;
        slot = 0;
        continue;
      }
    }
    have_prev = true;
    slot ^= 1;
  }
  HIP_TRY(ctx, ens_enqueue_copy_out(b, x_out, s));
  std::vector<SolverState> hs(L.sys.size());
  HIP_TRY(ctx, hipMemcpyAsync(hs.data(), b.st, hs.size() * sizeof(SolverState),
                              hipMemcpyDeviceToHost, s));
  HIP_TRY(ctx, hipStreamSynchronize(s));
  if (opt->trace && trace_cap > 0)
    HIP_TRY(ctx, hipMemcpy(opt->trace, b.trace,
                           hs.size() * static_cast<size_t>(trace_cap) * 3 * sizeof(double),
                           hipMemcpyDeviceToHost));
  if (n_dense > 0)
    HIP_TRY(ctx, hipMemcpy(opt->dense_states, b.dense_out,
                           static_cast<size_t>(n_dense) * L.nv * sizeof(double),
                           hipMemcpyDeviceToHost));
  end_work(ctx);
  odehip_ensemble_stats agg{};
  agg.tile_config = cfg;
  agg.interactions = L.interactions;
  agg.rhs_ms = tuned ? w->tuned_ms[cfg] : 0.0;
  for (size_t k = 0; k < hs.size(); ++k) {
    const SolverState& h = hs[k];
    agg.rounds = std::max(agg.rounds, h.attempts);
    agg.accepted += h.accepted;
    agg.rejected += h.rejected;
    agg.attempts += h.attempts;
    agg.rhs_evals += h.rhs_evals;
    agg.learned_steps += h.learned_steps;
    agg.particle_steps += static_cast<double>(L.sys[k].n) * h.accepted;
    if (h.status == ODEHIP_STATUS_RUNNING) agg.n_unfinished += 1;
    if (per_system) {
      odehip_system_stats& o = per_system[k];
      o.accepted = h.accepted;
      o.rejected = h.rejected;
      o.attempts = h.attempts;
      o.rhs_evals = h.rhs_evals;
      o.n_trace = std::min(h.attempts, trace_cap);
      o.learned_steps = h.learned_steps;
      o.status = h.status;
      o.pad = 0;
      o.t = h.t;
      o.dt = h.dt;
      o.dt_initial = h.dt_initial;
      o.last_err = h.err;
      o.n_dense = n_dense > 0 ? h.dense_next : 0;
    }
  }
  if (stats) *stats = agg;
  if (agg.n_unfinished)
    return fail(ctx, ODEHIP_ERR_INVALID, "ensemble did not finish within the attempt cap");
  return ODEHIP_OK;
}

}  // extern "C"
