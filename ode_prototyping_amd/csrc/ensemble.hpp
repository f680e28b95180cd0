// Internal layout of an ensemble solve: many independent N-body systems of
// different sizes packed into one set of device vectors.
#pragma once

#include "internal.hpp"

namespace odehip {

// One system of the ensemble (device table).
struct EnsSys {
  long eoff;     // first state element (6 x the particles of earlier systems)
  int n;         // particles
  int gbase;     // first reduction group of this system
  int n_groups;  // its groups of ODEHIP_REDUCE_GROUP elements
  int pad;
};

// One right-hand-side work item: the targets [i0, i0 + 64 tpl) of system
// `sys` (tpl = 2 for systems with n >= tpl2_min_n, else 1).  sys < 0 marks
// an empty slot.
struct EnsTile {
  int sys;
  int i0;
};

// Device buffers and parameters of one ensemble solve.  State-sized vectors
// hold all systems packed (nv elements); ks holds S of them, nv apart.
struct EnsBufs {
  const EnsSys* sys;
  const int2* groups;    // element workgroups: (system, reduction group id)
  const EnsTile* tiles;  // RHS work items in dispatch order
  int n_tiles;           // slots: RHS workgroups x waves per workgroup
  SolverState* st;       // one per system
  int* n_active;         // systems still running
  double* xb0;
  double* xb1;
  double* ks;
  double* ytmp;
  double* kfsal;
  double* sc0;
  double* partA;
  double* partB;
  double* trace;  // n_sys x trace_cap x 3, or null
  long trace_cap;
  const double* mass;  // packed, or null
  const void* mlp;     // learned controller's MlpInfer (device), or null
  int controller;      // ODEHIP_CONTROLLER_*
  double atol, rtol;
  size_t nv;
  int n_sys;
  int n_groups;  // entries of `groups` (workgroups of an element launch)
  int tableau;
  int variant;
  int wpb;         // waves per RHS workgroup
  int tpl2_min_n;  // systems with n >= this use 128-target tiles
  int split;       // > 1: waves sharing one tile's partners
  double eps2;
  double near_c;  // mixed variant: pairs with FP32 r^2 < near_c / n in FP64
  const double* dense_t;  // DOPRI5 dense output: n_dense times, or null
  double* dense_out;      // n_dense x nv (packed rows)
  long n_dense;
};

// Tile configurations: waves per workgroup, the size from which systems use
// 128-target tiles, whether tiles of one system stay on one XCD, and the
// waves that share one tile's partners (1: a wave per tile; the last
// ODEHIP_ENS_TILE_CONFIGS - ODEHIP_ENS_TILE_UNSPLIT configurations).  The
// autotuner times every one; tail_config moves a solve's running systems to
// a shared-tile one when they are few.
struct EnsTileConfig {
  int wpb;
  int tpl2_min_n;
  int xcd;
  int split;
};
constexpr int kEnsNoTpl2 = 1 << 30;
constexpr EnsTileConfig kEnsTileConfigs[ODEHIP_ENS_TILE_CONFIGS] = {
    {4, kEnsNoTpl2, 0, 1}, {4, kEnsNoTpl2, 1, 1}, {4, 1024, 0, 1}, {4, 1024, 1, 1},
    {4, 256, 0, 1},        {4, 256, 1, 1},        {1, kEnsNoTpl2, 0, 1}, {1, kEnsNoTpl2, 1, 1},
    {1, 1024, 0, 1},       {1, 1024, 1, 1},       {1, 256, 0, 1},  {1, 256, 1, 1},
    {1, kEnsNoTpl2, 0, 2}, {1, kEnsNoTpl2, 1, 2}, {1, kEnsNoTpl2, 0, 4},
    {1, kEnsNoTpl2, 1, 4}, {1, kEnsNoTpl2, 0, 8}, {1, kEnsNoTpl2, 1, 8}};
constexpr int kEnsDefaultTileConfig = 1;
// Partners per chunk of the fast and mixed sums (a multiple of the mixed
// variant's FP32 flush interval, 32).
constexpr int kEnsChunk = 128;
// Waves that give each of the 1024 SIMDs (256 CUs x 4) four, the depth that
// hides the RHS arithmetic's latency.
constexpr long kEnsFillWaves = 4096;

// y == null: each system's current x0 buffer (solve); else a packed state.
// st == null in b: plain evaluation of every system.
hipError_t ens_launch_rhs(const EnsBufs& b, const double* y, double* k,
                          int stage0, hipStream_t s);
hipError_t ens_enqueue_init(const EnsBufs& b, const double* x_in, double t0,
                            double tf, long max_acc, long max_att, int fsal,
                            int controller, hipStream_t s);
hipError_t ens_enqueue_initial_step(const EnsBufs& b, hipStream_t s);
hipError_t ens_enqueue_round(const EnsBufs& b, hipStream_t s);
hipError_t ens_enqueue_copy_out(const EnsBufs& b, double* x_out,
                                hipStream_t s);

}  // namespace odehip
