// The library context (odehip_context of odehip.h) and the error helpers
// shared by the C-ABI translation units.
#pragma once

#include <cstring>
#include <string>

#include "internal.hpp"

namespace odehip {
struct EnsembleWorkspace;
void destroy_ensemble_workspace(EnsembleWorkspace* w);
}  // namespace odehip

struct odehip_context;
namespace odehip {
// Grows the context's scratch buffer to at least n doubles (kept across
// calls; synchronises the stream before replacing it).
int ensure_scratch(odehip_context* ctx, size_t n);
}  // namespace odehip

struct GraphKey {
  int tableau = -1, rhs_kind = -1, variant = -1, k = 0, fsal = -1;
  size_t n_var = 0;
  double eps2 = 0.0;
  float near2 = 0.0f;
  const void* mass = nullptr;
  const void* atol = nullptr;
  const void* rtol = nullptr;
  const void* ws = nullptr;
  int controller = -1;
  const void* mlp = nullptr;
  // times, trace, states and dense-output capacities: they place the
  // buffers inside the workspace
  long caps[4] = {-1, -1, -1, -1};
  bool operator==(const GraphKey& o) const {
    return tableau == o.tableau && rhs_kind == o.rhs_kind &&
           variant == o.variant && k == o.k && fsal == o.fsal &&
           n_var == o.n_var && std::memcmp(&eps2, &o.eps2, sizeof eps2) == 0 &&
           std::memcmp(&near2, &o.near2, sizeof near2) == 0 &&
           mass == o.mass && atol == o.atol && rtol == o.rtol && ws == o.ws &&
           controller == o.controller && mlp == o.mlp &&
           std::memcmp(caps, o.caps, sizeof caps) == 0;
  }
};

struct odehip_context {
  int device = 0;
  hipStream_t own = nullptr;
  hipStream_t stream = nullptr;  // current stream (own or external)
  std::string err;
  // mixed RHS: pairs with r^2 < near_c / n are computed in FP64
  double near_c = ODEHIP_MIXED_NEAR_DEFAULT;
  // solve workspace
  double* ws = nullptr;
  size_t ws_cap = 0;
  odehip::SolverState* st = nullptr;
  // small scratch for reductions
  double* scratch = nullptr;
  size_t scratch_cap = 0;
  // this device's copy of the allocation-free norm's static array (odehip_abi.hip)
  double* noalloc_scratch = nullptr;
  // pinned host staging
  int* h_done = nullptr;           // 2 slots
  odehip::SolverState* h_st = nullptr;
  double* h_scalar = nullptr;      // 8 doubles
  hipEvent_t ev[2] = {nullptr, nullptr};
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  hipEvent_t t_start = nullptr, t_stop = nullptr;  // timing events
  // graph cache
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  GraphKey key;
  // ensemble solves (ensemble_abi.hip)
  odehip::EnsembleWorkspace* ens = nullptr;
};

inline int fail(odehip_context* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  return code;
}

inline int hip_fail(odehip_context* ctx, hipError_t e, const char* where) {
  return fail(ctx, ODEHIP_ERR_HIP,
              std::string(where) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(ctx, expr)                                   \
  do {                                                       \
    hipError_t _e = (expr);                                  \
    if (_e != hipSuccess) return hip_fail((ctx), _e, #expr); \
  } while (0)

// Makes the context's device current on the calling thread: allocations,
// copies and launches go to the current device, so a thread may hold contexts
// on several devices.  A null context is left to the entry point's checks.
inline int use_device(odehip_context* ctx) {
  if (!ctx) return ODEHIP_OK;
  int cur = -1;
  if (hipGetDevice(&cur) == hipSuccess && cur == ctx->device) return ODEHIP_OK;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  return ODEHIP_OK;
}

namespace odehip {
// Checks a tableau id (built in or registered) and makes a registered one
// readable by the kernels on the context's device (uploaded once per device).
int use_tableau(odehip_context* ctx, int id);
}  // namespace odehip

#define ODEHIP_USE_DEVICE(ctx)            \
  do {                                    \
    const int _rc = use_device(ctx);      \
    if (_rc != ODEHIP_OK) return _rc;     \
  } while (0)

// Work stream for graph capture: the context's own stream, ordered after the
// caller's stream when an external (possibly legacy default) stream is set.
inline hipStream_t begin_work(odehip_context* ctx) {
  if (ctx->stream != ctx->own) {
    (void)hipEventRecord(ctx->ev_in, ctx->stream);
    (void)hipStreamWaitEvent(ctx->own, ctx->ev_in, 0);
  }
  return ctx->own;
}

inline void end_work(odehip_context* ctx) {
  if (ctx->stream != ctx->own) {
    (void)hipEventRecord(ctx->ev_out, ctx->own);
    (void)hipStreamWaitEvent(ctx->stream, ctx->ev_out, 0);
  }
}

