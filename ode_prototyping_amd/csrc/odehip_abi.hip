// C ABI of libodehip (declarations and reference citations in odehip.h):
// context and stream management, device memory, executor-level kernels and
// the device-resident solve driver.
//
// The solve driver replaces the host loop of RKEmbeddedParallel::integrate
// (RKEmbeddedParallel.hpp:35-80).  The attempt's kernels (S right-hand sides,
// S-1 stage combinations, the fused final/error/norm pass and the controller)
// are captured once into a HIP graph holding K attempts; the graph is
// replayed until the device-side controller raises `done`.  Every kernel of a
// finished solve returns at once, so over-enqueueing is harmless.  The host
// only polls the done flag of the previous replay (one replay in flight), so
// the device never waits on the host between attempts.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "context.hpp"
#include "mlp.hpp"

using namespace odehip;

int odehip::ensure_scratch(odehip_context* ctx, size_t n) {
  if (ctx->scratch_cap >= n) return ODEHIP_OK;
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  if (ctx->scratch) HIP_TRY(ctx, hipFree(ctx->scratch));
  ctx->scratch = nullptr;
  ctx->scratch_cap = 0;
  size_t cap = std::max<size_t>(n, 4096);
  HIP_TRY(ctx, hipMalloc(&ctx->scratch, cap * sizeof(double)));
  ctx->scratch_cap = cap;
  return ODEHIP_OK;
}

// ------------------------------------------------------------ tableaux ----
// Registered tableaux are process-wide: host copies here, and each device's
// constant tables receive a slot on the first call that uses it there.
namespace {

struct TableauRegistry {
  std::mutex mu;
  Tableau tab[ODEHIP_MAX_TABLEAUS] = {};
  int count = ODEHIP_NUM_TABLEAUS;  // ids below this are valid
  std::vector<int> resident;        // per device: slots below this uploaded
};

TableauRegistry& registry() {
  static TableauRegistry r;
  return r;
}

}  // namespace

bool odehip::valid_tableau(int id) {
  if (id >= 0 && id < ODEHIP_NUM_TABLEAUS) return true;
  TableauRegistry& r = registry();
  std::lock_guard<std::mutex> lock(r.mu);
  return id >= ODEHIP_NUM_TABLEAUS && id < r.count;
}

const Tableau& odehip::host_tableau(int id) {
  // a registered slot is written before its id becomes valid and never again
  return id < ODEHIP_NUM_TABLEAUS ? kTableaus[id] : registry().tab[id];
}

int odehip::use_tableau(odehip_context* ctx, int id) {
  if (id >= 0 && id < ODEHIP_NUM_TABLEAUS) return ODEHIP_OK;
  TableauRegistry& r = registry();
  std::lock_guard<std::mutex> lock(r.mu);
  if (id < ODEHIP_NUM_TABLEAUS || id >= r.count)
    return fail(ctx, ODEHIP_ERR_INVALID, "bad tableau");
  const int d = ctx->device;
  if (static_cast<int>(r.resident.size()) <= d)
    r.resident.resize(d + 1, ODEHIP_NUM_TABLEAUS);
  for (int& next = r.resident[d]; next <= id; ++next) {
    HIP_TRY(ctx, upload_rk_tableau(next, r.tab[next]));
    HIP_TRY(ctx, upload_ens_tableau(next, r.tab[next]));
  }
  return ODEHIP_OK;
}

// ---------------------------------------------- allocation-free norm ----
// Group partials (from element 8) and result (element 0) of
// odehip_rk_norm_noalloc: a file-scope device array, so that call allocates
// nothing.  Every device holds its own copy until the process ends; host code
// uses the address hipGetSymbolAddress reports for a device, never the
// array's name (which on the host is only a stand-in symbol).
constexpr size_t kNoAllocSlots = 8 + ODEHIP_NOALLOC_MAX_N_VAR / ODEHIP_REDUCE_GROUP;
static __device__ __align__(256) double scratch[kNoAllocSlots];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};

// The array's address on ctx's device, resolved once per device (that
// device current: odehip_create has just selected it).
hipError_t resolve_noalloc_scratch(odehip_context* ctx) {
  if (ctx->device < 0 || ctx->device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[ctx->device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch[ctx->device] = static_cast<double*>(p);
  }
  ctx->noalloc_scratch = g_scratch[ctx->device];
  return hipSuccess;
}

}  // namespace

namespace {

int ensure_ws(odehip_context* ctx, size_t n) {
  if (ctx->ws_cap >= n) return ODEHIP_OK;
  HIP_TRY(ctx, hipDeviceSynchronize());
  if (ctx->exec) {
    (void)hipGraphExecDestroy(ctx->exec);
    ctx->exec = nullptr;
  }
  if (ctx->graph) {
    (void)hipGraphDestroy(ctx->graph);
    ctx->graph = nullptr;
  }
  ctx->key = GraphKey{};
  if (ctx->ws) HIP_TRY(ctx, hipFree(ctx->ws));
  ctx->ws = nullptr;
  ctx->ws_cap = 0;
  HIP_TRY(ctx, hipMalloc(&ctx->ws, n * sizeof(double)));
  ctx->ws_cap = n;
  return ODEHIP_OK;
}

int check_ode(odehip_context* ctx, const odehip_ode* ode) {
  if (!ode) return fail(ctx, ODEHIP_ERR_INVALID, "ode is null");
  if (ode->n_var == 0) return fail(ctx, ODEHIP_ERR_INVALID, "n_var is 0");
  if (ode->rhs_kind == ODEHIP_RHS_NBODY) {
    if (ode->n_var % 6 != 0)
      return fail(ctx, ODEHIP_ERR_INVALID, "N-body n_var must be a multiple of 6");
    if (ode->n_var / 6 > (1u << 26))
      return fail(ctx, ODEHIP_ERR_INVALID, "too many particles");
    if (ode->nbody_variant < 0 || ode->nbody_variant > 2)
      return fail(ctx, ODEHIP_ERR_INVALID, "bad nbody_variant");
  } else if (ode->rhs_kind == ODEHIP_RHS_VANDERPOL) {
    if (ode->n_var != 2) return fail(ctx, ODEHIP_ERR_INVALID, "Van der Pol has n_var 2");
  } else if (ode->rhs_kind != ODEHIP_RHS_EXPONENTIAL) {
    return fail(ctx, ODEHIP_ERR_INVALID, "unknown rhs_kind");
  }
  return ODEHIP_OK;
}

// Lays out the solve workspace and returns its length in doubles.
size_t layout(SolveBufs* b, double* base, SolverState* st, size_t n, int S,
              long times_cap, long trace_cap, long states_cap, long n_dense) {
  const size_t ng = static_cast<size_t>(n_groups_for(n));
  size_t off = 0;
  auto take = [&](size_t len) {
    double* p = base ? base + off : nullptr;
    off += (len + 31) & ~static_cast<size_t>(31);  // 256-byte alignment
    return p;
  };
  b->st = st;
  b->xb0 = take(n);
  b->xb1 = take(n);
  b->ks = take(static_cast<size_t>(S) * n);
  b->ytmp0 = take(n);
  b->ytmp1 = take(n);
  b->kfsal = take(n);
  b->sc0 = take(n);
  b->partA = take(ng);
  b->partB = take(ng);
  b->times = times_cap > 0 ? take(static_cast<size_t>(times_cap)) : nullptr;
  b->trace = trace_cap > 0 ? take(3 * static_cast<size_t>(trace_cap)) : nullptr;
  b->states = states_cap > 0 ? take(static_cast<size_t>(states_cap) * n) : nullptr;
  b->dense_t = n_dense > 0 ? take(static_cast<size_t>(n_dense)) : nullptr;
  b->dense_out = n_dense > 0 ? take(static_cast<size_t>(n_dense) * n) : nullptr;
  b->n_dense = n_dense;
  b->times_cap = times_cap;
  b->trace_cap = trace_cap;
  b->states_cap = states_cap;
  return off;
}

int prepare_bufs(odehip_context* ctx, int tableau, const odehip_ode* ode,
                 const double* atol, const double* rtol, long times_cap,
                 long trace_cap, long states_cap, long n_dense, SolveBufs* b) {
  const size_t n = ode->n_var;
  const int S = host_tableau(tableau).S;
  size_t need = layout(b, nullptr, ctx->st, n, S, times_cap, trace_cap, states_cap, n_dense);
  int rc = ensure_ws(ctx, need);
  if (rc) return rc;
  layout(b, ctx->ws, ctx->st, n, S, times_cap, trace_cap, states_cap, n_dense);
  b->atol = atol;
  b->rtol = rtol;
  b->n_var = n;
  b->tableau = tableau;
  b->rhs_kind = ode->rhs_kind;
  b->nbody_variant = ode->nbody_variant;
  b->n_particles = static_cast<int>(n / 6);
  b->eps2 = ode->softening * ode->softening;
  b->near2 = mixed_near2(ctx->near_c, b->n_particles);
  b->mass = ode->masses;
  return ODEHIP_OK;
}

}  // namespace

extern "C" {

const char* odehip_version(void) { return "odehip 1.0 gfx950"; }

int odehip_create(int device, odehip_context** out) {
  if (!out) return ODEHIP_ERR_INVALID;
  *out = nullptr;
  auto* ctx = new (std::nothrow) odehip_context();
  if (!ctx) return ODEHIP_ERR_NOMEM;
  ctx->device = device;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->own, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(&ctx->st, sizeof(SolverState));
  if (e == hipSuccess) e = hipHostMalloc(&ctx->h_done, 2 * sizeof(int));
  if (e == hipSuccess) e = hipHostMalloc(&ctx->h_st, sizeof(SolverState));
  if (e == hipSuccess) e = hipHostMalloc(&ctx->h_scalar, 8 * sizeof(double));
  if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->ev[0], hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->ev[1], hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->ev_in, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->ev_out, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreate(&ctx->t_start);
  if (e == hipSuccess) e = hipEventCreate(&ctx->t_stop);
  if (e == hipSuccess) e = resolve_noalloc_scratch(ctx);
  if (e != hipSuccess) {
    odehip_destroy(ctx);
    return ODEHIP_ERR_HIP;
  }
  ctx->stream = ctx->own;
  *out = ctx;
  return ODEHIP_OK;
}

int odehip_destroy(odehip_context* ctx) {
  if (!ctx) return ODEHIP_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->own) (void)hipStreamSynchronize(ctx->own);
  if (ctx->exec) (void)hipGraphExecDestroy(ctx->exec);
  if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
  if (ctx->ens) destroy_ensemble_workspace(ctx->ens);
  if (ctx->ws) (void)hipFree(ctx->ws);
  if (ctx->scratch) (void)hipFree(ctx->scratch);
  if (ctx->st) (void)hipFree(ctx->st);
  if (ctx->h_done) (void)hipHostFree(ctx->h_done);
  if (ctx->h_st) (void)hipHostFree(ctx->h_st);
  if (ctx->h_scalar) (void)hipHostFree(ctx->h_scalar);
  for (auto& ev : ctx->ev)
    if (ev) (void)hipEventDestroy(ev);
  if (ctx->ev_in) (void)hipEventDestroy(ctx->ev_in);
  if (ctx->ev_out) (void)hipEventDestroy(ctx->ev_out);
  if (ctx->t_start) (void)hipEventDestroy(ctx->t_start);
  if (ctx->t_stop) (void)hipEventDestroy(ctx->t_stop);
  if (ctx->own) (void)hipStreamDestroy(ctx->own);
  delete ctx;
  return ODEHIP_OK;
}

const char* odehip_last_error(const odehip_context* ctx) {
  return ctx ? ctx->err.c_str() : "null context";
}

int odehip_set_stream(odehip_context* ctx, void* hip_stream) {
  if (!ctx) return ODEHIP_ERR_INVALID;
  ctx->stream = hip_stream ? static_cast<hipStream_t>(hip_stream) : ctx->own;
  return ODEHIP_OK;
}

int odehip_set_near_field(odehip_context* ctx, double coefficient) {
  if (!ctx) return ODEHIP_ERR_INVALID;
  if (!(coefficient >= 0.0) || !std::isfinite(coefficient))
    return fail(ctx, ODEHIP_ERR_INVALID, "near-field coefficient must be finite and >= 0");
  ctx->near_c = coefficient;
  return ODEHIP_OK;
}

int odehip_get_near_field(const odehip_context* ctx, double* coefficient) {
  if (!ctx || !coefficient) return ODEHIP_ERR_INVALID;
  *coefficient = ctx->near_c;
  return ODEHIP_OK;
}

int odehip_synchronize(odehip_context* ctx) {
  ODEHIP_USE_DEVICE(ctx);
  if (!ctx) return ODEHIP_ERR_INVALID;
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->own));
  return ODEHIP_OK;
}

int odehip_timer_start(odehip_context* ctx) {
  ODEHIP_USE_DEVICE(ctx);
  if (!ctx) return ODEHIP_ERR_INVALID;
  HIP_TRY(ctx, hipEventRecord(ctx->t_start, ctx->stream));
  return ODEHIP_OK;
}

int odehip_timer_stop(odehip_context* ctx, double* elapsed_ms) {
  ODEHIP_USE_DEVICE(ctx);
  if (!ctx || !elapsed_ms) return ODEHIP_ERR_INVALID;
  HIP_TRY(ctx, hipEventRecord(ctx->t_stop, ctx->stream));
  HIP_TRY(ctx, hipEventSynchronize(ctx->t_stop));
  float ms = 0.0f;
  HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->t_start, ctx->t_stop));
  *elapsed_ms = ms;
  return ODEHIP_OK;
}

int odehip_probe_fp64(odehip_context* ctx, double* tflops, double* clock_mhz) {
  ODEHIP_USE_DEVICE(ctx);
  if (!ctx || !tflops || !clock_mhz) return ODEHIP_ERR_INVALID;
  const int blocks = 1024, threads = 1024, iters = 20000;
  const size_t need = static_cast<size_t>(blocks) * threads + 2;
  int rc = ensure_scratch(ctx, need);
  if (rc) return rc;
  long long* clk = reinterpret_cast<long long*>(ctx->scratch + static_cast<size_t>(blocks) * threads);
  double best = 0.0, mhz = 0.0;
  for (int rep = 0; rep < 3; ++rep) {
    HIP_TRY(ctx, hipEventRecord(ctx->t_start, ctx->stream));
    HIP_TRY(ctx, launch_probe_fp64(ctx->scratch, clk, blocks, iters, ctx->stream));
    HIP_TRY(ctx, hipEventRecord(ctx->t_stop, ctx->stream));
    HIP_TRY(ctx, hipEventSynchronize(ctx->t_stop));
    float ms = 0.0f;
    HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->t_start, ctx->t_stop));
    long long h[2];
    HIP_TRY(ctx, hipMemcpy(h, clk, sizeof h, hipMemcpyDeviceToHost));
    const double fl = 2.0 * blocks * threads * static_cast<double>(iters) * 8.0;
    const double tf = fl / (ms * 1e-3) / 1e12;
    if (tf > best) {
      best = tf;
      mhz = h[1] > 0 ? static_cast<double>(h[0]) / (static_cast<double>(h[1]) / 100.0) : 0.0;
    }
  }
  *tflops = best;
  *clock_mhz = mhz;
  return ODEHIP_OK;
}

// ------------------------------------------------------------------ memory --
int odehip_alloc(odehip_context* ctx, size_t n, double** out) {
  ODEHIP_USE_DEVICE(ctx);
  if (!ctx || !out) return fail(ctx, ODEHIP_ERR_INVALID, "null argument");
  *out = nullptr;
  hipError_t e = hipMalloc(out, std::max<size_t>(n, 1) * sizeof(double));
  if (e != hipSuccess) return fail(ctx, ODEHIP_ERR_NOMEM, hipGetErrorString(e));
  // CudaState() zero-fills (CudaState.cuh:15-18).
  e = launch_fill(*out, 0.0, n, ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "fill");
  return ODEHIP_OK;
}

int odehip_free(odehip_context* ctx, double* p) {
  ODEHIP_USE_DEVICE(ctx);
  if (!ctx) return ODEHIP_ERR_INVALID;
  if (!p) return ODEHIP_OK;
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  HIP_TRY(ctx, hipFree(p));
  return ODEHIP_OK;
}

int odehip_copy_h2d(odehip_context* ctx, double* dst, const double* src, size_t n) {
  ODEHIP_USE_DEVICE(ctx);
  if (!ctx || (n && (!dst || !src))) return fail(ctx, ODEHIP_ERR_INVALID, "null argument");
  HIP_TRY(ctx, hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return ODEHIP_OK;
}

int odehip_copy_d2h(odehip_context* ctx, double* dst, const double* src, size_t n) {
  ODEHIP_USE_DEVICE(ctx);
  if (!ctx || (n && (!dst || !src))) return fail(ctx, ODEHIP_ERR_INVALID, "null argument");
  HIP_TRY(ctx, hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return ODEHIP_OK;
}

int odehip_copy_d2d(odehip_context* ctx, double* dst, const double* src, size_t n) {
  ODEHIP_USE_DEVICE(ctx);
  if (!ctx || (n && (!dst || !src))) return fail(ctx, ODEHIP_ERR_INVALID, "null argument");
  HIP_TRY(ctx, hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
  return ODEHIP_OK;
}

int odehip_fill(odehip_context* ctx, double* dst, double v, size_t n) {
  ODEHIP_USE_DEVICE(ctx);
  if (!ctx || (n && !dst)) return fail(ctx, ODEHIP_ERR_INVALID, "null argument");
  HIP_TRY(ctx, launch_fill(dst, v, n, ctx->stream));
  return ODEHIP_OK;
}

// --------------------------------------------------------- executor ops ----
int odehip_nbody_rhs(odehip_context* ctx, const double* x, double* dxdt,
                     size_t n, double softening, const double* masses,
                     int variant) {
  ODEHIP_USE_DEVICE(ctx);
  if (!ctx || !x || !dxdt) return fail(ctx, ODEHIP_ERR_INVALID, "null argument");
  if (n == 0 || n > (1u << 26)) return fail(ctx, ODEHIP_ERR_INVALID, "bad particle count");
  if (variant < 0 || variant > 2) return fail(ctx, ODEHIP_ERR_INVALID, "bad variant");
  NBodyArgs a{};
  a.y = x;
  a.k = dxdt;
  a.mass = masses;
  a.n = static_cast<int>(n);
  a.eps2 = softening * softening;
  a.near2 = mixed_near2(ctx->near_c, a.n);
  HIP_TRY(ctx, launch_nbody_rhs(a, variant, ctx->stream));
  return ODEHIP_OK;
}

int odehip_nbody_acc(odehip_context* ctx, const double* pos, double* acc,
                     size_t n, double softening, const double* masses,
                     int variant) {
  ODEHIP_USE_DEVICE(ctx);
  if (!ctx || !pos || !acc) return fail(ctx, ODEHIP_ERR_INVALID, "null argument");
  if (n == 0 || n > (1u << 26)) return fail(ctx, ODEHIP_ERR_INVALID, "bad particle count");
  // The RHS kernels read a full [pos | vel] state and write [vel | acc]; stage
  // pos with zero velocities in scratch and extract acc.
  int rc = ensure_scratch(ctx, 12 * n);
  if (rc) return rc;
  double* xs = ctx->scratch;
  double* ks = ctx->scratch + 6 * n;
  HIP_TRY(ctx, hipMemcpyAsync(xs, pos, 3 * n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
  HIP_TRY(ctx, launch_fill(xs + 3 * n, 0.0, 3 * n, ctx->stream));
  rc = odehip_nbody_rhs(ctx, xs, ks, n, softening, masses, variant);
  if (rc) return rc;
  HIP_TRY(ctx, hipMemcpyAsync(acc, ks + 3 * n, 3 * n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
  return ODEHIP_OK;
}

int odehip_ode_rhs(odehip_context* ctx, int kind, const double* x, double* dxdt,
                   size_t n_var) {
  ODEHIP_USE_DEVICE(ctx);
  if (!ctx || !x || !dxdt) return fail(ctx, ODEHIP_ERR_INVALID, "null argument");
  if (kind == ODEHIP_RHS_NBODY) {
    if (n_var % 6) return fail(ctx, ODEHIP_ERR_INVALID, "n_var % 6 != 0");
    return odehip_nbody_rhs(ctx, x, dxdt, n_var / 6, 0.0, nullptr, ODEHIP_NBODY_EXACT);
  }
  if (kind == ODEHIP_RHS_VANDERPOL && n_var != 2)
    return fail(ctx, ODEHIP_ERR_INVALID, "Van der Pol has n_var 2");
  hipError_t e = launch_simple_rhs(kind, x, nullptr, nullptr, nullptr, dxdt,
                                   nullptr, 0, n_var, ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "ode_rhs");
  return ODEHIP_OK;
}

int odehip_tableau_register(odehip_context* ctx, int n_stages, int p, int pt,
                            const double* a, const double* b, const double* bt,
                            int* id) {
  ODEHIP_USE_DEVICE(ctx);
  if (!ctx) return ODEHIP_ERR_INVALID;
  if (!a || !b || !bt || !id) return fail(ctx, ODEHIP_ERR_INVALID, "null argument");
  const int S = n_stages;
  if (S < 2 || S > ODEHIP_MAX_STAGES)
    return fail(ctx, ODEHIP_ERR_INVALID, "n_stages must be in [2, ODEHIP_MAX_STAGES]");
  if (p < 1 || p > 20 || pt < 1 || pt > 20)
    return fail(ctx, ODEHIP_ERR_INVALID, "orders p, pt must be in [1, 20]");
  Tableau T{};
  T.S = S;
  T.p = p;
  T.pt = pt;
  for (int s = 1; s < S; ++s)
    for (int j = 0; j < s; ++j) {
      const double v = a[(s - 1) * (S - 1) + j];
      if (!std::isfinite(v)) return fail(ctx, ODEHIP_ERR_INVALID, "non-finite coefficient");
      T.a[s - 1][j] = v;
    }
  for (int j = 0; j < S; ++j) {
    if (!std::isfinite(b[j]) || !std::isfinite(bt[j]))
      return fail(ctx, ODEHIP_ERR_INVALID, "non-finite coefficient");
    T.b[j] = b[j];
    T.db[j] = b[j] - bt[j];  // as the reference's constexpr db (RKEmbeddedParallel.hpp:108-114)
  }
  bool fsal = b[S - 1] == 0.0;
  for (int j = 0; j < S - 1 && fsal; ++j) fsal = T.a[S - 2][j] == b[j];
  T.fsal = fsal ? 1 : 0;
  int found = -1;
  for (int k = 0; k < ODEHIP_NUM_TABLEAUS && found < 0; ++k)
    if (std::memcmp(&kTableaus[k], &T, sizeof T) == 0) found = k;
  if (found < 0) {
    TableauRegistry& r = registry();
    std::lock_guard<std::mutex> lock(r.mu);
    for (int k = ODEHIP_NUM_TABLEAUS; k < r.count && found < 0; ++k)
      if (std::memcmp(&r.tab[k], &T, sizeof T) == 0) found = k;
    if (found < 0) {
      if (r.count >= ODEHIP_MAX_TABLEAUS)
        return fail(ctx, ODEHIP_ERR_INVALID, "no free tableau slot");
      r.tab[r.count] = T;
      found = r.count++;
    }
  }
  if (const int rc = use_tableau(ctx, found)) return rc;
  *id = found;
  return ODEHIP_OK;
}

int odehip_tableau_info(int id, int* n_stages, int* p, int* pt) {
  if (!valid_tableau(id)) return ODEHIP_ERR_INVALID;
  const Tableau& T = host_tableau(id);
  if (n_stages) *n_stages = T.S;
  if (p) *p = T.p;
  if (pt) *pt = T.pt;
  return ODEHIP_OK;
}

int odehip_rk_stage(odehip_context* ctx, int tableau, int stage, double dt,
                    double* temp, const double* ks, const double* x0,
                    size_t n_var) {
  ODEHIP_USE_DEVICE(ctx);
  if (!ctx || !temp || !ks || !x0) return fail(ctx, ODEHIP_ERR_INVALID, "null argument");
  if (const int trc = use_tableau(ctx, tableau)) return trc;
  if (stage < 1 || stage >= host_tableau(tableau).S)
    return fail(ctx, ODEHIP_ERR_INVALID, "stage out of range");
  HIP_TRY(ctx, launch_rk_stage_host_dt(tableau, stage, dt, temp, ks, x0, n_var, ctx->stream));
  return ODEHIP_OK;
}

int odehip_rk_update_state_and_error(odehip_context* ctx, int tableau,
                                     double dt, double* x, double* err,
                                     const double* ks, const double* x0,
                                     size_t n_var) {
  ODEHIP_USE_DEVICE(ctx);
  if (!ctx || !x || !err || !ks || !x0) return fail(ctx, ODEHIP_ERR_INVALID, "null argument");
  if (const int trc = use_tableau(ctx, tableau)) return trc;
  HIP_TRY(ctx, launch_rk_update_host_dt(tableau, dt, x, err, ks, x0, n_var, ctx->stream));
  return ODEHIP_OK;
}

int odehip_rk_update_error_target(odehip_context* ctx, double* sc,
                                  const double* atol, const double* rtol,
                                  const double* x, const double* x0,
                                  size_t n_var) {
  ODEHIP_USE_DEVICE(ctx);
  if (!ctx || !sc || !atol || !rtol || !x || !x0)
    return fail(ctx, ODEHIP_ERR_INVALID, "null argument");
  HIP_TRY(ctx, launch_rk_target(sc, atol, rtol, x, x0, n_var, ctx->stream));
  return ODEHIP_OK;
}

int odehip_rk_norm(odehip_context* ctx, const double* v, const double* sc,
                   size_t n_var, double* host_result) {
  ODEHIP_USE_DEVICE(ctx);
  if (!ctx || !v || !sc || !host_result) return fail(ctx, ODEHIP_ERR_INVALID, "null argument");
  if (n_var == 0) return fail(ctx, ODEHIP_ERR_INVALID, "n_var is 0");
  const int ng = n_groups_for(n_var);
  int rc = ensure_scratch(ctx, static_cast<size_t>(ng) + 8);
  if (rc) return rc;
  HIP_TRY(ctx, launch_norm_partials(v, nullptr, sc, n_var, ctx->scratch + 8, ctx->stream));
  HIP_TRY(ctx, launch_fold_norm(ctx->scratch + 8, ng, n_var, ctx->scratch, ctx->stream));
  HIP_TRY(ctx, hipMemcpyAsync(ctx->h_scalar, ctx->scratch, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  *host_result = ctx->h_scalar[0];
  return ODEHIP_OK;
}

int odehip_rk_norm_noalloc(odehip_context* ctx, const double* v, const double* sc,
                           size_t n_var, double* host_result) {
  ODEHIP_USE_DEVICE(ctx);
  if (!ctx || !v || !sc || !host_result) return fail(ctx, ODEHIP_ERR_INVALID, "null argument");
  if (n_var == 0) return fail(ctx, ODEHIP_ERR_INVALID, "n_var is 0");
  if (n_var > ODEHIP_NOALLOC_MAX_N_VAR)
    return fail(ctx, ODEHIP_ERR_INVALID, "n_var exceeds ODEHIP_NOALLOC_MAX_N_VAR");
  if (!ctx->noalloc_scratch) return fail(ctx, ODEHIP_ERR_HIP, "static scratch not resolved");
  // the kernels of odehip_rk_norm; every partial and the result are written
  // before they are read, so nothing needs clearing between calls
  double* part = ctx->noalloc_scratch + 8;
  const int ng = n_groups_for(n_var);
  HIP_TRY(ctx, launch_norm_partials(v, nullptr, sc, n_var, part, ctx->stream));
  HIP_TRY(ctx, launch_fold_norm(part, ng, n_var, ctx->noalloc_scratch, ctx->stream));
  HIP_TRY(ctx, hipMemcpyAsync(ctx->h_scalar, ctx->noalloc_scratch, sizeof(double),
                              hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  *host_result = ctx->h_scalar[0];
  return ODEHIP_OK;
}

int odehip_rk_compute_error_target(odehip_context* ctx, double* sc,
                                   const double* x0, const double* atol,
                                   const double* rtol, size_t n_var) {
  ODEHIP_USE_DEVICE(ctx);
  if (!ctx || !sc || !x0 || !atol || !rtol) return fail(ctx, ODEHIP_ERR_INVALID, "null argument");
  HIP_TRY(ctx, launch_init_target(sc, x0, atol, rtol, n_var, ctx->stream));
  return ODEHIP_OK;
}

int odehip_rk_euler_step(odehip_context* ctx, double* x1, const double* x0,
                         const double* f0, double dt0, size_t n_var) {
  ODEHIP_USE_DEVICE(ctx);
  if (!ctx || !x1 || !x0 || !f0) return fail(ctx, ODEHIP_ERR_INVALID, "null argument");
  HIP_TRY(ctx, launch_euler_host_dt(x1, x0, f0, dt0, n_var, ctx->stream));
  return ODEHIP_OK;
}

int odehip_rk_difference(odehip_context* ctx, double* df, const double* f0,
                         size_t n_var) {
  ODEHIP_USE_DEVICE(ctx);
  if (!ctx || !df || !f0) return fail(ctx, ODEHIP_ERR_INVALID, "null argument");
  HIP_TRY(ctx, launch_difference(df, f0, n_var, ctx->stream));
  return ODEHIP_OK;
}

int odehip_rk_initial_step(odehip_context* ctx, int tableau, int rhs_kind,
                           const double* x0, const double* atol,
                           const double* rtol, size_t n_var, double softening,
                           const double* masses, int variant, double* host_dt) {
  ODEHIP_USE_DEVICE(ctx);
  if (!ctx || !x0 || !atol || !rtol || !host_dt)
    return fail(ctx, ODEHIP_ERR_INVALID, "null argument");
  if (const int trc = use_tableau(ctx, tableau)) return trc;
  odehip_ode ode{rhs_kind, variant, n_var, softening, masses};
  int rc = check_ode(ctx, &ode);
  if (rc) return rc;
  SolveBufs b{};
  rc = prepare_bufs(ctx, tableau, &ode, atol, rtol, 0, 0, 0, 0, &b);
  if (rc) return rc;
  hipStream_t s = begin_work(ctx);
  HIP_TRY(ctx, enqueue_solve_init(b, x0, 0.0, 1.0, 0, 0, 0, s));
  HIP_TRY(ctx, enqueue_initial_step(b, s));
  HIP_TRY(ctx, hipMemcpyAsync(ctx->h_st, ctx->st, sizeof(SolverState), hipMemcpyDeviceToHost, s));
  HIP_TRY(ctx, hipStreamSynchronize(s));
  end_work(ctx);
  *host_dt = ctx->h_st->dt_initial;
  return ODEHIP_OK;
}

// ---------------------------------------------------------------- solve ----
int odehip_rk_solve(odehip_context* ctx, int tableau, const odehip_ode* ode,
                    const odehip_solve_options* opt, double* x,
                    odehip_solve_stats* stats) {
  ODEHIP_USE_DEVICE(ctx);
  if (!ctx) return ODEHIP_ERR_INVALID;
  if (!opt || !x) return fail(ctx, ODEHIP_ERR_INVALID, "null argument");
  if (const int trc = use_tableau(ctx, tableau)) return trc;
  int rc = check_ode(ctx, ode);
  if (rc) return rc;
  if (!opt->atol || !opt->rtol) return fail(ctx, ODEHIP_ERR_INVALID, "atol/rtol null");
  if (opt->max_accepted < 0 || opt->max_attempts < 0 || opt->times_cap < 0 ||
      opt->trace_cap < 0 || opt->states_cap < 0)
    return fail(ctx, ODEHIP_ERR_INVALID, "negative budget or capacity");
  if (opt->controller < ODEHIP_CONTROLLER_REFERENCE ||
      opt->controller > ODEHIP_CONTROLLER_LEARNED)
    return fail(ctx, ODEHIP_ERR_INVALID, "bad controller");
  if (opt->controller == ODEHIP_CONTROLLER_LEARNED &&
      (!opt->mlp || ode->rhs_kind != ODEHIP_RHS_NBODY))
    return fail(ctx, ODEHIP_ERR_INVALID,
                "the learned controller needs a network and an N-body problem");
  if (opt->controller == ODEHIP_CONTROLLER_LEARNED && mlp_device(opt->mlp) != ctx->device)
    return fail(ctx, ODEHIP_ERR_INVALID, "the network lives on another device");
  const long times_cap = opt->times ? opt->times_cap : 0;
  const long trace_cap = opt->trace ? opt->trace_cap : 0;
  const long states_cap = opt->states ? opt->states_cap : 0;
  const long n_dense = opt->n_dense;
  if (n_dense < 0) return fail(ctx, ODEHIP_ERR_INVALID, "negative n_dense");
  if (n_dense > 0) {
    if (tableau != ODEHIP_TABLEAU_DOP853 && tableau != ODEHIP_TABLEAU_DOPRI5)
      return fail(ctx, ODEHIP_ERR_INVALID, "dense output needs DOP853 or DOPRI5");
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
  SolveBufs b{};
  rc = prepare_bufs(ctx, tableau, ode, opt->atol, opt->rtol, times_cap,
                    trace_cap, states_cap, n_dense, &b);
  if (rc) return rc;
  b.controller = opt->controller;
  b.mlp = opt->controller == ODEHIP_CONTROLLER_LEARNED ? mlp_infer_ptr(opt->mlp) : nullptr;
  const int S = host_tableau(tableau).S;
  const size_t n = ode->n_var;

  // attempts per replay: aim for several ms of work per replay
  int K = opt->attempts_per_launch;
  if (K == 0) {
    const double work = ode->rhs_kind == ODEHIP_RHS_NBODY
                            ? static_cast<double>(S) * (n / 6.0) * (n / 6.0)
                            : static_cast<double>(S) * n;
    K = work > 4.0e9 ? 2 : (work > 1.0e8 ? 8 : 32);
  }

  hipStream_t s = begin_work(ctx);
  if (n_dense > 0)
    HIP_TRY(ctx, hipMemcpyAsync(b.dense_t, opt->dense_times, n_dense * sizeof(double),
                                hipMemcpyHostToDevice, s));
  HIP_TRY(ctx, enqueue_solve_init(b, x, opt->t0, opt->tf, opt->max_accepted,
                                  opt->max_attempts, opt->fsal ? 1 : 0, s));
  HIP_TRY(ctx, enqueue_initial_step(b, s));

  if (K > 0) {
    GraphKey key;
    key.tableau = tableau;
    key.rhs_kind = ode->rhs_kind;
    key.variant = ode->nbody_variant;
    key.k = K;
    key.fsal = opt->fsal ? 1 : 0;
    key.n_var = n;
    key.eps2 = b.eps2;
    key.near2 = b.near2;
    key.mass = ode->masses;
    key.atol = opt->atol;
    key.rtol = opt->rtol;
    key.ws = ctx->ws;
    // times/trace pointers live in ws; their presence changes the layout:
    key.ws = reinterpret_cast<const char*>(ctx->ws) + (b.times ? 1 : 0) + (b.trace ? 2 : 0) +
             (b.states ? 4 : 0);
    key.controller = b.controller;
    key.mlp = b.mlp;
    key.caps[0] = times_cap;
    key.caps[1] = trace_cap;
    key.caps[2] = states_cap;
    key.caps[3] = n_dense;
    if (!ctx->exec || !(key == ctx->key)) {
      // a replay of the cached graph may still run (e.g. after a failed call)
      HIP_TRY(ctx, hipStreamSynchronize(s));
      if (ctx->exec) {
        (void)hipGraphExecDestroy(ctx->exec);
        ctx->exec = nullptr;
      }
      if (ctx->graph) {
        (void)hipGraphDestroy(ctx->graph);
        ctx->graph = nullptr;
      }
      HIP_TRY(ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      hipError_t ce = hipSuccess;
      for (int a = 0; a < K && ce == hipSuccess; ++a) ce = enqueue_attempt(b, s);
      hipGraph_t g = nullptr;
      hipError_t ee = hipStreamEndCapture(s, &g);
      if (ce != hipSuccess || ee != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return ce != hipSuccess ? hip_fail(ctx, ce, "capture attempt")
                                : hip_fail(ctx, ee, "hipStreamEndCapture");
      }
      ctx->graph = g;
      HIP_TRY(ctx, hipGraphInstantiate(&ctx->exec, g, nullptr, nullptr, 0));
      ctx->key = key;
    }
  }

  const long host_cap_attempts =
      opt->max_attempts > 0 ? opt->max_attempts : (1L << 40);
  long launched = 0;
  int slot = 0;
  bool have_prev = false;
  for (;;) {
    if (K > 0) {
      HIP_TRY(ctx, hipGraphLaunch(ctx->exec, s));
      launched += K;
    } else {
      const int direct = 16;
      for (int a = 0; a < direct; ++a) HIP_TRY(ctx, enqueue_attempt(b, s));
      launched += direct;
    }
    HIP_TRY(ctx, hipMemcpyAsync(&ctx->h_done[slot], &ctx->st->done, sizeof(int),
                                hipMemcpyDeviceToHost, s));
    HIP_TRY(ctx, hipEventRecord(ctx->ev[slot], s));
    if (have_prev) {
      HIP_TRY(ctx, hipEventSynchronize(ctx->ev[slot ^ 1]));
      if (ctx->h_done[slot ^ 1]) break;
    }
    have_prev = true;
    slot ^= 1;
    if (launched >= host_cap_attempts + 2L * std::max(K, 16)) break;
  }
  HIP_TRY(ctx, enqueue_solve_finish(b, x, s));
  HIP_TRY(ctx, hipMemcpyAsync(ctx->h_st, ctx->st, sizeof(SolverState),
                              hipMemcpyDeviceToHost, s));
  HIP_TRY(ctx, hipStreamSynchronize(s));
  const SolverState& hs = *ctx->h_st;
  const long n_times = std::min(hs.accepted + 1, times_cap);
  const long n_trace = std::min(hs.attempts, trace_cap);
  const long n_states = std::min(hs.accepted + 1, states_cap);
  const long n_dense_out = n_dense > 0 ? hs.dense_next : 0;
  if (n_dense_out > 0)
    HIP_TRY(ctx, hipMemcpy(opt->dense_states, b.dense_out,
                           static_cast<size_t>(n_dense_out) * n * sizeof(double),
                           hipMemcpyDeviceToHost));
  if (opt->states && n_states > 0)
    HIP_TRY(ctx, hipMemcpy(opt->states, b.states, static_cast<size_t>(n_states) * n * sizeof(double),
                           hipMemcpyDeviceToHost));
  if (opt->times && n_times > 0)
    HIP_TRY(ctx, hipMemcpy(opt->times, b.times, n_times * sizeof(double), hipMemcpyDeviceToHost));
  if (opt->trace && n_trace > 0)
    HIP_TRY(ctx, hipMemcpy(opt->trace, b.trace, 3 * n_trace * sizeof(double), hipMemcpyDeviceToHost));
  end_work(ctx);
  if (stats) {
    stats->accepted = hs.accepted;
    stats->rejected = hs.rejected;
    stats->attempts = hs.attempts;
    stats->rhs_evals = hs.rhs_evals;
    stats->n_times = n_times;
    stats->n_trace = n_trace;
    stats->status = hs.status;
    stats->t = hs.t;
    stats->dt = hs.dt;
    stats->dt_initial = hs.dt_initial;
    stats->last_err = hs.err;
    stats->learned_steps = hs.learned_steps;
    stats->n_states = n_states;
    stats->n_dense = n_dense_out;
    stats->dense_steps = hs.dense_steps;
    stats->dense_rhs_evals = hs.dense_rhs_evals;
  }
  if (hs.status == ODEHIP_STATUS_RUNNING)
    return fail(ctx, ODEHIP_ERR_INVALID, "solve did not finish within the attempt cap");
  return ODEHIP_OK;
}

// ------------------------------------------------------ conserved quantities --
int odehip_nbody_energy(odehip_context* ctx, const double* x, size_t n,
                        double softening, const double* masses,
                        double* host_out) {
  ODEHIP_USE_DEVICE(ctx);
  if (!ctx || !x || !host_out) return fail(ctx, ODEHIP_ERR_INVALID, "null argument");
  if (n == 0) return fail(ctx, ODEHIP_ERR_INVALID, "n is 0");
  const size_t ng = static_cast<size_t>(n_groups_for(n));
  int rc = ensure_scratch(ctx, 2 * n + 2 * (ng + 1) + 8);
  if (rc) return rc;
  double* per = ctx->scratch;
  double* sums = ctx->scratch + 2 * n;
  HIP_TRY(ctx, launch_nbody_energy(x, static_cast<int>(n), softening * softening,
                                   masses, per, ctx->stream));
  HIP_TRY(ctx, launch_sum_tree(per, n, sums, ctx->stream));
  HIP_TRY(ctx, launch_sum_tree(per + n, n, sums + ng + 1, ctx->stream));
  HIP_TRY(ctx, hipMemcpyAsync(ctx->h_scalar, sums, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipMemcpyAsync(ctx->h_scalar + 1, sums + ng + 1, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  host_out[0] = ctx->h_scalar[0];
  host_out[1] = ctx->h_scalar[1];
  return ODEHIP_OK;
}

int odehip_nbody_momentum(odehip_context* ctx, const double* x, size_t n,
                          const double* masses, double* host_out) {
  ODEHIP_USE_DEVICE(ctx);
  if (!ctx || !x || !host_out) return fail(ctx, ODEHIP_ERR_INVALID, "null argument");
  if (n == 0) return fail(ctx, ODEHIP_ERR_INVALID, "n is 0");
  const size_t ng = static_cast<size_t>(n_groups_for(n));
  int rc = ensure_scratch(ctx, 3 * n + 3 * (ng + 1) + 8);
  if (rc) return rc;
  double* per = ctx->scratch;
  double* sums = ctx->scratch + 3 * n;
  HIP_TRY(ctx, launch_momentum_terms(x, static_cast<int>(n), masses, per, ctx->stream));
  for (int c = 0; c < 3; ++c) {
    HIP_TRY(ctx, launch_sum_tree(per + c * n, n, sums + c * (ng + 1), ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(ctx->h_scalar + c, sums + c * (ng + 1), sizeof(double),
                                hipMemcpyDeviceToHost, ctx->stream));
  }
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  for (int c = 0; c < 3; ++c) host_out[c] = ctx->h_scalar[c];
  return ODEHIP_OK;
}

}  // extern "C"
