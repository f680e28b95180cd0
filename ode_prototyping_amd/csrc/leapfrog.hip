// Fixed-step leapfrog (velocity Verlet) of an N-body state on the device:
// ParallelLeapfrogIntegrator::integrate (LeapfrogParallel.hpp:8-41, kernels
// :57-73) with the library's N-body right-hand side for the accelerations
// (NBodyODE::acc, NBodyODE.hpp:41-46).  Per step: the position kernel, one
// RHS launch, the velocity kernel; the two acceleration buffers swap roles
// instead of the reference's copy acc0 = acc1.  The step times are computed
// on the host with the reference's own t += dt loop (no clip of the last
// step), so the number of steps and every saved time match it exactly.
#include <algorithm>
#include <cmath>
#include <vector>

#include "context.hpp"

using namespace odehip;

namespace {

constexpr int kBlock = 256;

// update_position_kernel (LeapfrogParallel.hpp:57-62).
__global__ void k_lf_position(double* x, const double* acc0, double dt,
                              size_t n_dof) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n_dof) x[i] += dt * (x[n_dof + i] + 0.5 * dt * acc0[i]);
}

// update_velocity_kernel (:64-68).
__global__ void k_lf_velocity(double* x, const double* acc0, const double* acc1,
                              double dt, size_t n_dof) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n_dof) x[n_dof + i] += 0.5 * dt * (acc0[i] + acc1[i]);
}

}  // namespace

extern "C" int odehip_leapfrog(odehip_context* ctx, const odehip_ode* ode,
                               double t0, double tf, double dt, double* x,
                               double* states, double* times, long states_cap,
                               odehip_leapfrog_stats* stats) {
  ODEHIP_USE_DEVICE(ctx);
  if (!ctx) return ODEHIP_ERR_INVALID;
  if (!ode || !x) return fail(ctx, ODEHIP_ERR_INVALID, "null argument");
  if (ode->rhs_kind != ODEHIP_RHS_NBODY || ode->n_var == 0 || ode->n_var % 6)
    return fail(ctx, ODEHIP_ERR_INVALID, "leapfrog needs an N-body state");
  if (ode->nbody_variant < 0 || ode->nbody_variant > 2)
    return fail(ctx, ODEHIP_ERR_INVALID, "bad nbody_variant");
  if (!(dt > 0.0) || !std::isfinite(t0) || !std::isfinite(tf))
    return fail(ctx, ODEHIP_ERR_INVALID, "dt must be positive and t0, tf finite");
  if (states_cap < 0) return fail(ctx, ODEHIP_ERR_INVALID, "negative capacity");
  // the reference's time loop, on the host
  std::vector<double> t_list{t0};
  for (double t = t0; t < tf;) {
    t += dt;
    t_list.push_back(t);
    if (t_list.size() > (1u << 30))
      return fail(ctx, ODEHIP_ERR_INVALID, "more than 2^30 steps");
  }
  const long n_steps = static_cast<long>(t_list.size()) - 1;
  const size_t n_var = ode->n_var;
  const size_t n_dof = n_var / 2;
  const long keep = std::min<long>(states_cap, n_steps + 1);
  // the context's scratch buffer (kept across calls: nothing is allocated
  // per call once it is large enough)
  const size_t need = 2 * n_var + (states ? static_cast<size_t>(keep) * n_var : 0);
  int rc = ensure_scratch(ctx, need);
  if (rc) return rc;
  double* buf = ctx->scratch;
  double* k0 = buf;           // [vel | acc] of the current positions
  double* k1 = buf + n_var;   // of the new positions
  double* sv = buf + 2 * n_var;
  hipStream_t s = ctx->stream;
  NBodyArgs a{};
  a.mass = ode->masses;
  a.n = static_cast<int>(n_var / 6);
  a.eps2 = ode->softening * ode->softening;
  a.near2 = mixed_near2(ctx->near_c, a.n);
  a.y = x;
  auto rhs = [&](double* k) {
    a.k = k;
    return launch_nbody_rhs(a, ode->nbody_variant, s);
  };
  const unsigned blocks = static_cast<unsigned>((n_dof + kBlock - 1) / kBlock);
  hipError_t e = hipSuccess;
  auto save = [&](long idx) {
    if (states && idx < keep)
      return hipMemcpyAsync(sv + static_cast<size_t>(idx) * n_var, x,
                            n_var * sizeof(double), hipMemcpyDeviceToDevice, s);
    return hipSuccess;
  };
  if ((e = save(0)) == hipSuccess) e = rhs(k0);
  for (long step = 0; step < n_steps && e == hipSuccess; ++step) {
    hipLaunchKernelGGL(k_lf_position, dim3(blocks), dim3(kBlock), 0, s, x, k0 + n_dof, dt,
                       n_dof);
    if ((e = rhs(k1)) != hipSuccess) break;
    hipLaunchKernelGGL(k_lf_velocity, dim3(blocks), dim3(kBlock), 0, s, x, k0 + n_dof,
                       k1 + n_dof, dt, n_dof);
    std::swap(k0, k1);
    if ((e = hipGetLastError()) != hipSuccess) break;
    e = save(step + 1);
  }
  if (e == hipSuccess && states && keep > 0)
    e = hipMemcpyAsync(states, sv, static_cast<size_t>(keep) * n_var * sizeof(double),
                       hipMemcpyDeviceToHost, s);
  hipError_t se = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_fail(ctx, e, "leapfrog");
  if (se != hipSuccess) return hip_fail(ctx, se, "leapfrog sync");
  if (times)
    for (long i = 0; i < keep; ++i) times[i] = t_list[i];
  if (stats) {
    stats->steps = n_steps;
    stats->rhs_evals = n_steps + 1;
    stats->t = t_list.back();
    stats->n_states = states ? keep : 0;
  }
  return ODEHIP_OK;
}
