/* odehip.h -- C ABI of the MI355X (gfx950) embedded Runge-Kutta library.
 *
 * The HIP execution policy for the adaptive embedded RK path of the
 * reference (noether17/ode_prototyping).  Every entry point is extern "C",
 * takes plain pointers and sizes, returns an int status (ODEHIP_OK == 0) and
 * never throws; odehip_last_error() gives the message of the last failure on
 * a context.  Pointers documented as "device" are HIP device allocations
 * (from odehip_alloc or any other HIP allocator on the context's device);
 * work is enqueued on the context's stream in call order.
 *
 * Each declaration names the reference code it replaces (file:line in the
 * reference tree).
 */
#ifndef ODEHIP_H
#define ODEHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ----------------------------------------------------------- status codes */
#define ODEHIP_OK 0
#define ODEHIP_ERR_INVALID (-1)     /* bad argument (null pointer, bad size) */
#define ODEHIP_ERR_HIP (-2)         /* a HIP runtime call failed */
#define ODEHIP_ERR_NOMEM (-3)       /* device or host allocation failed */
#define ODEHIP_ERR_UNSUPPORTED (-4) /* combination not implemented */

/* Tableaux (src/integrator/RKEmbeddedParallel/butcher_tableaus/BT*.hpp; DOP853
 * in the same format from NumericalRecipes/Dopr853_constants.hpp). */
#define ODEHIP_TABLEAU_HE21 0
#define ODEHIP_TABLEAU_RKF45 1
#define ODEHIP_TABLEAU_DOPRI5 2
#define ODEHIP_TABLEAU_DVERK 3
#define ODEHIP_TABLEAU_RKF78 4
#define ODEHIP_TABLEAU_DOP853 5
#define ODEHIP_NUM_TABLEAUS 6   /* built-in ids: 0 .. ODEHIP_NUM_TABLEAUS - 1 */
#define ODEHIP_MAX_TABLEAUS 32  /* registered ids: ODEHIP_NUM_TABLEAUS .. 31 */
#define ODEHIP_MAX_STAGES 13

/* Right-hand sides. */
#define ODEHIP_RHS_NBODY 0       /* src/ode/NBodyODE.hpp:31-39 */
#define ODEHIP_RHS_EXPONENTIAL 1 /* src/ode/ExponentialODE.hpp:14-18 */
#define ODEHIP_RHS_VANDERPOL 2   /* src/ode/VanDerPolODE.hpp:16-25 */

/* Arithmetic of the N-body pair term. */
#define ODEHIP_NBODY_EXACT 0 /* the reference's FP64 expression, correctly
                                rounded sqrt and divisions, partners summed
                                in increasing index: bit-identical to
                                NBodyODE with SingleThreadedExecutor */
#define ODEHIP_NBODY_FAST 1  /* FP64 with a refined hardware rsqrt and
                                parallel partial sums (~1 ulp per term) */
#define ODEHIP_NBODY_MIXED 2 /* displacement rounded to FP32, pair force in
                                FP32, accumulation in FP64; in single-system
                                calls near pairs (r^2 < c / n,
                                odehip_set_near_field) in FP64 */
/* Near-field coefficient c of the mixed variant: with masses ~1/n a pair
   closer than sqrt(c / n) is one whose single term is a sizeable part of a
   body's acceleration, where FP32's ~1e-7 relative error would dominate the
   solve's error at tight tolerances.  Off by default (all pairs FP32, the
   fastest form); ODEHIP_MIXED_NEAR_CLOSE is the setting for systems with
   close encounters. */
#define ODEHIP_MIXED_NEAR_DEFAULT 0.0
#define ODEHIP_MIXED_NEAR_CLOSE 20.0

/* Solve exit status (the reference only ever exits with REACHED_END). */
/* Step-size controllers.  REFERENCE is RKEmbeddedParallel.hpp:64-79
   (dt * (0.38 / err)^(1/(1+q)), clamps [0.33, 6], floor, clip to tf).  PI
   multiplies by (0.38/err)^(0.3/k) (err_prev/err)^(0.4/k), k = 1 + q
   (Gustafsson's PI.3.4, err_prev of the last accepted step), with the same
   clamps.  LEARNED asks a trained network for
   the ratio dt_new / dt and falls back to PI (see odehip_mlp_*).  Every
   controller accepts a step exactly when the embedded error norm is <= 1. */
#define ODEHIP_CONTROLLER_REFERENCE 0
#define ODEHIP_CONTROLLER_PI 1
#define ODEHIP_CONTROLLER_LEARNED 2

/* A trained step-size network (odehip_mlp_create); opaque. */
struct odehip_mlp;

#define ODEHIP_STATUS_RUNNING (-1)
#define ODEHIP_STATUS_REACHED_END 0  /* t >= tf */
#define ODEHIP_STATUS_MAX_ACCEPTED 1 /* accepted-step budget reached */
#define ODEHIP_STATUS_NONFINITE 2    /* error norm not finite */
#define ODEHIP_STATUS_MAX_ATTEMPTS 3 /* attempt budget exhausted */

/* The canonical order of every reduction the library performs (the error
 * norm): elements are taken in groups of 256 (zero padded); each group is the
 * sum of its four 64-element quads combined as (q0 + q1) + (q2 + q3), each quad
 * reduced by halving strides 32, 16, ..., 1 (v[i] += v[i + s]); the group
 * sums are then added in group order starting from 0.0.  (The reference's
 * order is executor-defined: src/par_exec/SingleThreadedExecutor.hpp:17-29,
 * ThreadPoolExecutor.hpp:81-114, CudaExecutor.cuh:21-78.) */
#define ODEHIP_REDUCE_GROUP 256

typedef struct odehip_context odehip_context;

/* --------------------------------------------------------------- context
 * Replaces the CUDA executor object (src/par_exec/CudaExecutor.cuh:80-126),
 * which used the legacy default stream; a context owns a HIP stream and the
 * scratch the reductions and the device-resident controller need, so no call
 * allocates per step (CudaExecutor.cuh:103-107,121-122 does).  A context
 * belongs to one device: every call on it first makes that device current on
 * the calling thread, and device buffers and networks passed to it must live
 * on that device (a network from another device is rejected). */
int odehip_create(int device, odehip_context** out);
int odehip_destroy(odehip_context* ctx);
const char* odehip_last_error(const odehip_context* ctx);
/* Use an external hipStream_t (e.g. torch's current stream); NULL restores
 * the context's own stream. */
int odehip_set_stream(odehip_context* ctx, void* hip_stream);
/* Near-field coefficient c of the mixed RHS on this context (no reference
 * counterpart: the reference has only FP64 pair terms, NBodyODE.hpp:89-98).
 * Pairs with FP32 r^2 < c / n are computed and summed in FP64; c = 0 gives
 * all pairs in FP32.  c must be finite and >= 0; the default is
 * ODEHIP_MIXED_NEAR_DEFAULT.  Applies to every later mixed-variant call on
 * the context: single systems (RHS, acc, solve, leapfrog) and ensembles
 * (odehip_ensemble_rhs / _solve, with n each system's own size). */
int odehip_set_near_field(odehip_context* ctx, double coefficient);
int odehip_get_near_field(const odehip_context* ctx, double* coefficient);
int odehip_synchronize(odehip_context* ctx);
/* Library version string, e.g. "odehip 1.0 gfx950". */
const char* odehip_version(void);
/* HIP-event timing on the context's stream (no reference counterpart; the
 * reference times with Google Benchmark, src/benchmark/n_body_benchmark.cu):
 * odehip_timer_start records a start event, odehip_timer_stop records a stop
 * event, waits for it and writes the elapsed milliseconds. */
int odehip_timer_start(odehip_context* ctx);
int odehip_timer_stop(odehip_context* ctx, double* elapsed_ms);
/* Diagnostic (no reference counterpart): sustained FP64 FMA throughput of
 * the device and its shader clock under that load, from a dependent-free FMA
 * kernel that fills every SIMD (s_memtime / s_memrealtime).  Writes TFLOP/s
 * and MHz. */
int odehip_probe_fp64(odehip_context* ctx, double* tflops, double* clock_mhz);

/* ----------------------------------------------------------------- memory
 * Backing store of HipState (replaces src/state/CudaState.cuh:15-50). */
int odehip_alloc(odehip_context* ctx, size_t n_doubles, double** dev_out);
int odehip_free(odehip_context* ctx, double* dev);
int odehip_copy_h2d(odehip_context* ctx, double* dev_dst, const double* host_src,
                    size_t n);
int odehip_copy_d2h(odehip_context* ctx, double* host_dst, const double* dev_src,
                    size_t n);
int odehip_copy_d2d(odehip_context* ctx, double* dev_dst, const double* dev_src,
                    size_t n);
/* src/state/StateUtils.hpp:5-16 (fill). */
int odehip_fill(odehip_context* ctx, double* dev, double value, size_t n);

/* ------------------------------------------------ executor-level kernels
 * One call per reference kernel launch; all pointers are device pointers. */

/* NBodyODE::operator() (src/ode/NBodyODE.hpp:31-39, kernels :51-56 and
 * :68-133): dxdt = [vel | acc(pos)], x = [pos xyz | vel xyz], n_var = 6 n.
 * masses may be NULL (unit masses, as the reference; otherwise the pair term
 * is multiplied by m_j after the division, the reference's commented-out mass
 * code NBodyODE.hpp:104-124).  G = 1. */
int odehip_nbody_rhs(odehip_context* ctx, const double* x, double* dxdt,
                     size_t n_particles, double softening, const double* masses,
                     int variant);
/* NBodyODE::acc (src/ode/NBodyODE.hpp:41-46): acc (3 n) from pos (3 n). */
int odehip_nbody_acc(odehip_context* ctx, const double* pos, double* acc,
                     size_t n_particles, double softening, const double* masses,
                     int variant);
/* ExponentialODE / VanDerPolODE (src/ode/ExponentialODE.hpp:14-18,
 * src/ode/VanDerPolODE.hpp:16-25). */
int odehip_ode_rhs(odehip_context* ctx, int rhs_kind, const double* x,
                   double* dxdt, size_t n_var);

/* A Butcher tableau in the reference's struct format
 * (butcher_tableaus/BTDOPRI5.hpp:5-26: a, b, bt, p, pt, n_stages), registered
 * for every context of the process.  Replaces the integrator's tableau type
 * argument (RKEmbeddedParallel<ButcherTableau>, RKEmbeddedParallel.hpp:11-24):
 * the returned id is taken wherever a tableau id is (stage, update, initial
 * step, device and ensemble solves; dense output stays limited to the
 * built-in DOPRI5 / DOP853).
 *   a   (n_stages - 1) x (n_stages - 1), row major: row s - 1 holds the
 *       weights of stage s (entries j >= s are ignored, zero in the
 *       reference's arrays)
 *   b   n_stages propagated weights (order p); bt n_stages embedded weights
 *       (order pt); the error weights b - bt are formed in double as
 *       RKEmbeddedParallel.hpp:108-114 does.
 * 2 <= n_stages <= ODEHIP_MAX_STAGES, 1 <= p, pt <= 20, finite coefficients.
 * Registering the same tableau again returns the same id; at most
 * ODEHIP_MAX_TABLEAUS - ODEHIP_NUM_TABLEAUS different ones per process. */
int odehip_tableau_register(odehip_context* ctx, int n_stages, int p, int pt,
                            const double* a, const double* b, const double* bt,
                            int* id);
/* Stage count and orders of a built-in or registered tableau. */
int odehip_tableau_info(int id, int* n_stages, int* p, int* pt);

/* detail::rk_stage_kernel (RKEmbeddedParallel.hpp:91-101): for stage >= 1,
 * temp = x0 + dt * sum_{j<stage} a[stage-1][j] * ks[j*n_var + .]. */
int odehip_rk_stage(odehip_context* ctx, int tableau, int stage, double dt,
                    double* temp, const double* ks, const double* x0,
                    size_t n_var);
/* detail::update_state_and_error_kernel (RKEmbeddedParallel.hpp:103-123). */
int odehip_rk_update_state_and_error(odehip_context* ctx, int tableau,
                                     double dt, double* x, double* err,
                                     const double* ks, const double* x0,
                                     size_t n_var);
/* detail::update_error_target_kernel (RKEmbeddedParallel.hpp:125-130). */
int odehip_rk_update_error_target(odehip_context* ctx, double* sc,
                                  const double* atol, const double* rtol,
                                  const double* x, const double* x0,
                                  size_t n_var);
/* detail::rk_norm (RKEmbeddedParallel.hpp:141-149) = sqrt(sum (v/sc)^2 / n)
 * in the canonical order above; blocks until the host value is available
 * (as CudaExecutor::transform_reduce does, CudaExecutor.cuh:117-119). */
int odehip_rk_norm(odehip_context* ctx, const double* v, const double* sc,
                   size_t n_var, double* host_result);
/* odehip_rk_norm without any allocation: no counterpart in the reference;
 * for callers that cannot allocate (code under graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same kernels and result bit for
 * bit; its group partials live in a static device array the library holds on
 * each device until the process ends (8 + ODEHIP_NOALLOC_MAX_N_VAR / 256
 * doubles, 512 KiB), so n_var > ODEHIP_NOALLOC_MAX_N_VAR is rejected
 * (ODEHIP_ERR_INVALID).  Every call on a device shares that array: calls that
 * overlap on one device (other contexts, other host threads) race. */
#define ODEHIP_NOALLOC_MAX_N_VAR (1 << 24)
int odehip_rk_norm_noalloc(odehip_context* ctx, const double* v, const double* sc,
                           size_t n_var, double* host_result);
/* detail::compute_error_target_kernel (RKEmbeddedParallel.hpp:151-157):
 * sc = |x0| * rtol + atol. */
int odehip_rk_compute_error_target(odehip_context* ctx, double* sc,
                                   const double* x0, const double* atol,
                                   const double* rtol, size_t n_var);
/* detail::euler_step_kernel (RKEmbeddedParallel.hpp:159-163):
 * x1 = x0 + f0 * dt0. */
int odehip_rk_euler_step(odehip_context* ctx, double* x1, const double* x0,
                         const double* f0, double dt0, size_t n_var);
/* detail::difference_kernel (RKEmbeddedParallel.hpp:165-168): df -= f0. */
int odehip_rk_difference(odehip_context* ctx, double* df, const double* f0,
                         size_t n_var);
/* detail::estimate_initial_step (RKEmbeddedParallel.hpp:170-200) on the
 * device; writes the step to *host_dt. */
int odehip_rk_initial_step(odehip_context* ctx, int tableau, int rhs_kind,
                           const double* x0, const double* atol,
                           const double* rtol, size_t n_var, double softening,
                           const double* masses, int variant, double* host_dt);

/* --------------------------------------------------- device-resident solve
 * RKEmbeddedParallel<BT>::integrate (RKEmbeddedParallel.hpp:13-81) with the
 * attempt loop, error norm, accept/reject and step-size controller kept on
 * the device: the host only enqueues captured graphs of attempts and reads a
 * completion flag; no per-step host round trip. */
typedef struct {
  int rhs_kind;          /* ODEHIP_RHS_* */
  int nbody_variant;     /* ODEHIP_NBODY_* (N-body only) */
  size_t n_var;          /* state length (6 n for N-body) */
  double softening;      /* N-body softening length */
  const double* masses;  /* device, n entries, or NULL (unit masses) */
} odehip_ode;

typedef struct {
  double t0, tf;
  const double* atol;    /* device, n_var entries (per component, as the
                            reference's StateType atol) */
  const double* rtol;    /* device, n_var entries */
  long max_accepted;     /* stop after this many accepted steps; 0 = none */
  long max_attempts;     /* stop after this many attempts; 0 = none */
  int fsal;              /* 1: reuse the last stage as the next first stage
                            when the tableau is FSAL (DOPRI5), and keep k1
                            across a rejection -- same results, fewer RHS
                            evaluations.  0: the reference's S evaluations
                            per attempt. */
  int attempts_per_launch; /* attempts per captured graph; 0 = automatic;
                              negative: no graph, attempts launched directly */
  double* times;         /* host, times_cap entries: t0 then t after each
                            accepted step (RawOutput::times()); may be NULL */
  long times_cap;
  double* trace;         /* host, trace_cap x 3: per attempt (err, dt,
                            accepted); may be NULL */
  long trace_cap;
  int controller;        /* ODEHIP_CONTROLLER_*; 0 = the reference's */
  const struct odehip_mlp* mlp; /* trained network (LEARNED only) */
  double* states;        /* host, states_cap x n_var: x0 then the state after
                            each accepted step (RawOutput::save_state,
                            RawOutput.hpp:15-19), kept on the device during
                            the solve; may be NULL */
  long states_cap;
  /* Dense output, DOP853 or DOPRI5 (StepperDopr853::prepare_dense /
   * dense_out, NumericalRecipes/StepperDopr853.hpp:200-289; StepperDopr5,
   * StepperDopr5.hpp:181-221; and the DenseOutput policy, Output.hpp:99-114):
   * the state at each of n_dense strictly increasing times after t0, from
   * the continuous extension of the first accepted step that ends strictly
   * after it -- steps are not shortened to hit them.  dense_times: host,
   * n_dense; dense_states: host, n_dense x n_var, rows written for the times
   * reached (stats.n_dense).  DOP853 costs three extra RHS evaluations per
   * step that covers an output time, plus one for f at the step's end, which
   * the next attempt reuses as its k1 when fsal is set; DOPRI5 costs none. */
  const double* dense_times;
  long n_dense;
  double* dense_states;
} odehip_solve_options;

typedef struct {
  long accepted, rejected, attempts, rhs_evals;
  long n_times;          /* entries written to times (accepted + 1, capped) */
  long n_trace;          /* entries written to trace */
  int status;            /* ODEHIP_STATUS_* */
  double t, dt, dt_initial, last_err;
  long learned_steps;    /* steps sized by the learned controller */
  long n_states;         /* states written (accepted + 1, capped) */
  long n_dense;          /* dense-output rows written */
  long dense_steps;      /* accepted steps that produced dense output */
  long dense_rhs_evals;  /* RHS evaluations of the dense output (4 per such
                            step for DOP853, 0 for DOPRI5; not included in
                            rhs_evals) */
} odehip_solve_stats;

/* x: device, n_var; in: x0, out: the state at the final time. */
int odehip_rk_solve(odehip_context* ctx, int tableau, const odehip_ode* ode,
                    const odehip_solve_options* opt, double* x,
                    odehip_solve_stats* stats);

/* ---------------------------------------------------------------- leapfrog
 * ParallelLeapfrogIntegrator::integrate (src/integrator/leapfrog/
 * LeapfrogParallel.hpp:8-41): fixed-step velocity Verlet of an N-body state
 * with the accelerations of the RHS variant ode->nbody_variant (NBodyODE::acc,
 * NBodyODE.hpp:41-46); t advances by t += dt while t < tf (no clip of the last
 * step, as the reference).  x: device, in/out.  states (host, states_cap x
 * n_var) and times (host, states_cap) receive the first states_cap saved
 * states (RawOutput::save_state at t0 and after every step); either may be
 * NULL. */
typedef struct {
  long steps;            /* steps taken */
  long rhs_evals;        /* steps + 1 */
  long n_states;         /* states written */
  double t;              /* final time */
} odehip_leapfrog_stats;

int odehip_leapfrog(odehip_context* ctx, const odehip_ode* ode, double t0,
                    double tf, double dt, double* x, double* states,
                    double* times, long states_cap,
                    odehip_leapfrog_stats* stats);

/* ------------------------------------------------------ trajectory output
 * output_to_file (src/demo/nbody_io.hpp:4-22; read by
 * scripts/nbody_io/bin_reader.py:4-18): little-endian u64 n_times, u64 n_var,
 * f64 softening, then n_times rows of f64 t followed by n_var f64 state
 * values.  times, states: host.  Returns ODEHIP_ERR_INVALID if the file cannot
 * be written. */
int odehip_write_trajectory(const char* path, size_t n_times, size_t n_var,
                            double softening, const double* times,
                            const double* states);

/* ----------------------------------------------------- conserved quantities
 * scripts/nbody_physics/energy.py:118-192,224-241 (softened PE + KE), with
 * masses; and total momentum (absent from the reference).  host_out: ke, pe /
 * px, py, pz. */
int odehip_nbody_energy(odehip_context* ctx, const double* x, size_t n,
                        double softening, const double* masses,
                        double* host_out);
int odehip_nbody_momentum(odehip_context* ctx, const double* x, size_t n,
                          const double* masses, double* host_out);

/* ------------------------------------------------- learned step controller
 * A small network proposes the next step ratio (no reference counterpart;
 * the reference's controller is RKEmbeddedParallel.hpp:64-79, whose test
 * still decides acceptance: a step is accepted exactly when its error norm
 * is <= 1; the
 * proposed ratio is capped at 1 right after a rejection, the reference's
 * clamps and clip to tf apply, and the PI controller replaces the network
 * before three attempts of history exist, after two rejections in a row, or
 * on a non-finite output).
 * Network: 10 -> 64 -> 64 -> 1, tanh, linear output = log(dt_new / dt),
 * clamped to [log 0.2, log 5].  Inputs (raw features, normalised by the
 * stored mean / std): log err of the last three attempts, log(dt / (tf -
 * t0)) of the last, log of the two earlier dt over the last, the three
 * accept flags (1/0), log N.  Parameters: FP32 flat array
 * W1[64][10] | b1[64] | W2[64][64] | b2[64] | w3[64] | b3.  Training: FP32
 * forward/backward kernels + Adam; inference: bfloat16 matrix cores. */
#define ODEHIP_MLP_IN 10
#define ODEHIP_MLP_HIDDEN 64
#define ODEHIP_MLP_PARAMS 4929

typedef struct odehip_mlp odehip_mlp;

typedef struct {
  int epochs;
  int batch;             /* samples per Adam step */
  float lr;
  float beta1, beta2, eps; /* Adam; 0 = 0.9, 0.999, 1e-8 */
  uint64_t seed;         /* sample order */
} odehip_mlp_train_options;

/* params: host, ODEHIP_MLP_PARAMS floats, or NULL (Glorot-uniform weights
 * from seed, zero biases); mean, stdv: host, ODEHIP_MLP_IN floats. */
int odehip_mlp_create(odehip_context* ctx, const float* params, uint64_t seed,
                      const float* mean, const float* stdv, odehip_mlp** out);
int odehip_mlp_destroy(odehip_mlp* mlp);
/* params_out: host, ODEHIP_MLP_PARAMS floats (FP32 master weights). */
int odehip_mlp_get_params(odehip_context* ctx, const odehip_mlp* mlp,
                          float* params_out);
/* Mean-squared loss 0.5 mean((y - t)^2) of n samples and its gradient
 * (host, ODEHIP_MLP_PARAMS floats) at the current weights.  x: device
 * [n][ODEHIP_MLP_IN] raw features, t: device [n] targets. */
int odehip_mlp_loss_grad(odehip_context* ctx, odehip_mlp* mlp, const float* x,
                         const float* t, long n, float* grad_out,
                         double* loss_out);
/* Adam over the n samples for opt->epochs epochs; epoch_loss (host, may be
 * NULL) receives each epoch's mean loss.  Publishes the bfloat16 inference
 * copy at the end. */
int odehip_mlp_train(odehip_context* ctx, odehip_mlp* mlp, const float* x,
                     const float* t, long n, const odehip_mlp_train_options* opt,
                     double* epoch_loss);
/* Network outputs (device, n floats) for raw features x (device), on the
 * matrix cores with the bfloat16 weights. */
int odehip_mlp_infer(odehip_context* ctx, const odehip_mlp* mlp,
                     const float* x, long n, float* y);

/* ------------------------------------------------------------- ensembles
 * Many independent N-body systems of different sizes solved together, each
 * with its own step size, accept/reject sequence, counters and termination
 * (the reference integrates one system per integrate() call,
 * RKEmbeddedParallel.hpp:13-81; this is that call for every system at once).
 * States are packed: system k occupies 6 n_k doubles [pos | vel] starting at
 * 6 * (n_0 + ... + n_{k-1}); masses (optional) are packed the same way with
 * n_k entries per system.  Each system's result is bit-identical to
 * odehip_rk_solve on that system alone with the same options (same right-hand
 * side variant, FSAL setting and controller).
 *
 * The attempt of every active system advances in lockstep rounds: one
 * launch per stage covers all systems, a round ends with a controller pass
 * (one thread per system), and K rounds are captured in a HIP graph that is
 * replayed until every system has finished.  The right-hand side runs one
 * wavefront per tile of 64 (or 128) target particles; which systems use
 * 128-target tiles, the waves per workgroup and the XCD placement of tiles
 * are chosen by an autotuner that times the candidates on the ensemble and,
 * among those within 3 % of the fastest, keeps the one with the most target
 * tiles.  When fewer than 4096 tiles of running systems remain, 128-target
 * tiles are split in two for the rest of the solve (results unchanged).
 *
 * The fast and mixed variants sum each target's partners in chunks of 128
 * (increasing j within a chunk, chunk sums added in chunk order).
 * Configurations ODEHIP_ENS_TILE_UNSPLIT .. ODEHIP_ENS_TILE_CONFIGS-1 share
 * one 64-target tile among 2, 4 or 8 waves of a workgroup, with the same
 * bits: fast and mixed fold the waves' chunk sums in chunk order; exact
 * computes the terms of blocks of partners on all waves and adds them in
 * increasing j on three.  The autotuner times all configurations; a solve
 * also moves to a shared-tile one, at the start and as systems finish, when
 * its running systems are few or their longest partner walks dominate (a
 * decision table measured on this device). */
#define ODEHIP_ENS_TILE_CONFIGS 18
#define ODEHIP_ENS_TILE_UNSPLIT 12
#define ODEHIP_ENS_TILE_AUTO (-1)

typedef struct {
  double t0, tf;
  double atol, rtol;     /* scalar tolerances, every system and component */
  long max_accepted;     /* per system; 0 = none */
  long max_attempts;     /* per system; 0 = none */
  int fsal;              /* as odehip_solve_options.fsal */
  int rounds_per_launch; /* lockstep rounds per graph replay; 0 = auto */
  int tile_config;       /* ODEHIP_ENS_TILE_AUTO or 0..ODEHIP_ENS_TILE_CONFIGS-1 */
  int controller;        /* ODEHIP_CONTROLLER_* */
  const struct odehip_mlp* mlp; /* trained network (LEARNED only) */
  double* trace;         /* host, n_systems x trace_cap x 3: per attempt
                            (err, dt, accepted); may be NULL */
  long trace_cap;
  double* tune_ms;       /* host, ODEHIP_ENS_TILE_CONFIGS entries: ms per
                            right-hand side of every candidate (autotune
                            only; untried entries 0); may be NULL */
  /* Dense output (DOPRI5 or DOP853) of every system at common times, as
   * odehip_solve_options.dense_times: dense_times host, n_dense strictly
   * increasing times after t0; dense_states host, n_dense x (total state
   * length), each row packed like x0.  A system's row r is written when one
   * of its accepted steps covers dense_times[r]; rows it does not reach
   * (odehip_system_stats.n_dense) are NaN.  DOPRI5 adds no RHS evaluation;
   * DOP853 adds, for the systems with a covering step, the single-system
   * pass's evaluations. */
  const double* dense_times;
  long n_dense;
  double* dense_states;
} odehip_ensemble_options;

typedef struct {
  long accepted, rejected, attempts, rhs_evals, n_trace, learned_steps;
  int status;            /* ODEHIP_STATUS_* */
  int pad;
  double t, dt, dt_initial, last_err;
  long n_dense;          /* dense-output rows written for this system */
} odehip_system_stats;

typedef struct {
  long rounds;           /* lockstep rounds launched */
  long accepted, rejected, attempts, rhs_evals, learned_steps; /* sums */
  double particle_steps; /* sum over systems of n_k * accepted_k */
  double interactions;   /* sum n_k (n_k - 1): pair terms per ensemble RHS */
  int tile_config;       /* the tuned (or requested) configuration; the running
                            systems may move to a shared-tile one (same results) */
  int n_unfinished;      /* systems still running at return (0 on success) */
  double rhs_ms;         /* autotuner's time of one RHS in that configuration
                            (0 when not tuned) */
} odehip_ensemble_stats;

/* n_particles: host, n_systems entries (>= 1).  masses, x0, x_out: device,
 * packed (masses may be NULL: unit masses).  per_system: host array of
 * n_systems or NULL.  Only ODEHIP_RHS_NBODY systems. */
int odehip_ensemble_solve(odehip_context* ctx, int tableau, int n_systems,
                          const int* n_particles, int nbody_variant,
                          double softening, const double* masses,
                          const odehip_ensemble_options* opt,
                          const double* x0, double* x_out,
                          odehip_ensemble_stats* stats,
                          odehip_system_stats* per_system);

/* One batched evaluation of every system's right-hand side (NBodyODE
 * operator(), NBodyODE.hpp:31-39, per system): dxdt = f(x), both packed and
 * on the device.  tile_config as above (AUTO: the tuned or default one);
 * ms_out (host, may be NULL) receives the kernel time. */
int odehip_ensemble_rhs(odehip_context* ctx, int n_systems,
                        const int* n_particles, int nbody_variant,
                        double softening, const double* masses,
                        int tile_config, const double* x, double* dxdt,
                        double* ms_out);

/* ------------------------------------------------------ initial conditions
 * Host-side generators of the synthetic systems (seeded mt19937_64 with an
 * explicit 53-bit mapping and Box-Muller).  x_out: 6 n doubles
 * [pos | vel], masses_out: n doubles (may be NULL). */
/* Positions uniform in [-half_width, half_width]^3, velocities N(0, sigma^2)
 * per component, masses total_mass / n, momentum removed. */
int odehip_ic_uniform_cube(size_t n, uint64_t seed, double half_width,
                           double vel_sigma, double total_mass, double* x_out,
                           double* masses_out);
/* Plummer sphere (Aarseth, Henon & Wielen 1974), scale radius a, total mass
 * M, G = 1, isotropic, virial equilibrium; centre of mass and momentum
 * removed. */
int odehip_ic_plummer(size_t n, uint64_t seed, double scale_radius,
                      double total_mass, double* x_out, double* masses_out);
/* Two Plummer spheres of n/2 bodies each, centres at (+-separation/2, 0, 0)
 * moving with -+ rel_speed/2 along x; total mass M. */
int odehip_ic_colliding_plummer(size_t n, uint64_t seed, double scale_radius,
                                double total_mass, double separation,
                                double rel_speed, double* x_out,
                                double* masses_out);

#ifdef __cplusplus
}
#endif

#endif /* ODEHIP_H */
