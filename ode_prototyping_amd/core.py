"""Execution policy, state, ODE, output and integrator mirrors over libodehip.

Reference interfaces mirrored here (file:line in noether17/ode_prototyping):

* ``HipExecutor``  -- ``src/par_exec/CudaExecutor.cuh:80-126`` (an executor
  object the integrator and ODEs run their kernels on).  Owns an odehip
  context (HIP stream + scratch).
* ``HipState``     -- ``src/state/CudaState.cuh:10-64`` (owning device buffer:
  zero-initialised default, construction from host data, deep copies,
  ``copy_to_span``).
* ``NBodyODE`` / ``ExponentialODE`` / ``VanDerPolODE`` -- ``src/ode/*.hpp``;
  calling one as ``ode(exe, x, dxdt)`` evaluates the right-hand side on the
  device exactly like ``operator()(exe, span x, span dxdt)``.
* ``RawOutput`` / ``MostRecentOutput`` -- ``src/output/RawOutput.hpp:8-27`` and
  ``src/benchmark/n_body_benchmark.cu:23-34``.
* ``integrate``    -- ``RKEmbeddedParallel<BT>::integrate``
  (``RKEmbeddedParallel.hpp:13-81``) with the attempt loop and controller on
  the device; ``integrate_host_driven`` runs the reference's host loop
  literally (one executor call per reference kernel launch, host controller
  with the C library's pow), the way the reference's CUDA executor drives it.
"""
import ctypes as C
import math

import numpy as np

from . import _lib
from ._lib import (CONTROLLERS, NBODY_VARIANTS, RHS_EXPONENTIAL, RHS_NBODY, RHS_VANDERPOL, STAGES,
                   STATUS, TABLEAUS, dptr, lib)


class OdeHipError(RuntimeError):
    pass


class ButcherTableau:
    """A Butcher tableau in the reference's struct format
    (``butcher_tableaus/BTDOPRI5.hpp:5-26``: ``a``, ``b``, ``bt``, ``p``, ``pt``,
    ``n_stages``), usable wherever a built-in tableau name is.

    ``a`` is (n_stages - 1) x (n_stages - 1), row s - 1 the weights of stage s
    (entries j >= s ignored, zero in the reference's arrays); ``b`` (order ``p``)
    and ``bt`` (order ``pt``) have n_stages entries.  The library registers it
    (``odehip_tableau_register``) on first use; the same coefficients as a
    built-in method give that method's id.
    """

    def __init__(self, a, b, bt, p, pt, name=None):
        self.b = np.ascontiguousarray(b, dtype=np.float64).ravel()
        self.bt = np.ascontiguousarray(bt, dtype=np.float64).ravel()
        S = self.b.size
        if not 2 <= S <= _lib.MAX_STAGES or self.bt.size != S:
            raise ValueError("b and bt need the same length, 2 .. 13 stages")
        rows = [list(r) for r in a]
        if len(rows) != S - 1:
            raise ValueError(f"a needs {S - 1} rows for {S} stages")
        self.a = np.zeros((S - 1, S - 1))
        for s_, r in enumerate(rows):
            if len(r) < s_ + 1:
                raise ValueError(f"row {s_} of a needs {s_ + 1} weights")
            self.a[s_, :s_ + 1] = r[:s_ + 1]
        self.p, self.pt = int(p), int(pt)
        self.name = name or "registered"
        self._id = None

    @property
    def n_stages(self):
        return self.b.size

    def id(self, exe):
        """The library's tableau id (registered on the first call)."""
        if self._id is None:
            out = C.c_int(-1)
            exe.check(lib().odehip_tableau_register(exe.handle, self.n_stages, self.p, self.pt,
                                                    dptr(self.a), dptr(self.b), dptr(self.bt),
                                                    C.byref(out)), "tableau_register")
            self._id = out.value
        return self._id

    def __repr__(self):
        return f"ButcherTableau({self.name}, stages={self.n_stages}, p={self.p}, pt={self.pt})"


def tableau_id(exe, tableau):
    """Library id of a built-in tableau name or a :class:`ButcherTableau`."""
    if isinstance(tableau, ButcherTableau):
        return tableau.id(exe)
    if tableau not in TABLEAUS:
        raise ValueError(f"unknown tableau {tableau}")
    return TABLEAUS[tableau]


def tableau_shape(tableau):
    """(n_stages, p, pt) of a built-in tableau name or a :class:`ButcherTableau`."""
    if isinstance(tableau, ButcherTableau):
        return tableau.n_stages, tableau.p, tableau.pt
    if tableau not in TABLEAUS:
        raise ValueError(f"unknown tableau {tableau}")
    return (STAGES[tableau],) + _lib.ORDERS[tableau]


class HipExecutor:
    """HIP execution policy: one odehip context on one device."""

    def __init__(self, device=0):
        L = lib()
        h = C.c_void_p()
        rc = L.odehip_create(device, C.byref(h))
        if rc != 0:
            raise OdeHipError(f"odehip_create(device={device}) failed with {rc}: no usable "
                              "HIP device (the library has no CPU fallback)")
        self._h = h
        self.device = device

    @property
    def handle(self):
        return self._h

    def check(self, rc, what):
        if rc != 0:
            msg = lib().odehip_last_error(self._h)
            raise OdeHipError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def synchronize(self):
        self.check(lib().odehip_synchronize(self._h), "synchronize")

    def timer_start(self):
        """Record a HIP start event on the executor's stream."""
        self.check(lib().odehip_timer_start(self._h), "timer_start")

    def timer_stop(self):
        """Record a stop event, wait for it; elapsed milliseconds since timer_start."""
        ms = C.c_double()
        self.check(lib().odehip_timer_stop(self._h, C.byref(ms)), "timer_stop")
        return ms.value

    @property
    def near_field(self):
        """Near-field coefficient c of the mixed RHS: pairs with r^2 < c / n in FP64.

        0 (MIXED_NEAR_DEFAULT) keeps every pair in FP32; MIXED_NEAR_CLOSE (20) is
        the setting for systems with close encounters.  Single-system calls only:
        the ensemble solver's mixed RHS keeps every pair in FP32."""
        c = C.c_double()
        self.check(lib().odehip_get_near_field(self._h, C.byref(c)), "get_near_field")
        return c.value

    @near_field.setter
    def near_field(self, coefficient):
        self.check(lib().odehip_set_near_field(self._h, float(coefficient)), "set_near_field")

    def probe_fp64(self):
        """Sustained FP64 FMA TFLOP/s and shader clock (MHz) of the device."""
        tf, mhz = C.c_double(), C.c_double()
        self.check(lib().odehip_probe_fp64(self._h, C.byref(tf), C.byref(mhz)), "probe_fp64")
        return tf.value, mhz.value

    def close(self):
        if getattr(self, "_h", None):
            lib().odehip_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- reference kernels (one call per reference kernel launch) ------------
    def rk_stage(self, tableau, stage, dt, temp, ks, x0):
        self.check(lib().odehip_rk_stage(self._h, tableau_id(self, tableau), stage, dt, temp.ptr,
                                         ks.ptr, x0.ptr, x0.size), "rk_stage")

    def rk_update_state_and_error(self, tableau, dt, x, err, ks, x0):
        self.check(lib().odehip_rk_update_state_and_error(
            self._h, tableau_id(self, tableau), dt, x.ptr, err.ptr, ks.ptr, x0.ptr, x0.size),
            "rk_update_state_and_error")

    def rk_update_error_target(self, sc, atol, rtol, x, x0):
        self.check(lib().odehip_rk_update_error_target(self._h, sc.ptr, atol.ptr, rtol.ptr,
                                                       x.ptr, x0.ptr, x0.size),
                   "rk_update_error_target")

    def rk_norm(self, v, sc):
        out = C.c_double()
        self.check(lib().odehip_rk_norm(self._h, v.ptr, sc.ptr, v.size, C.byref(out)),
                   "rk_norm")
        return out.value

    def fill(self, state, value):
        self.check(lib().odehip_fill(self._h, state.ptr, float(value), state.size), "fill")


class HipState:
    """Owning device buffer of ``size`` doubles (CudaState counterpart)."""

    def __init__(self, exe, size=None, data=None):
        if data is not None:
            data = np.ascontiguousarray(data, dtype=np.float64).ravel()
            size = data.size
        if size is None:
            raise ValueError("size or data required")
        self.exe = exe
        self.size = int(size)
        p = C.c_void_p()
        exe.check(lib().odehip_alloc(exe.handle, self.size, C.byref(p)), "alloc")
        self._p = p
        if data is not None:
            exe.check(lib().odehip_copy_h2d(exe.handle, self._p, dptr(data), self.size), "h2d")

    @property
    def ptr(self):
        return self._p

    def copy(self):
        other = HipState(self.exe, self.size)
        self.exe.check(lib().odehip_copy_d2d(self.exe.handle, other.ptr, self._p, self.size),
                       "d2d")
        return other

    def assign(self, other):
        if other.size != self.size:
            raise ValueError("size mismatch")
        self.exe.check(lib().odehip_copy_d2d(self.exe.handle, self._p, other.ptr, self.size),
                       "d2d")

    def upload(self, data):
        data = np.ascontiguousarray(data, dtype=np.float64).ravel()
        if data.size != self.size:
            raise ValueError("size mismatch")
        self.exe.check(lib().odehip_copy_h2d(self.exe.handle, self._p, dptr(data), self.size),
                       "h2d")

    def copy_to_span(self, out=None):
        if out is None:
            out = np.empty(self.size)
        self.exe.check(lib().odehip_copy_d2h(self.exe.handle, dptr(out), self._p, self.size),
                       "d2h")
        return out

    def to_numpy(self):
        return self.copy_to_span()

    def free(self):
        if getattr(self, "_p", None):
            lib().odehip_free(self.exe.handle, self._p)
            self._p = None

    def __del__(self):
        try:
            if self.exe.handle:
                self.free()
        except Exception:
            pass


class NBodyODE:
    """N-body gravity RHS (src/ode/NBodyODE.hpp), G = 1.

    ``masses``: None for unit masses (the reference), else a length-N array
    (held on the device).  ``variant``: "exact" (bit-identical to the
    reference's single-threaded evaluation), "fast" (FP64, refined rsqrt) or
    "mixed" (FP32 pair force, FP64 accumulation; near pairs in FP64 when the
    executor's ``near_field`` is set).
    """
    rhs_kind = RHS_NBODY

    def __init__(self, softening=0.0, masses=None, variant="exact", exe=None):
        self.softening = float(softening)
        self.variant = variant
        self._variant_id = NBODY_VARIANTS[variant]
        self.masses_host = None if masses is None else np.ascontiguousarray(masses, np.float64)
        self._masses_dev = None
        self._exe = exe

    def masses_device(self, exe):
        if self.masses_host is None:
            return None
        if self._masses_dev is None or self._masses_dev.exe is not exe:
            self._masses_dev = HipState(exe, data=self.masses_host)
        return self._masses_dev

    def __call__(self, exe, x, dxdt):
        m = self.masses_device(exe)
        exe.check(lib().odehip_nbody_rhs(exe.handle, x.ptr, dxdt.ptr, x.size // 6,
                                         self.softening, m.ptr if m else None,
                                         self._variant_id), "nbody_rhs")

    def acc(self, exe, pos_vel, acc):
        """NBodyODE::acc (NBodyODE.hpp:41-46): acc (3N) from the positions."""
        m = self.masses_device(exe)
        exe.check(lib().odehip_nbody_acc(exe.handle, pos_vel.ptr, acc.ptr, pos_vel.size // 6,
                                         self.softening, m.ptr if m else None,
                                         self._variant_id), "nbody_acc")

    def ode_struct(self, exe, n_var):
        m = self.masses_device(exe)
        return _lib.OdeStruct(RHS_NBODY, self._variant_id, n_var, self.softening,
                              m.ptr.value if m else None)


class _SimpleODE:
    rhs_kind = None

    def __call__(self, exe, x, dxdt):
        exe.check(lib().odehip_ode_rhs(exe.handle, self.rhs_kind, x.ptr, dxdt.ptr, x.size),
                  "ode_rhs")

    def ode_struct(self, exe, n_var):
        return _lib.OdeStruct(self.rhs_kind, 0, n_var, 0.0, None)


class ExponentialODE(_SimpleODE):
    """dx/dt = x (src/ode/ExponentialODE.hpp)."""
    rhs_kind = RHS_EXPONENTIAL


class VanDerPolODE(_SimpleODE):
    """Van der Pol oscillator with eps = 1 (src/ode/VanDerPolODE.hpp)."""
    rhs_kind = RHS_VANDERPOL


class RawOutput:
    """Keeps every saved (t, state) on the host (src/output/RawOutput.hpp)."""

    def __init__(self):
        self._times = []
        self._states = []

    def save_state(self, t, x):
        self._times.append(float(t))
        self._states.append(x.copy_to_span() if isinstance(x, HipState) else np.array(x))

    def times(self):
        return self._times

    def states(self):
        return self._states


class MostRecentOutput:
    """Keeps only the last saved state (n_body_benchmark.cu:23-34)."""

    def __init__(self):
        self.time = None
        self.state = None

    def save_state(self, t, x):
        self.time = float(t)
        self.state = x.copy_to_span() if isinstance(x, HipState) else np.array(x)


class DenseOutput:
    """DOP853 / DOPRI5 dense output at equally spaced points (the reference's
    DenseOutput policy, NumericalRecipes/Output.hpp:63-147): the initial
    state, then the state at t0 + k (tf - t0) / n_intervals for every such
    point strictly before the end of the integration, from the continuous
    extension of the step that covers it (steps are not shortened).  The
    points are accumulated by repeated addition, as the reference does, so
    they are bit-identical to its x values."""

    default_n_intervals = 20

    def __init__(self, n_intervals=default_n_intervals):
        self.set_n_intervals(n_intervals)
        self._x = []
        self._y = []

    def set_n_intervals(self, n_intervals):
        if n_intervals < 1:
            raise ValueError("n_intervals must be at least 1.")
        self.n_intervals = int(n_intervals)

    def output_times(self, t0, tf):
        """The candidate points after t0 (those before tf)."""
        width = (tf - t0) / self.n_intervals
        out = []
        t = t0 + width
        while t < tf and len(out) <= self.n_intervals:
            out.append(t)
            t += width
        return np.array(out, dtype=np.float64)

    def save_state(self, t, x):
        self._x.append(float(t))
        self._y.append(x.copy_to_span() if isinstance(x, HipState) else np.array(x))

    def n_steps(self):
        return len(self._x)

    def x_values(self):
        return self._x

    def y_values(self):
        """Per component, the saved values (y_values()[i][k])."""
        return np.array(self._y).T if self._y else np.zeros((0, 0))


class SolveResult:
    """Outcome of a device-resident solve."""

    def __init__(self, stats, x_final, times, trace):
        self.accepted = stats.accepted
        self.rejected = stats.rejected
        self.attempts = stats.attempts
        self.rhs_evals = stats.rhs_evals
        self.status = STATUS.get(stats.status, str(stats.status))
        self.t = stats.t
        self.dt = stats.dt
        self.dt_initial = stats.dt_initial
        self.last_err = stats.last_err
        self.x_final = x_final
        self.times = times
        self.trace = trace
        self.n_dense = stats.n_dense
        self.dense_steps = stats.dense_steps
        self.dense_rhs_evals = stats.dense_rhs_evals
        self.dense_times = None
        self.dense_states = None

    def __repr__(self):
        return (f"SolveResult(status={self.status}, accepted={self.accepted}, "
                f"rejected={self.rejected}, rhs_evals={self.rhs_evals}, t={self.t})")


def _as_state(exe, v, n, name):
    if isinstance(v, HipState):
        if v.size != n:
            raise ValueError(f"{name} has {v.size} entries, expected {n}")
        return v
    arr = np.broadcast_to(np.asarray(v, dtype=np.float64), (n,))
    return HipState(exe, data=np.ascontiguousarray(arr))


def integrate(tableau, x0, t0, tf, atol, rtol, ode, output=None, exe=None, *,
              max_accepted=0, max_attempts=0, fsal=False, attempts_per_launch=0,
              times_cap=None, trace_cap=0, x_out=None, fetch=True, controller="reference",
              mlp=None, states_cap=None, dense_times=None):
    """RKEmbeddedParallel<BT>::integrate with the attempt loop on the device.

    ``x0``: HipState or host array (n_var); ``atol``/``rtol``: scalars, host
    arrays or HipStates (per component, as the reference).  ``output`` (if
    given) receives ``save_state(t, x)`` for the initial state and every
    accepted step, as the reference's RawOutput (the states are kept on the
    device during the solve: ``states_cap`` of them, default ``times_cap``;
    a list larger than 8 GiB must be asked for explicitly).
    Returns a :class:`SolveResult` (final state on the host in ``x_final``
    unless ``fetch`` is False -- it stays in ``x_out`` on the device --,
    accepted times in ``times``, per-attempt ``trace`` rows (err, dt, accepted)
    if ``trace_cap``).

    Dense output (DOP853, DOPRI5): ``dense_times`` (increasing, after t0) gives
    ``res.dense_states`` rows at ``res.dense_times`` -- those reached before
    the end --, from each covering step's continuous extension
    (StepperDopr853 / StepperDopr5 prepare_dense, dense_out).  An ``output``
    that is a
    :class:`DenseOutput` receives the initial state and its equally spaced
    points instead of the accepted steps.
    """
    tableau_shape(tableau)  # a known tableau (ValueError otherwise)
    if controller not in CONTROLLERS:
        raise ValueError(f"unknown controller {controller}")
    exe = exe or HipExecutor()
    x0_host = None if isinstance(x0, HipState) else np.ascontiguousarray(x0, np.float64)
    n = x0.size if isinstance(x0, HipState) else x0_host.size
    work = x_out if x_out is not None else HipState(exe, n)
    if isinstance(x0, HipState):
        work.assign(x0)
    else:
        work.upload(x0_host)
    at = _as_state(exe, atol, n, "atol")
    rt = _as_state(exe, rtol, n, "rtol")
    if times_cap is None:
        times_cap = (max_accepted + 1) if max_accepted > 0 else 1 << 16
    times = np.zeros(max(int(times_cap), 1))
    trace = np.zeros((max(int(trace_cap), 1), 3))
    dense_policy = isinstance(output, DenseOutput)
    if dense_policy:
        if dense_times is not None:
            raise ValueError("pass either dense_times or a DenseOutput, not both")
        dense_times = output.output_times(float(t0), float(tf))
        x0_saved = x0_host.copy() if x0_host is not None else work.copy_to_span()
    if states_cap is None:
        states_cap = times_cap if (output is not None and not dense_policy) else 0
        if states_cap * n * 8 > (8 << 30):
            raise ValueError("the saved-state list would exceed 8 GiB: pass states_cap")
    states = np.zeros((max(int(states_cap), 1), n)) if states_cap > 0 else None
    dt_arr = None
    dense_buf = None
    if dense_times is not None and len(dense_times) > 0:
        dt_arr = np.ascontiguousarray(dense_times, dtype=np.float64)
        if dt_arr.size * n * 8 > (8 << 30):
            raise ValueError("the dense output would exceed 8 GiB")
        dense_buf = np.zeros((dt_arr.size, n))
    ode_s = ode.ode_struct(exe, n)
    opt = _lib.SolveOptions(float(t0), float(tf), at.ptr.value, rt.ptr.value,
                            int(max_accepted), int(max_attempts), 1 if fsal else 0,
                            int(attempts_per_launch),
                            times.ctypes.data if times_cap > 0 else None, int(times_cap),
                            trace.ctypes.data if trace_cap > 0 else None, int(trace_cap),
                            CONTROLLERS[controller], mlp.handle if mlp is not None else None,
                            states.ctypes.data if states is not None else None,
                            int(states_cap),
                            dt_arr.ctypes.data if dt_arr is not None else None,
                            int(dt_arr.size) if dt_arr is not None else 0,
                            dense_buf.ctypes.data if dense_buf is not None else None)
    stats = _lib.SolveStats()
    exe.check(lib().odehip_rk_solve(exe.handle, tableau_id(exe, tableau), C.byref(ode_s),
                                    C.byref(opt), work.ptr, C.byref(stats)), "rk_solve")
    x_final = work.copy_to_span() if fetch else None
    res = SolveResult(stats, x_final, times[:stats.n_times].copy(),
                      trace[:stats.n_trace].copy() if trace_cap else None)
    if dt_arr is not None:
        res.dense_times = dt_arr[:stats.n_dense].copy()
        res.dense_states = dense_buf[:stats.n_dense]
    if dense_policy:
        output.save_state(float(t0), x0_saved)
        for k in range(stats.n_dense):
            output.save_state(dt_arr[k], dense_buf[k])
    elif output is not None:
        for k in range(min(stats.n_states, stats.n_times)):
            output.save_state(times[k], states[k])
    return res


def _host_pow(x, y):
    return math.pow(x, y) if x > 0 else (0.0 if x == 0 else float("nan"))


def integrate_host_driven(tableau, x0, t0, tf, atol, rtol, ode, output, exe, *,
                          max_accepted=0, max_attempts=0):
    """RKEmbeddedParallel<BT>::integrate driven from the host.

    Literal restatement of the reference loop (RKEmbeddedParallel.hpp:13-81,
    initial step :170-200): every reference kernel launch becomes one executor
    call on the device, the error norm is read back each attempt and the
    controller runs on the host with the C library's pow, as the reference's
    CUDA executor does.  ``output.save_state`` is called for every accepted
    state.  Returns (accepted, rejected, rhs_evals).
    """
    S, p, pt = tableau_shape(tableau)
    x0s = x0.copy() if isinstance(x0, HipState) else HipState(exe, data=x0)
    n = x0s.size
    at = _as_state(exe, atol, n, "atol")
    rt = _as_state(exe, rtol, n, "rtol")
    max_step_scale, min_step_scale = 6.0, 0.33
    q = min(p, pt)
    safety_factor = math.pow(0.38, 1.0 / (1.0 + q))
    ks = HipState(exe, S * n)
    rhs_evals = 0

    # estimate_initial_step (:170-200), its kernels as executor-level calls
    L = lib()
    sc, f0, x1, df = (HipState(exe, n) for _ in range(4))
    exe.check(L.odehip_rk_compute_error_target(exe.handle, sc.ptr, x0s.ptr, at.ptr, rt.ptr, n),
              "compute_error_target")
    ode(exe, x0s, f0)
    d0 = exe.rk_norm(x0s, sc)
    d1 = exe.rk_norm(f0, sc)
    dt0 = 1.0e-6 if (d0 < 1.0e-5 or d1 < 1.0e-5) else 0.01 * (d0 / d1)
    exe.check(L.odehip_rk_euler_step(exe.handle, x1.ptr, x0s.ptr, f0.ptr, dt0, n), "euler_step")
    ode(exe, x1, df)
    exe.check(L.odehip_rk_difference(exe.handle, df.ptr, f0.ptr, n), "difference")
    d2 = exe.rk_norm(df, sc) / dt0
    dmax = d2 if d1 < d2 else d1
    dt1 = (max(1.0e-6, dt0 * 1.0e-3) if dmax <= 1.0e-15
           else _host_pow(0.01 / dmax, 1.0 / (1.0 + p)))
    dt = min(100.0 * dt0, dt1)
    rhs_evals += 2
    t = t0
    x = x0s.copy()
    temp = HipState(exe, n)
    err = HipState(exe, n)
    tgt = HipState(exe, n)
    output.save_state(t, x)
    accepted = rejected = attempts = 0
    views = [_Sub(ks, j * n, n) for j in range(S)]
    while t < tf:
        if max_attempts and attempts >= max_attempts:
            break
        attempts += 1
        ode(exe, x0s, views[0])
        for stage in range(1, S):
            exe.rk_stage(tableau, stage, dt, temp, ks, x0s)
            ode(exe, temp, views[stage])
        rhs_evals += S
        exe.rk_update_state_and_error(tableau, dt, x, err, ks, x0s)
        exe.rk_update_error_target(tgt, at, rt, x, x0s)
        scaled_error = exe.rk_norm(err, tgt)
        if not math.isfinite(scaled_error):
            rejected += 1
            break
        accept = scaled_error <= 1.0
        if accept:
            t += dt
            x0s.assign(x)
            output.save_state(t, x)
            accepted += 1
        else:
            rejected += 1
        dtnew = dt * safety_factor / _host_pow(scaled_error, 1.0 / (1.0 + q))
        if abs(dtnew) > max_step_scale * abs(dt):
            dt *= max_step_scale
        elif abs(dtnew) < min_step_scale * abs(dt):
            dt *= min_step_scale
        elif dtnew / (tf - t0) < 1.0e-12:
            dt = (tf - t0) * 1.0e-12
        else:
            dt = dtnew
        if t + dt > tf:
            dt = tf - t
        if accept and max_accepted and accepted >= max_accepted:
            break
    return accepted, rejected, rhs_evals


class _Sub:
    """Non-owning view of a HipState range (a std::span into ks)."""

    def __init__(self, base, offset, size):
        self.base = base
        self.size = size
        self.ptr = C.c_void_p(base.ptr.value + 8 * offset)


def leapfrog(x0, t0, tf, dt, ode, output=None, exe=None, *, states_cap=None, x_out=None,
             fetch=True):
    """ParallelLeapfrogIntegrator::integrate (``LeapfrogParallel.hpp:8-41``):
    fixed-step velocity Verlet of an N-body state on the device, the
    accelerations from ``ode`` (an :class:`NBodyODE`).  ``output`` (if given)
    receives ``save_state(t, x)`` for the initial state and after every step,
    as in the reference; ``states_cap`` limits how many are kept (default:
    all when ``output`` is given, none otherwise).  Returns
    ``(x_final, stats)`` with stats a dict (steps, rhs_evals, t).
    """
    if not isinstance(ode, NBodyODE):
        raise TypeError("leapfrog needs an NBodyODE")
    exe = exe or HipExecutor()
    x0_host = None if isinstance(x0, HipState) else np.ascontiguousarray(x0, np.float64)
    n = x0.size if isinstance(x0, HipState) else x0_host.size
    work = x_out if x_out is not None else HipState(exe, n)
    if isinstance(x0, HipState):
        work.assign(x0)
    else:
        work.upload(x0_host)
    if states_cap is None:
        if output is not None:
            t, steps = float(t0), 0
            while t < tf:  # the reference's time loop
                t += dt
                steps += 1
            states_cap = steps + 1
        else:
            states_cap = 0
    states = np.zeros((max(int(states_cap), 1), n))
    times = np.zeros(max(int(states_cap), 1))
    st = _lib.LeapfrogStats()
    ode_s = ode.ode_struct(exe, n)
    exe.check(lib().odehip_leapfrog(exe.handle, C.byref(ode_s), float(t0), float(tf),
                                    float(dt), work.ptr,
                                    dptr(states) if states_cap > 0 else None,
                                    dptr(times) if states_cap > 0 else None,
                                    int(states_cap), C.byref(st)), "leapfrog")
    if output is not None:
        for k in range(st.n_states):
            output.save_state(times[k], states[k])
    x_final = work.copy_to_span() if fetch else None
    return x_final, {"steps": st.steps, "rhs_evals": st.rhs_evals, "t": st.t,
                     "n_states": st.n_states}
