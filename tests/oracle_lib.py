"""ctypes binding of the CPU oracle (oracle/liboracle.so) for the tests.

Test infrastructure only: the oracle is the checker, never the product.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_LIB = None

RHS_NBODY, RHS_EXPONENTIAL, RHS_VANDERPOL = 0, 1, 2
REDUCE_SEQUENTIAL, REDUCE_CHUNKED, REDUCE_THREADPOOL, REDUCE_HIP = 0, 1, 2, 3
HIP_ORDER = (REDUCE_HIP, 0)
POW_LIBM, POW_CORRECT = 0, 1
CONTROLLER_REFERENCE, CONTROLLER_PI = 0, 1
STATUS_REACHED_END, STATUS_MAX_ACCEPTED, STATUS_NONFINITE, STATUS_MAX_ATTEMPTS = 0, 1, 2, 3

_dp = C.POINTER(C.c_double)


def _ptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(_dp)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(ROOT, "oracle", "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        L = C.CDLL(path)
        L.oracle_rhs.argtypes = [C.c_int, _dp, _dp, C.c_size_t, C.c_double, _dp, C.c_int]
        L.oracle_rk_norm.argtypes = [_dp, _dp, C.c_size_t, C.c_int, C.c_long]
        L.oracle_rk_norm.restype = C.c_double
        L.oracle_integrate.argtypes = [C.c_char_p, C.c_int, _dp, C.c_size_t, C.c_double, C.c_double,
                                       _dp, _dp, C.c_double, _dp, C.c_int, C.c_long, C.c_int,
                                       C.c_long, C.c_long, C.c_int, C.c_int, C.c_int, C.c_int,
                                       _dp, C.c_long]
        L.oracle_integrate.restype = C.c_void_p
        L.oracle_result_n_dense.argtypes = [C.c_void_p]
        L.oracle_result_n_dense.restype = C.c_long
        L.oracle_result_dense.argtypes = [C.c_void_p, _dp, C.POINTER(C.c_long)]
        L.oracle_nr_dop853_dense.argtypes = ([C.c_int] + [C.c_double] * 9 +
                                             [C.c_int, C.c_long, C.POINTER(C.c_long)] +
                                             [_dp] * 4)
        L.oracle_dop853_step_dense.argtypes = [C.c_int, _dp, C.c_size_t, C.c_double,
                                               C.c_double, _dp, _dp, C.c_long, _dp, _dp]
        L.oracle_dopri5_step_dense.argtypes = L.oracle_dop853_step_dense.argtypes
        L.oracle_nr_dopr5.argtypes = L.oracle_nr_dop853_dense.argtypes
        L.oracle_register_tableau.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, _dp, _dp,
                                              _dp]
        L.oracle_result_count.argtypes = [C.c_void_p]
        L.oracle_result_count.restype = C.c_long
        L.oracle_result_times.argtypes = [C.c_void_p, _dp]
        L.oracle_result_state.argtypes = [C.c_void_p, C.c_long, _dp]
        L.oracle_result_final.argtypes = [C.c_void_p, _dp]
        L.oracle_result_stats.argtypes = [C.c_void_p, C.POINTER(C.c_long), _dp]
        L.oracle_result_trace_len.argtypes = [C.c_void_p]
        L.oracle_result_trace_len.restype = C.c_long
        L.oracle_result_trace.argtypes = [C.c_void_p, _dp, _dp, C.POINTER(C.c_int)]
        L.oracle_result_free.argtypes = [C.c_void_p]
        L.oracle_nr_dop853.argtypes = [C.c_int] + [C.c_double] * 8 + [C.c_double, C.POINTER(C.c_long), _dp]
        L.oracle_energy.argtypes = [_dp, C.c_size_t, C.c_double, _dp, _dp, _dp]
        L.oracle_momentum.argtypes = [_dp, C.c_size_t, _dp, _dp]
        L.oracle_tableau.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                     C.POINTER(C.c_int), _dp, _dp, _dp]
        L.oracle_spinning_cube_ic.argtypes = [C.c_long, _dp, _dp]
        L.oracle_pow.argtypes = [C.c_double, C.c_double, C.c_int]
        L.oracle_pow.restype = C.c_double
        L.oracle_leapfrog.argtypes = [_dp, C.c_size_t, C.c_double, C.c_double, C.c_double,
                                      C.c_double, _dp, C.c_int, C.c_int]
        L.oracle_leapfrog.restype = C.c_void_p
        L.oracle_mlp_forward.argtypes = [_dp, _dp, _dp, _dp, C.c_long, C.c_int, _dp]
        L.oracle_mlp_loss_grad.argtypes = [_dp, _dp, _dp, _dp, _dp, C.c_long, _dp]
        L.oracle_mlp_loss_grad.restype = C.c_double
        _LIB = L
    return _LIB


class Result:
    """Outputs of one oracle integration (RawOutput-like)."""

    def __init__(self, handle, n_var):
        L = lib()
        self.n_var = n_var
        n = L.oracle_result_count(handle)
        self.times = np.zeros(n)
        L.oracle_result_times(handle, _ptr(self.times))
        self.x_final = np.zeros(n_var)
        L.oracle_result_final(handle, _ptr(self.x_final))
        st = (C.c_long * 4)()
        ds = np.zeros(3)
        L.oracle_result_stats(handle, st, _ptr(ds))
        self.accepted, self.rejected, self.rhs_evals, self.status = list(st)
        self.t, self.dt, self.dt_initial = ds
        nt = L.oracle_result_trace_len(handle)
        self.trace_err = np.zeros(nt)
        self.trace_dt = np.zeros(nt)
        acc = (C.c_int * max(nt, 1))()
        if nt:
            L.oracle_result_trace(handle, _ptr(self.trace_err), _ptr(self.trace_dt), acc)
        self.trace_acc = np.array(list(acc)[:nt], dtype=np.int32)
        self._states = {}
        self._handle = handle
        nd = L.oracle_result_n_dense(handle)
        self.dense_states = np.zeros((nd, n_var))
        cnt = (C.c_long * 3)()
        L.oracle_result_dense(handle, _ptr(self.dense_states), cnt)
        self.n_dense, self.dense_steps, self.dense_rhs_evals = list(cnt)

    def state(self, k):
        if k < 0:
            k += len(self.times)
        out = np.zeros(self.n_var)
        if lib().oracle_result_state(self._handle, k, _ptr(out)) != 0:
            raise IndexError(k)
        return out

    def __del__(self):
        try:
            lib().oracle_result_free(self._handle)
        except Exception:
            pass


def integrate(tableau, rhs_kind, x0, t0, tf, atol, rtol, softening=0.0, masses=None,
              reduce=(REDUCE_SEQUENTIAL, 0), pow_mode=POW_LIBM, max_accepted=0,
              max_attempts=0, store_states=True, store_trace=False, threads=1,
              controller=CONTROLLER_REFERENCE, dense_times=None):
    x0 = np.ascontiguousarray(x0, dtype=np.float64)
    dt_ = None if dense_times is None else np.ascontiguousarray(dense_times, dtype=np.float64)
    n = x0.size
    atol = np.ascontiguousarray(np.broadcast_to(np.asarray(atol, dtype=np.float64), (n,)))
    rtol = np.ascontiguousarray(np.broadcast_to(np.asarray(rtol, dtype=np.float64), (n,)))
    m = None if masses is None else np.ascontiguousarray(masses, dtype=np.float64)
    h = lib().oracle_integrate(tableau.encode(), rhs_kind, _ptr(x0), n, t0, tf, _ptr(atol),
                               _ptr(rtol), softening, _ptr(m), reduce[0], reduce[1], pow_mode,
                               max_accepted, max_attempts, int(store_states), int(store_trace),
                               threads, controller, _ptr(dt_), 0 if dt_ is None else dt_.size)
    if not h:
        raise RuntimeError("oracle_integrate failed")
    return Result(h, n)


def leapfrog(x0, t0, tf, dt, softening=0.0, masses=None, store_states=True, threads=1):
    """ParallelLeapfrogIntegrator on an N-body state (accepted = steps)."""
    x0 = np.ascontiguousarray(x0, dtype=np.float64)
    m = None if masses is None else np.ascontiguousarray(masses, dtype=np.float64)
    h = lib().oracle_leapfrog(_ptr(x0), x0.size, t0, tf, dt, softening, _ptr(m),
                              int(store_states), threads)
    if not h:
        raise RuntimeError("oracle_leapfrog failed")
    return Result(h, x0.size)


def rhs(rhs_kind, x, softening=0.0, masses=None, threads=1):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros_like(x)
    m = None if masses is None else np.ascontiguousarray(masses, dtype=np.float64)
    if lib().oracle_rhs(rhs_kind, _ptr(x), _ptr(out), x.size, softening, _ptr(m), threads) != 0:
        raise RuntimeError("oracle_rhs failed")
    return out


def rk_norm(v, sc, reduce=(REDUCE_SEQUENTIAL, 0)):
    v = np.ascontiguousarray(v, dtype=np.float64)
    sc = np.ascontiguousarray(sc, dtype=np.float64)
    return lib().oracle_rk_norm(_ptr(v), _ptr(sc), v.size, reduce[0], reduce[1])


def nr_dop853(problem, y0, x1, x2, atol, rtol, h1, hmin=0.0, mu_eps=1e-3):
    cnt = C.c_long()
    sel = np.zeros(9)
    st = lib().oracle_nr_dop853(problem, y0[0], y0[1], x1, x2, atol, rtol, h1, hmin, mu_eps,
                                C.byref(cnt), _ptr(sel))
    return st, cnt.value, sel


def _nr_call(fn, problem, y0, x1, x2, atol, rtol, h1, intervals, hmin, mu_eps, cap):
    cnt = C.c_long()
    xs, y0s, y1s, yf = np.zeros(cap), np.zeros(cap), np.zeros(cap), np.full(2, np.nan)
    st = fn(problem, y0[0], y0[1], x1, x2, atol, rtol, h1, hmin, mu_eps, intervals, cap,
            C.byref(cnt), _ptr(xs), _ptr(y0s), _ptr(y1s), _ptr(yf))
    c = min(cnt.value, cap)
    return st, xs[:c], np.stack([y0s[:c], y1s[:c]], axis=1), yf


def nr_dop853_dense(problem, y0, x1, x2, atol, rtol, h1, intervals, hmin=0.0, mu_eps=1e-3,
                    cap=100000):
    """NR DOP853 with the DenseOutput policy: (status, xs, ys (count x 2), y_final)."""
    return _nr_call(lib().oracle_nr_dop853_dense, problem, y0, x1, x2, atol, rtol, h1,
                    intervals, hmin, mu_eps, cap)


def nr_dopr5(problem, y0, x1, x2, atol, rtol, h1, intervals=0, hmin=0.0, mu_eps=1e-3,
             cap=100000):
    """NR DOPRI5 with the RawOutput (intervals = 0) or DenseOutput policy:
    (status, xs, ys (count x 2), y_final)."""
    return _nr_call(lib().oracle_nr_dopr5, problem, y0, x1, x2, atol, rtol, h1, intervals,
                    hmin, mu_eps, cap)


def step_dense(tableau, rhs_kind, y, h, s_list, softening=0.0, masses=None):
    """One RK-form step (DOP853 or DOPRI5) of length h and its continuous
    extension at the fractions s_list: (yout, dense (len(s_list) x n))."""
    fn = {"DOP853": lib().oracle_dop853_step_dense,
          "DOPRI5": lib().oracle_dopri5_step_dense}[tableau]
    y = np.ascontiguousarray(y, dtype=np.float64)
    s_list = np.ascontiguousarray(s_list, dtype=np.float64)
    m = None if masses is None else np.ascontiguousarray(masses, dtype=np.float64)
    yout = np.zeros_like(y)
    out = np.zeros((s_list.size, y.size))
    if fn(rhs_kind, _ptr(y), y.size, h, softening, _ptr(m), _ptr(s_list), s_list.size,
          _ptr(yout), _ptr(out)) != 0:
        raise RuntimeError("oracle step_dense failed")
    return yout, out


def dop853_step_dense(rhs_kind, y, h, s_list, softening=0.0, masses=None):
    return step_dense("DOP853", rhs_kind, y, h, s_list, softening, masses)


def energy(x, n, softening, masses=None):
    x = np.ascontiguousarray(x, dtype=np.float64)
    ke, pe = np.zeros(1), np.zeros(1)
    m = None if masses is None else np.ascontiguousarray(masses, dtype=np.float64)
    lib().oracle_energy(_ptr(x), n, softening, _ptr(m), _ptr(ke), _ptr(pe))
    return float(ke[0]), float(pe[0])


def momentum(x, n, masses=None):
    x = np.ascontiguousarray(x, dtype=np.float64)
    p = np.zeros(3)
    m = None if masses is None else np.ascontiguousarray(masses, dtype=np.float64)
    lib().oracle_momentum(_ptr(x), n, _ptr(m), _ptr(p))
    return p


def tableau(name):
    S, p, pt = C.c_int(), C.c_int(), C.c_int()
    if lib().oracle_tableau(name.encode(), C.byref(S), C.byref(p), C.byref(pt), None, None, None):
        raise KeyError(name)
    s = S.value
    a = np.zeros((s - 1) * (s - 1))
    b = np.zeros(s)
    bt = np.zeros(s)
    lib().oracle_tableau(name.encode(), C.byref(S), C.byref(p), C.byref(pt), _ptr(a), _ptr(b), _ptr(bt))
    return dict(S=s, p=p.value, pt=pt.value, a=a.reshape(s - 1, s - 1), b=b, bt=bt)


def register_tableau(name, a, b, bt, p, pt):
    """Make a tableau (the reference's struct fields) available to integrate() by name."""
    b = np.ascontiguousarray(b, dtype=np.float64)
    bt = np.ascontiguousarray(bt, dtype=np.float64)
    s = b.size
    full = np.zeros((s - 1, s - 1))
    for r, row in enumerate(a):
        full[r, :r + 1] = list(row)[:r + 1]
    if lib().oracle_register_tableau(name.encode(), s, int(p), int(pt), _ptr(full.ravel()),
                                     _ptr(b), _ptr(bt)):
        raise ValueError("oracle_register_tableau failed")


def pow_(x, y, mode=POW_LIBM):
    return lib().oracle_pow(x, y, mode)


def spinning_cube_ic(n):
    """SpinningCubeNBodyTest<n> state and (softening, tf, tol)."""
    x = np.zeros(6 * n)
    p = np.zeros(3)
    lib().oracle_spinning_cube_ic(n, _ptr(x), _ptr(p))
    return x, float(p[0]), float(p[1]), float(p[2])


MLP_IN, MLP_HID = 10, 64
MLP_PARAMS = MLP_HID * MLP_IN + MLP_HID + MLP_HID * MLP_HID + MLP_HID + MLP_HID + 1


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def mlp_forward(params, mean, std, raw, bf16=False):
    raw = _f64(raw).reshape(-1, MLP_IN)
    out = np.zeros(raw.shape[0])
    lib().oracle_mlp_forward(_ptr(_f64(params)), _ptr(_f64(mean)), _ptr(_f64(std)),
                             _ptr(raw), raw.shape[0], int(bf16), _ptr(out))
    return out


def mlp_loss_grad(params, mean, std, raw, target):
    raw = _f64(raw).reshape(-1, MLP_IN)
    g = np.zeros(MLP_PARAMS)
    loss = lib().oracle_mlp_loss_grad(_ptr(_f64(params)), _ptr(_f64(mean)), _ptr(_f64(std)),
                                      _ptr(raw), _ptr(_f64(target)), raw.shape[0], _ptr(g))
    return loss, g
