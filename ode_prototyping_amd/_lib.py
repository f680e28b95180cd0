"""ctypes declarations of libodehip.so (include/odehip.h).

The library is built in-tree (``__graft_entry__.build()`` or ``make -C
ode_prototyping_amd``); importing this module never builds anything and there
is no CPU fallback: if the shared library is missing, :func:`lib` raises.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libodehip.so")

OK = 0
ERR_INVALID, ERR_HIP, ERR_NOMEM, ERR_UNSUPPORTED = -1, -2, -3, -4

TABLEAUS = {"HE21": 0, "RKF45": 1, "DOPRI5": 2, "DVERK": 3, "RKF78": 4, "DOP853": 5}
STAGES = {"HE21": 2, "RKF45": 6, "DOPRI5": 7, "DVERK": 8, "RKF78": 13, "DOP853": 12}
ORDERS = {"HE21": (2, 1), "RKF45": (4, 5), "DOPRI5": (5, 4), "DVERK": (6, 5), "RKF78": (7, 8),
          "DOP853": (8, 5)}  # (p, pt)
MAX_STAGES = 13  # ODEHIP_MAX_STAGES
RHS_NBODY, RHS_EXPONENTIAL, RHS_VANDERPOL = 0, 1, 2
NBODY_VARIANTS = {"exact": 0, "fast": 1, "mixed": 2}
MIXED_NEAR_DEFAULT = 0.0  # ODEHIP_MIXED_NEAR_DEFAULT
MIXED_NEAR_CLOSE = 20.0  # ODEHIP_MIXED_NEAR_CLOSE
CONTROLLERS = {"reference": 0, "pi": 1, "learned": 2}
STATUS = {-1: "running", 0: "reached_end", 1: "max_accepted", 2: "nonfinite", 3: "max_attempts"}
REDUCE_GROUP = 256

_dp = C.POINTER(C.c_double)
_ctxp = C.c_void_p


class OdeStruct(C.Structure):
    _fields_ = [("rhs_kind", C.c_int), ("nbody_variant", C.c_int), ("n_var", C.c_size_t),
                ("softening", C.c_double), ("masses", C.c_void_p)]


class SolveOptions(C.Structure):
    _fields_ = [("t0", C.c_double), ("tf", C.c_double), ("atol", C.c_void_p),
                ("rtol", C.c_void_p), ("max_accepted", C.c_long), ("max_attempts", C.c_long),
                ("fsal", C.c_int), ("attempts_per_launch", C.c_int),
                ("times", C.c_void_p), ("times_cap", C.c_long),
                ("trace", C.c_void_p), ("trace_cap", C.c_long),
                ("controller", C.c_int), ("mlp", C.c_void_p),
                ("states", C.c_void_p), ("states_cap", C.c_long),
                ("dense_times", C.c_void_p), ("n_dense", C.c_long),
                ("dense_states", C.c_void_p)]


class SolveStats(C.Structure):
    _fields_ = [("accepted", C.c_long), ("rejected", C.c_long), ("attempts", C.c_long),
                ("rhs_evals", C.c_long), ("n_times", C.c_long), ("n_trace", C.c_long),
                ("status", C.c_int), ("t", C.c_double), ("dt", C.c_double),
                ("dt_initial", C.c_double), ("last_err", C.c_double),
                ("learned_steps", C.c_long), ("n_states", C.c_long),
                ("n_dense", C.c_long), ("dense_steps", C.c_long),
                ("dense_rhs_evals", C.c_long)]


class EnsembleOptions(C.Structure):
    _fields_ = [("t0", C.c_double), ("tf", C.c_double), ("atol", C.c_double),
                ("rtol", C.c_double), ("max_accepted", C.c_long), ("max_attempts", C.c_long),
                ("fsal", C.c_int), ("rounds_per_launch", C.c_int), ("tile_config", C.c_int),
                ("controller", C.c_int), ("mlp", C.c_void_p), ("trace", C.c_void_p),
                ("trace_cap", C.c_long), ("tune_ms", C.c_void_p),
                ("dense_times", C.c_void_p), ("n_dense", C.c_long),
                ("dense_states", C.c_void_p)]


class SystemStats(C.Structure):
    _fields_ = [("accepted", C.c_long), ("rejected", C.c_long), ("attempts", C.c_long),
                ("rhs_evals", C.c_long), ("n_trace", C.c_long), ("learned_steps", C.c_long),
                ("status", C.c_int), ("pad", C.c_int), ("t", C.c_double), ("dt", C.c_double),
                ("dt_initial", C.c_double), ("last_err", C.c_double),
                ("n_dense", C.c_long)]


class EnsembleStats(C.Structure):
    _fields_ = [("rounds", C.c_long), ("accepted", C.c_long), ("rejected", C.c_long),
                ("attempts", C.c_long), ("rhs_evals", C.c_long), ("learned_steps", C.c_long),
                ("particle_steps", C.c_double), ("interactions", C.c_double),
                ("tile_config", C.c_int), ("n_unfinished", C.c_int), ("rhs_ms", C.c_double)]


class MlpTrainOptions(C.Structure):
    _fields_ = [("epochs", C.c_int), ("batch", C.c_int), ("lr", C.c_float),
                ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float),
                ("seed", C.c_uint64)]


MLP_IN, MLP_HIDDEN, MLP_PARAMS = 10, 64, 4929
_fp = C.POINTER(C.c_float)

class LeapfrogStats(C.Structure):
    _fields_ = [("steps", C.c_long), ("rhs_evals", C.c_long), ("n_states", C.c_long),
                ("t", C.c_double)]


ENS_TILE_CONFIGS = 18
ENS_TILE_UNSPLIT = 12  # configurations ENS_TILE_UNSPLIT.. share a tile among 2, 4, 8 waves
ENS_TILE_AUTO = -1

_ip = C.POINTER(C.c_int)

# name -> (restype, argtypes)
SIGNATURES = {
    "odehip_version": (C.c_char_p, []),
    "odehip_create": (C.c_int, [C.c_int, C.POINTER(_ctxp)]),
    "odehip_destroy": (C.c_int, [_ctxp]),
    "odehip_last_error": (C.c_char_p, [_ctxp]),
    "odehip_set_stream": (C.c_int, [_ctxp, C.c_void_p]),
    "odehip_set_near_field": (C.c_int, [_ctxp, C.c_double]),
    "odehip_get_near_field": (C.c_int, [_ctxp, _dp]),
    "odehip_synchronize": (C.c_int, [_ctxp]),
    "odehip_timer_start": (C.c_int, [_ctxp]),
    "odehip_timer_stop": (C.c_int, [_ctxp, _dp]),
    "odehip_probe_fp64": (C.c_int, [_ctxp, _dp, _dp]),
    "odehip_alloc": (C.c_int, [_ctxp, C.c_size_t, C.POINTER(C.c_void_p)]),
    "odehip_free": (C.c_int, [_ctxp, C.c_void_p]),
    "odehip_copy_h2d": (C.c_int, [_ctxp, C.c_void_p, _dp, C.c_size_t]),
    "odehip_copy_d2h": (C.c_int, [_ctxp, _dp, C.c_void_p, C.c_size_t]),
    "odehip_copy_d2d": (C.c_int, [_ctxp, C.c_void_p, C.c_void_p, C.c_size_t]),
    "odehip_fill": (C.c_int, [_ctxp, C.c_void_p, C.c_double, C.c_size_t]),
    "odehip_nbody_rhs": (C.c_int, [_ctxp, C.c_void_p, C.c_void_p, C.c_size_t, C.c_double,
                                   C.c_void_p, C.c_int]),
    "odehip_nbody_acc": (C.c_int, [_ctxp, C.c_void_p, C.c_void_p, C.c_size_t, C.c_double,
                                   C.c_void_p, C.c_int]),
    "odehip_ode_rhs": (C.c_int, [_ctxp, C.c_int, C.c_void_p, C.c_void_p, C.c_size_t]),
    "odehip_tableau_register": (C.c_int, [_ctxp, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp,
                                          C.POINTER(C.c_int)]),
    "odehip_tableau_info": (C.c_int, [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                      C.POINTER(C.c_int)]),
    "odehip_rk_stage": (C.c_int, [_ctxp, C.c_int, C.c_int, C.c_double, C.c_void_p, C.c_void_p,
                                  C.c_void_p, C.c_size_t]),
    "odehip_rk_update_state_and_error": (C.c_int, [_ctxp, C.c_int, C.c_double, C.c_void_p,
                                                   C.c_void_p, C.c_void_p, C.c_void_p,
                                                   C.c_size_t]),
    "odehip_rk_update_error_target": (C.c_int, [_ctxp, C.c_void_p, C.c_void_p, C.c_void_p,
                                                C.c_void_p, C.c_void_p, C.c_size_t]),
    "odehip_rk_norm": (C.c_int, [_ctxp, C.c_void_p, C.c_void_p, C.c_size_t, _dp]),
    "odehip_rk_norm_noalloc": (C.c_int, [_ctxp, C.c_void_p, C.c_void_p, C.c_size_t, _dp]),
    "odehip_rk_compute_error_target": (C.c_int, [_ctxp, C.c_void_p, C.c_void_p, C.c_void_p,
                                                 C.c_void_p, C.c_size_t]),
    "odehip_rk_euler_step": (C.c_int, [_ctxp, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
                                       C.c_size_t]),
    "odehip_rk_difference": (C.c_int, [_ctxp, C.c_void_p, C.c_void_p, C.c_size_t]),
    "odehip_rk_initial_step": (C.c_int, [_ctxp, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.c_size_t, C.c_double, C.c_void_p,
                                         C.c_int, _dp]),
    "odehip_rk_solve": (C.c_int, [_ctxp, C.c_int, C.POINTER(OdeStruct),
                                  C.POINTER(SolveOptions), C.c_void_p, C.POINTER(SolveStats)]),
    "odehip_ensemble_solve": (C.c_int, [_ctxp, C.c_int, C.c_int, _ip, C.c_int, C.c_double,
                                        C.c_void_p, C.POINTER(EnsembleOptions), C.c_void_p,
                                        C.c_void_p, C.POINTER(EnsembleStats),
                                        C.POINTER(SystemStats)]),
    "odehip_ensemble_rhs": (C.c_int, [_ctxp, C.c_int, _ip, C.c_int, C.c_double, C.c_void_p,
                                      C.c_int, C.c_void_p, C.c_void_p, _dp]),
    "odehip_mlp_create": (C.c_int, [_ctxp, _fp, C.c_uint64, _fp, _fp, C.POINTER(C.c_void_p)]),
    "odehip_mlp_destroy": (C.c_int, [C.c_void_p]),
    "odehip_mlp_get_params": (C.c_int, [_ctxp, C.c_void_p, _fp]),
    "odehip_mlp_loss_grad": (C.c_int, [_ctxp, C.c_void_p, C.c_void_p, C.c_void_p, C.c_long,
                                       _fp, _dp]),
    "odehip_mlp_train": (C.c_int, [_ctxp, C.c_void_p, C.c_void_p, C.c_void_p, C.c_long,
                                   C.POINTER(MlpTrainOptions), _dp]),
    "odehip_mlp_infer": (C.c_int, [_ctxp, C.c_void_p, C.c_void_p, C.c_long, C.c_void_p]),
    "odehip_leapfrog": (C.c_int, [_ctxp, C.POINTER(OdeStruct), C.c_double, C.c_double,
                                  C.c_double, C.c_void_p, _dp, _dp, C.c_long,
                                  C.POINTER(LeapfrogStats)]),
    "odehip_write_trajectory": (C.c_int, [C.c_char_p, C.c_size_t, C.c_size_t, C.c_double,
                                          _dp, _dp]),
    "odehip_nbody_energy": (C.c_int, [_ctxp, C.c_void_p, C.c_size_t, C.c_double, C.c_void_p,
                                      _dp]),
    "odehip_nbody_momentum": (C.c_int, [_ctxp, C.c_void_p, C.c_size_t, C.c_void_p, _dp]),
    "odehip_ic_uniform_cube": (C.c_int, [C.c_size_t, C.c_uint64, C.c_double, C.c_double,
                                         C.c_double, _dp, _dp]),
    "odehip_ic_plummer": (C.c_int, [C.c_size_t, C.c_uint64, C.c_double, C.c_double, _dp, _dp]),
    "odehip_ic_colliding_plummer": (C.c_int, [C.c_size_t, C.c_uint64, C.c_double, C.c_double,
                                              C.c_double, C.c_double, _dp, _dp]),
}

_LIB = None


def lib():
    """Load libodehip.so (raises RuntimeError if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `make -C ode_prototyping_amd` "
                "(or __graft_entry__.build()); there is no CPU fallback")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def dptr(a):
    """ctypes double* of a contiguous float64 numpy array (or None)."""
    if a is None:
        return None
    return a.ctypes.data_as(_dp)
