"""Batched ensembles of independent N-body systems.

The reference integrates one system per ``RKEmbeddedParallel<BT>::integrate``
call (``src/integrator/RKEmbeddedParallel/RKEmbeddedParallel.hpp:13-81``) with
``NBodyODE`` (``src/ode/NBodyODE.hpp``).  :func:`solve_ensemble` makes that
call for many systems of different sizes at once: every system keeps its own
step size, accept/reject sequence, counters and end condition, and its result
is bit-identical to :func:`ode_prototyping_amd.integrate` on that system alone
with the same options.

States are packed: system k occupies ``6 n_k`` doubles ``[pos | vel]``
starting at ``6 (n_0 + ... + n_{k-1})``; masses (optional) likewise with
``n_k`` entries per system.  :func:`pack` / :func:`unpack` convert from / to
lists of per-system arrays.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import CONTROLLERS, NBODY_VARIANTS, STATUS, lib
from .core import HipExecutor, HipState, tableau_id, tableau_shape

ENS_TILE_CONFIGS = _lib.ENS_TILE_CONFIGS
ENS_TILE_UNSPLIT = _lib.ENS_TILE_UNSPLIT
ENS_TILE_AUTO = _lib.ENS_TILE_AUTO


def pack(arrays):
    """Concatenate per-system arrays into one packed float64 array."""
    return np.ascontiguousarray(np.concatenate([np.asarray(a, np.float64).ravel()
                                                for a in arrays]))


def unpack(packed, sizes, per_particle=6):
    """Split a packed array into per-system views (6 n_k or n_k entries)."""
    out, off = [], 0
    for n in sizes:
        m = per_particle * int(n)
        out.append(packed[off:off + m])
        off += m
    if off != packed.size:
        raise ValueError(f"packed array has {packed.size} entries, sizes imply {off}")
    return out


def _sizes_arg(sizes):
    sz = np.ascontiguousarray(np.asarray(sizes, dtype=np.int32))
    if sz.ndim != 1 or sz.size == 0:
        raise ValueError("sizes must be a non-empty 1-D sequence")
    return sz, sz.ctypes.data_as(C.POINTER(C.c_int))


def _device(exe, v, n, name):
    if isinstance(v, HipState):
        if v.size != n:
            raise ValueError(f"{name} has {v.size} entries, expected {n}")
        return v
    arr = np.ascontiguousarray(np.asarray(v, np.float64).ravel())
    if arr.size != n:
        raise ValueError(f"{name} has {arr.size} entries, expected {n}")
    return HipState(exe, data=arr)


class EnsembleResult:
    """Per-system counters (numpy arrays indexed by system) and totals."""

    def __init__(self, sizes, per, agg, x_final, trace, tune_ms):
        self.sizes = np.asarray(sizes)
        for f, _ in _lib.SystemStats._fields_:
            if f == "pad":
                continue
            setattr(self, f, np.array([getattr(p, f) for p in per]))
        self.status_names = [STATUS.get(int(s), str(s)) for s in self.status]
        self.rounds = agg.rounds
        self.total_accepted = agg.accepted
        self.total_rejected = agg.rejected
        self.total_attempts = agg.attempts
        self.total_rhs_evals = agg.rhs_evals
        self.total_learned_steps = agg.learned_steps
        self.particle_steps = agg.particle_steps
        self.interactions = agg.interactions
        self.tile_config = agg.tile_config
        self.n_unfinished = agg.n_unfinished
        self.rhs_ms = agg.rhs_ms
        self.tune_ms = tune_ms
        self.x_final = x_final
        self.trace = trace
        self.dense_times = None
        self.dense_states = None  # n_dense x packed state (NaN where not reached)

    def states(self):
        """Final states as a list of per-system arrays."""
        return unpack(self.x_final, self.sizes)

    def dense_rows(self, k):
        """System k's dense-output rows (n_dense[k] x 6 n_k)."""
        off = 6 * int(self.sizes[:k].sum())
        n = 6 * int(self.sizes[k])
        return self.dense_states[:int(self.n_dense[k]), off:off + n]

    def __repr__(self):
        return (f"EnsembleResult(systems={self.sizes.size}, rounds={self.rounds}, "
                f"accepted={self.total_accepted}, rejected={self.total_rejected}, "
                f"rhs_evals={self.total_rhs_evals})")


def solve_ensemble(tableau, sizes, x0, t0, tf, atol, rtol, softening, masses=None, *,
                   variant="exact", exe=None, max_accepted=0, max_attempts=0, fsal=False,
                   rounds_per_launch=0, tile_config=ENS_TILE_AUTO, controller="reference",
                   mlp=None, trace_cap=0, x_out=None, fetch=True, dense_times=None):
    """Solve every system of the ensemble from t0 to tf (device-resident).

    ``x0``: packed host array or HipState; ``masses``: packed (n_k per system)
    host array, HipState or None (unit masses); ``atol``/``rtol``: scalars.
    ``mlp``: a trained :class:`ode_prototyping_amd.controller.StepMLP` for
    ``controller="learned"``.  Returns an :class:`EnsembleResult`; the final
    packed state is in ``x_final`` (host) unless ``fetch`` is False (then it
    stays in ``x_out``).  ``dense_times`` (DOPRI5, DOP853): common output times after
    t0; every system's state there, from its continuous extension, is in
    ``dense_states`` (packed rows) and ``dense_rows(k)``.
    """
    tableau_shape(tableau)  # a known tableau (ValueError otherwise)
    if variant not in NBODY_VARIANTS:
        raise ValueError(f"unknown variant {variant}")
    if controller not in CONTROLLERS:
        raise ValueError(f"unknown controller {controller}")
    exe = exe or HipExecutor()
    sz, szp = _sizes_arg(sizes)
    nv = 6 * int(sz.sum())
    xin = _device(exe, x0, nv, "x0")
    m = None if masses is None else _device(exe, masses, int(sz.sum()), "masses")
    out = x_out if x_out is not None else HipState(exe, nv)
    trace = np.zeros((sz.size, max(int(trace_cap), 1), 3))
    tune = np.zeros(ENS_TILE_CONFIGS)
    dt_arr = dense = None
    if dense_times is not None and len(dense_times) > 0:
        dt_arr = np.ascontiguousarray(dense_times, dtype=np.float64)
        if dt_arr.size * nv * 8 > (16 << 30):
            raise ValueError("the dense output would exceed 16 GiB")
        dense = np.zeros((dt_arr.size, nv))
    opt = _lib.EnsembleOptions(float(t0), float(tf), float(atol), float(rtol),
                               int(max_accepted), int(max_attempts), 1 if fsal else 0,
                               int(rounds_per_launch), int(tile_config),
                               CONTROLLERS[controller],
                               mlp.handle if mlp is not None else None,
                               trace.ctypes.data if trace_cap > 0 else None, int(trace_cap),
                               tune.ctypes.data,
                               dt_arr.ctypes.data if dt_arr is not None else None,
                               int(dt_arr.size) if dt_arr is not None else 0,
                               dense.ctypes.data if dense is not None else None)
    agg = _lib.EnsembleStats()
    per = (_lib.SystemStats * sz.size)()
    exe.check(lib().odehip_ensemble_solve(exe.handle, tableau_id(exe, tableau), int(sz.size),
                                          szp,
                                          NBODY_VARIANTS[variant], float(softening),
                                          m.ptr if m is not None else None, C.byref(opt),
                                          xin.ptr, out.ptr, C.byref(agg), per),
              "ensemble_solve")
    x_final = out.copy_to_span() if fetch else None
    tr = None
    if trace_cap > 0:
        tr = [trace[k, :per[k].n_trace].copy() for k in range(sz.size)]
    res = EnsembleResult(sz, list(per), agg, x_final, tr, tune)
    if dt_arr is not None:
        res.dense_times = dt_arr
        res.dense_states = dense
    return res


def ensemble_rhs(exe, sizes, x, softening, masses=None, variant="exact",
                 tile_config=ENS_TILE_AUTO, dxdt=None):
    """One batched right-hand side of every system (packed in and out).

    Returns ``(dxdt, ms)``: the derivative on the host (or written into the
    HipState ``dxdt`` and returned as that) and the kernel time in ms.
    """
    if variant not in NBODY_VARIANTS:
        raise ValueError(f"unknown variant {variant}")
    sz, szp = _sizes_arg(sizes)
    nv = 6 * int(sz.sum())
    xin = _device(exe, x, nv, "x")
    m = None if masses is None else _device(exe, masses, int(sz.sum()), "masses")
    out = dxdt if dxdt is not None else HipState(exe, nv)
    ms = C.c_double()
    exe.check(lib().odehip_ensemble_rhs(exe.handle, int(sz.size), szp, NBODY_VARIANTS[variant],
                                        float(softening), m.ptr if m is not None else None,
                                        int(tile_config), xin.ptr, out.ptr, C.byref(ms)),
              "ensemble_rhs")
    return (out if dxdt is not None else out.copy_to_span()), ms.value
