"""Learned step-size controller: network handle, training data from solver
traces, training on the GPU.

The reference's controller (``RKEmbeddedParallel.hpp:64-79``) sizes the next
step as ``dt (0.38 / err)^(1/k)``, i.e. it assumes the error constant of the
last step holds for the next.  The learned controller replaces that proposal
with a network that sees the last three attempts (error norms, step sizes,
accept flags) and the system size; the embedded error estimate still decides on
every step and the PI controller takes over before three attempts of history
exist, after two rejections in a row and on a non-finite output (see
``include/odehip.h``).

Training target.  After attempt ``a`` the ideal next step is the one that
would have produced the error ``tau`` at attempt ``a + 1``; under the local
error model ``err ~ C dt^k`` observed at attempt ``a + 1`` that step is
``dt_{a+1} (tau / err_{a+1})^(1/k)``.  The network regresses
``log(dt*_{a+1} / dt_a)`` (clamped to ``[log 0.2, log 5]``) from the features
at attempt ``a``, on traces that the PI controller generates.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import lib
from .core import HipExecutor, HipState

MLP_IN, MLP_HIDDEN, MLP_PARAMS = _lib.MLP_IN, _lib.MLP_HIDDEN, _lib.MLP_PARAMS
LOG_RATIO_MIN, LOG_RATIO_MAX = np.log(0.2), np.log(5.0)


def f32_device(exe, arr):
    """A HipState holding the bytes of a float32 array (padded to 8 bytes)."""
    a = np.ascontiguousarray(np.asarray(arr, dtype=np.float32).ravel())
    if a.size % 2:
        a = np.concatenate([a, np.zeros(1, np.float32)])
    return HipState(exe, data=a.view(np.float64))


def f32_host(state, n):
    """The first n float32 values held by a HipState."""
    return state.copy_to_span().view(np.float32)[:n].copy()


class StepMLP:
    """The controller network on one device (an ``odehip_mlp`` handle)."""

    def __init__(self, exe, mean, std, params=None, seed=0):
        self.exe = exe
        mean = np.ascontiguousarray(mean, dtype=np.float32)
        std = np.ascontiguousarray(std, dtype=np.float32)
        if mean.size != MLP_IN or std.size != MLP_IN:
            raise ValueError("mean and std need 10 entries")
        p = None
        if params is not None:
            p = np.ascontiguousarray(params, dtype=np.float32)
            if p.size != MLP_PARAMS:
                raise ValueError(f"params need {MLP_PARAMS} entries")
        h = C.c_void_p()
        fp = C.POINTER(C.c_float)
        exe.check(lib().odehip_mlp_create(exe.handle,
                                          p.ctypes.data_as(fp) if p is not None else None,
                                          int(seed), mean.ctypes.data_as(fp),
                                          std.ctypes.data_as(fp), C.byref(h)), "mlp_create")
        self._h = h
        self.mean, self.std = mean, std

    @property
    def handle(self):
        return self._h

    def params(self):
        out = np.zeros(MLP_PARAMS, np.float32)
        self.exe.check(lib().odehip_mlp_get_params(self.exe.handle, self._h,
                                                   out.ctypes.data_as(C.POINTER(C.c_float))),
                       "mlp_get_params")
        return out

    def loss_grad(self, x, t):
        """Loss 0.5 mean((y - t)^2) and its gradient (FP32 kernels)."""
        x = np.asarray(x, np.float32).reshape(-1, MLP_IN)
        xd, td = f32_device(self.exe, x), f32_device(self.exe, t)
        g = np.zeros(MLP_PARAMS, np.float32)
        loss = C.c_double()
        self.exe.check(lib().odehip_mlp_loss_grad(self.exe.handle, self._h, xd.ptr, td.ptr,
                                                  x.shape[0],
                                                  g.ctypes.data_as(C.POINTER(C.c_float)),
                                                  C.byref(loss)), "mlp_loss_grad")
        return loss.value, g

    def train(self, x, t, epochs=10, batch=4096, lr=1e-3, seed=0):
        """Adam on the GPU; returns the mean loss of every epoch."""
        x = np.asarray(x, np.float32).reshape(-1, MLP_IN)
        xd, td = f32_device(self.exe, x), f32_device(self.exe, t)
        o = _lib.MlpTrainOptions(int(epochs), int(batch), float(lr), 0.0, 0.0, 0.0, int(seed))
        losses = np.zeros(max(int(epochs), 1))
        self.exe.check(lib().odehip_mlp_train(self.exe.handle, self._h, xd.ptr, td.ptr,
                                              x.shape[0], C.byref(o), _lib.dptr(losses)),
                       "mlp_train")
        return losses[:epochs]

    def infer(self, x):
        """Outputs for raw features x (BF16 matrix-core kernel)."""
        x = np.asarray(x, np.float32).reshape(-1, MLP_IN)
        xd = f32_device(self.exe, x)
        yd = f32_device(self.exe, np.zeros(x.shape[0], np.float32))
        self.exe.check(lib().odehip_mlp_infer(self.exe.handle, self._h, xd.ptr, x.shape[0],
                                              yd.ptr), "mlp_infer")
        return f32_host(yd, x.shape[0])

    def close(self):
        if getattr(self, "_h", None):
            lib().odehip_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def trace_features(trace, n_particles, span):
    """Raw features after every attempt of one trace (rows (err, dt, acc)),
    exactly as the device computes them; rows 0 and 1 lack history."""
    err = np.clip(trace[:, 0], 1e-12, 1e12)
    dt = trace[:, 1]
    acc = trace[:, 2]
    m = trace.shape[0]
    f = np.zeros((m, MLP_IN))
    with np.errstate(divide="ignore", invalid="ignore"):
        for h in range(3):
            f[h:, h] = np.log(err[:m - h])
            f[h:, 6 + h] = acc[:m - h]
        f[:, 3] = np.log(dt / span)
        f[1:, 4] = np.log(dt[:-1] / dt[1:])
        f[2:, 5] = np.log(dt[:-2] / dt[2:])
    f[:, 9] = np.log(float(n_particles))
    f[~np.isfinite(f)] = 0.0
    return f


def training_pairs(traces, sizes, span, k, tau=0.5):
    """(features, target) pairs from per-system traces: features after
    attempt a >= 2, target log(dt_{a+1} / dt_a) + log(tau / err_{a+1}) / k
    clamped to [log 0.2, log 5]."""
    xs, ts = [], []
    for tr, n in zip(traces, sizes):
        if tr is None or tr.shape[0] < 4:
            continue
        f = trace_features(tr, n, span)
        err = np.clip(tr[:, 0], 1e-12, 1e12)
        dt = tr[:, 1]
        a = np.arange(2, tr.shape[0] - 1)
        tgt = np.log(dt[a + 1] / dt[a]) + np.log(tau / err[a + 1]) / k
        ok = np.isfinite(tgt)
        xs.append(f[a][ok])
        ts.append(np.clip(tgt[ok], LOG_RATIO_MIN, LOG_RATIO_MAX))
    x = np.concatenate(xs).astype(np.float32)
    t = np.concatenate(ts).astype(np.float32)
    return x, t


def normalisation(x):
    mean = x.mean(axis=0)
    std = x.std(axis=0)
    std = np.where(std > 1e-3, std, 1.0)
    return mean.astype(np.float32), std.astype(np.float32)
