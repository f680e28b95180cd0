"""odehip_rk_norm_noalloc: the error norm with no allocation (its group partials in a
static device array of the library), through the C ABI, against odehip_rk_norm.

Two calls in a row, the second on a smaller input (the array keeps the first call's
partials), each bit for bit as odehip_rk_norm on the same vectors, which itself is
checked against the oracle's canonical-order norm; the largest accepted size; and the
size limit rejected with a clear error.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

MAX_N = 1 << 24  # ODEHIP_NOALLOC_MAX_N_VAR


def norms(exe, v, sc):
    import ode_prototyping_amd as oh
    vd, sd = oh.HipState(exe, data=v), oh.HipState(exe, data=sc)
    a, b = C.c_double(), C.c_double()
    exe.check(oh.lib().odehip_rk_norm_noalloc(exe.handle, vd.ptr, sd.ptr, v.size, C.byref(a)),
              "rk_norm_noalloc")
    exe.check(oh.lib().odehip_rk_norm(exe.handle, vd.ptr, sd.ptr, v.size, C.byref(b)),
              "rk_norm")
    return a.value, b.value


def test_noalloc_norm_matches_rk_norm_twice(exe):
    rng = np.random.default_rng(5)
    for n in (1_000_003, 70_001):  # second call smaller than the first
        v = rng.standard_normal(n) * 1e-9
        sc = 1e-10 + 1e-10 * np.abs(rng.standard_normal(n))
        got, want = norms(exe, v, sc)
        ref = O.rk_norm(v, sc, O.HIP_ORDER)
        print(f"n={n}: noalloc {got!r} rk_norm {want!r} oracle {ref!r}")
        assert np.float64(got).view(np.uint64) == np.float64(want).view(np.uint64)
        assert got == ref


def test_noalloc_norm_largest_size_and_limit(exe):
    import ode_prototyping_amd as oh
    v = np.full(MAX_N, 3e-10)
    sc = np.full(MAX_N, 1e-10)
    got, want = norms(exe, v, sc)
    assert got == want and abs(got - 3.0) < 1e-12
    small = np.array([1.0, -2.0, 2.0])
    got, want = norms(exe, small, np.ones(3))
    assert got == want == np.sqrt(3.0)
    big = oh.HipState(exe, MAX_N + 1)
    out = C.c_double()
    rc = oh.lib().odehip_rk_norm_noalloc(exe.handle, big.ptr, big.ptr, MAX_N + 1, C.byref(out))
    assert rc == oh._lib.ERR_INVALID
    assert b"ODEHIP_NOALLOC_MAX_N_VAR" in oh.lib().odehip_last_error(exe.handle)
