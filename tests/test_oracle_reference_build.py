"""The oracle restatement against the reference's own code (CPU).

oracle/_ref/libref.so is built by `make -C oracle ref` from the reference's
headers where they lie (oracle/ref_driver.cpp instantiates them); these tests
are skipped when it has not been built.  Every comparison is bit for bit:
the N-body RHS at N = 64 and 1024, whole DOPRI5 and DOP853 (RK form) solves
with rejections, and the leapfrog integrator.
"""
import ctypes as C
import os

import numpy as np
import pytest

import oracle_lib as O

REF = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   "oracle", "_ref", "libref.so")
pytestmark = pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref not built")

_dp = C.POINTER(C.c_double)


def ref():
    L = C.CDLL(REF)
    L.ref_nbody_rhs.argtypes = [C.c_long, _dp, _dp, C.c_double]
    L.ref_integrate_nbody64.argtypes = [C.c_int, _dp, C.c_double, C.c_double, C.c_double,
                                        C.c_double, _dp, _dp, C.c_long]
    L.ref_integrate_nbody64.restype = C.c_long
    L.ref_leapfrog_nbody64.argtypes = [_dp, C.c_double, C.c_double, C.c_double, C.c_double,
                                       _dp, _dp, C.c_long]
    L.ref_leapfrog_nbody64.restype = C.c_long
    return L


def p(a):
    return a.ctypes.data_as(_dp)


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def cube(n, seed):
    rng = np.random.default_rng(seed)
    x = np.concatenate([rng.uniform(-1, 1, 3 * n), rng.normal(0, 0.1, 3 * n)])
    return x


@pytest.mark.parametrize("n,eps", [(64, 1e-3), (64, 0.0), (1024, 1e-2)])
def test_rhs_matches_reference(n, eps):
    x = cube(n, n)
    d = np.zeros(6 * n)
    assert ref().ref_nbody_rhs(n, p(x), p(d), eps) == 0
    assert np.array_equal(bits(d), bits(O.rhs(O.RHS_NBODY, x, eps)))


@pytest.mark.parametrize("tableau,code", [("DOPRI5", 0), ("DOP853", 1)])
def test_solve_matches_reference(tableau, code):
    x = cube(64, 3)
    xo = np.zeros(64 * 6)
    times = np.zeros(100000)
    n = ref().ref_integrate_nbody64(code, p(x), 0.0, 0.02, 1e-10, 1e-3, p(xo), p(times),
                                    times.size)
    r = O.integrate(tableau, O.RHS_NBODY, x, 0.0, 0.02, 1e-10, 1e-10, softening=1e-3,
                    store_states=False)
    assert n == len(r.times) and r.accepted > 300 and r.rejected > 10
    assert np.array_equal(bits(times[:n]), bits(r.times))
    assert np.array_equal(bits(xo), bits(r.x_final))


def test_leapfrog_matches_reference():
    x = cube(64, 5)
    xo = np.zeros(64 * 6)
    times = np.zeros(10000)
    n = ref().ref_leapfrog_nbody64(p(x), 0.0, 0.05, 1e-3, 1e-3, p(xo), p(times), times.size)
    r = O.leapfrog(x, 0.0, 0.05, 1e-3, softening=1e-3, store_states=False)
    assert n == len(r.times) == r.accepted + 1
    assert np.array_equal(bits(times[:n]), bits(r.times))
    assert np.array_equal(bits(xo), bits(r.x_final))
