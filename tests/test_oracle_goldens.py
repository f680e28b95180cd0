"""The CPU oracle against the reference's own golden values (CPU only).

The oracle (oracle/) restates the reference's integration path; these tests
pin it to every consistency golden of the reference's test suites
(src/test/SingleThreadedRKEmbeddedTest.cpp, ThreadPoolRKEmbeddedTest.cpp,
NumericalRecipes/test/*), loaded from tests/golden/reference_goldens.json.
The reference asserts with EXPECT_DOUBLE_EQ (within 4 ulp); the oracle is
held to bit equality for the single-threaded goldens.
"""
import json
import os

import numpy as np
import pytest

import oracle_lib as O

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                   "reference_goldens.json")))
ST = GOLD["single_threaded"]


def ulps(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.max(np.abs(a - b) / np.spacing(np.maximum(np.abs(a), np.abs(b))))


def run_case(name, pow_mode=O.POW_LIBM, reduce=(O.REDUCE_SEQUENTIAL, 0)):
    tab, prob = name.split("_", 1)
    if prob == "VanDerPol":
        return O.integrate(tab, O.RHS_VANDERPOL, [2.0, 0.0], 0.0, 2.0, 1e-10, 1e-10,
                           pow_mode=pow_mode, reduce=reduce)
    if prob == "Exponential":
        return O.integrate(tab, O.RHS_EXPONENTIAL, np.arange(10.0), 0.0, 10.0, 1e-6, 1e-6,
                           pow_mode=pow_mode, reduce=reduce)
    if prob == "NBody":
        return O.integrate(tab, O.RHS_NBODY, GOLD["nbody5_x0"]["x0"], 0.0, 6.3, 1e-10, 1e-10,
                           pow_mode=pow_mode, reduce=reduce)
    if prob == "SpinningCube1024":
        x0, soft, tf, tol = O.spinning_cube_ic(1024)
        return O.integrate(tab, O.RHS_NBODY, x0, 0.0, tf, tol, tol, softening=soft,
                           pow_mode=pow_mode, reduce=reduce, threads=8)
    raise KeyError(name)


def check_golden(r, g, name, exact=True):
    cmp = (lambda a, b: a == b) if exact else (lambda a, b: ulps(a, b) <= 4)
    assert len(r.times) == g["count"], name
    assert r.times[0] == 0.0
    mid = g["count"] // 2
    assert cmp(r.times[mid], g["t_mid"]), (name, r.times[mid], g["t_mid"])
    s_mid, s_last = r.state(mid), r.state(-1)
    if name.endswith("VanDerPol"):
        assert r.times[-1] == 2.0
        assert cmp(s_mid[0], g["mid"][0]) and cmp(s_mid[1], g["mid"][1]), name
        assert cmp(s_last[0], g["last"][0]) and cmp(s_last[1], g["last"][1]), name
    elif name.endswith("Exponential"):
        assert r.times[-1] == 10.0
        idx = [0, 5, 9]
        for k, i in enumerate(idx):
            assert cmp(s_mid[i], g["mid"][k]), (name, i)
            assert cmp(s_last[i], g["last"][k]), (name, i)
    elif name.endswith("NBody"):
        assert r.times[-1] == 6.3
        for k, i in enumerate(g["first_idx"]):
            assert r.state(0)[i] == g["first"][k]
            assert cmp(s_mid[i], g["mid"][k]), (name, i)
            assert cmp(s_last[i], g["last"][k]), (name, i)
    else:
        assert r.times[-1] == g["t_last"]
        for k, i in enumerate(g["idx"]):
            assert r.state(0)[i] == g["first"][k]
            assert cmp(s_mid[i], g["mid"][k]), (name, i)
            assert cmp(s_last[i], g["last"][k]), (name, i)


@pytest.mark.parametrize("name", [k for k in ST if k != "RKF78_SpinningCube1024"])
def test_single_threaded_goldens_bit_exact(name):
    """SingleThreadedRKEmbeddedTest consistency goldens, bit for bit."""
    check_golden(run_case(name), ST[name], name, exact=True)


def test_spinning_cube_1024_golden_bit_exact():
    """RKF78 on the 1024-body spinning cube (SingleThreadedRKEmbeddedTest.cpp:342-382)."""
    name = "RKF78_SpinningCube1024"
    check_golden(run_case(name), ST[name], name, exact=True)


def test_thread_pool_exponential_golden():
    """The 8-thread pool's chunked reduction order has its own golden
    (ThreadPoolRKEmbeddedTest.cpp:163-193); the oracle's ThreadPool order
    reproduces it exactly, and differs from the single-threaded one."""
    g = GOLD["thread_pool_8"]["DOPRI5_Exponential"]
    r = run_case("DOPRI5_Exponential", reduce=(O.REDUCE_THREADPOOL, 8))
    assert len(r.times) == g["count"]
    mid = g["count"] // 2
    assert r.times[mid] == g["t_mid"]
    assert r.state(mid)[5] == g["mid"][1] and r.state(mid)[9] == g["mid"][2]
    assert r.state(-1)[5] == g["last"][1] and r.state(-1)[9] == g["last"][2]
    assert g["t_mid"] != ST["DOPRI5_Exponential"]["t_mid"]


@pytest.mark.parametrize("name", ["DOPRI5_VanDerPol", "RKF78_NBody", "DVERK_Exponential",
                                  "RKF45_VanDerPol"])
def test_correctly_rounded_pow_mode_matches_goldens(name):
    """With a correctly rounded pow (the device controller's) the goldens
    still hold to EXPECT_DOUBLE_EQ precision (glibc's pow is not correctly
    rounded in ~0.1% of cases, so bit equality is not guaranteed)."""
    check_golden(run_case(name, pow_mode=O.POW_CORRECT), ST[name], name, exact=False)


@pytest.mark.parametrize("problem", ["oscillator", "vanderpol"])
def test_nr_dop853_goldens(problem):
    """The DOP853 constants, through the NR stepper's own semantics, reproduce
    NumericalRecipes/test/StepperDopr853_test.cpp bit for bit."""
    g = GOLD["nr_dop853"][problem]
    st, cnt, sel = O.nr_dop853(0 if problem == "oscillator" else 1, g["y0"], 0.0, 2.0,
                               g["tol"], g["tol"], g["h1"], 0.0, g.get("mu_eps", 1e-3))
    assert st == 0
    assert cnt == g["count"]
    assert sel[0] == 0.0 and sel[1] == g["y0"][0] and sel[2] == g["y0"][1]
    assert sel[3] == g["x_mid"]
    assert sel[4] == g["y_mid"][0] and sel[5] == g["y_mid"][1]
    assert sel[6] == 2.0
    assert sel[7] == g["y_last"][0] and sel[8] == g["y_last"][1]


def test_inner_product_reduction_orders():
    """transform_reduce of iota*(iota+1), n=10 = 330 (CudaExecutorTest.cu:52-61)
    in every reduction order the oracle knows."""
    g = GOLD["cuda_inner_product"]
    a = np.arange(10.0)
    b = np.arange(1.0, 11.0)
    # rk_norm computes sqrt(sum((v/sc)^2)/n); choose v/sc = sqrt(a*b).
    v = np.sqrt(a * b)
    for red in [(O.REDUCE_SEQUENTIAL, 0), (O.REDUCE_THREADPOOL, 8), (O.REDUCE_CHUNKED, 3),
                O.HIP_ORDER]:
        val = O.rk_norm(v, np.ones(10), red) ** 2 * g["n"]
        assert abs(val - g["value"]) < 1e-10
