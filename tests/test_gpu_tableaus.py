"""Tableaux in the reference's struct format crossing the C ABI
(odehip_tableau_register), against the oracle.

The reference takes its tableau as the integrator's template argument
(RKEmbeddedParallel<ButcherTableau>, RKEmbeddedParallel.hpp:11-24; the struct of
butcher_tableaus/BTDOPRI5.hpp:5-26).  Here a user's struct is registered at run time
and its id used by every entry point: device solve, host-driven loop and ensembles,
each bit for bit against the oracle integrating the same coefficients.
"""
import numpy as np
import pytest

import oracle_lib as O
from tableaus_extra import BS23, CASH_KARP

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def user(t, name):
    import ode_prototyping_amd as oh
    O.register_tableau(name, t["a"], t["b"], t["bt"], t["p"], t["pt"])
    return oh.ButcherTableau(t["a"], t["b"], t["bt"], t["p"], t["pt"], name=name)


def test_reference_rkf45_struct_is_the_builtin_method(exe):
    """RKF45's coefficients passed as a struct (here the oracle's restatement of
    BTRKF45.hpp) register as the built-in RKF45 and solve bit for bit as the oracle."""
    import ode_prototyping_amd as oh
    T = O.tableau("RKF45")
    rows = [list(T["a"][r, :r + 1]) for r in range(T["S"] - 1)]
    tab = oh.ButcherTableau(rows, T["b"], T["bt"], T["p"], T["pt"], name="RKF45 struct")
    assert tab.id(exe) == oh.TABLEAUS["RKF45"]
    x0, m = oh.scenarios.uniform_cube(70, seed=2)
    res = oh.integrate(tab, x0, 0.0, 10.0, 1e-10, 1e-10, oh.NBodyODE(1e-3, m, "exact"),
                       exe=exe, max_accepted=60, trace_cap=256)
    ref = O.integrate("RKF45", O.RHS_NBODY, x0, 0.0, 10.0, 1e-10, 1e-10, softening=1e-3,
                      masses=m, reduce=O.HIP_ORDER, pow_mode=O.POW_CORRECT, max_accepted=60,
                      store_states=False, store_trace=True, threads=8)
    assert (res.accepted, res.rejected, res.rhs_evals) == (ref.accepted, ref.rejected,
                                                           ref.rhs_evals)
    assert np.array_equal(bits(res.x_final), bits(ref.x_final))


@pytest.mark.parametrize("t,name,S", [(BS23, "BS23", 4), (CASH_KARP, "CASHKARP", 6)])
def test_user_tableau_device_solve_bitwise(exe, t, name, S):
    import ode_prototyping_amd as oh
    tab = user(t, name)
    tid = tab.id(exe)
    assert tid >= len(oh.TABLEAUS)
    assert tab.id(exe) == tid and user(t, name).id(exe) == tid  # same coefficients, same id
    x0, m = oh.scenarios.uniform_cube(64, seed=4)
    res = oh.integrate(tab, x0, 0.0, 10.0, 1e-10, 1e-10, oh.NBodyODE(1e-3, m, "exact"),
                       exe=exe, max_accepted=80, trace_cap=512)
    ref = O.integrate(name, O.RHS_NBODY, x0, 0.0, 10.0, 1e-10, 1e-10, softening=1e-3,
                      masses=m, reduce=O.HIP_ORDER, pow_mode=O.POW_CORRECT, max_accepted=80,
                      store_states=False, store_trace=True, threads=8)
    assert res.accepted == 80
    assert (res.accepted, res.rejected, res.rhs_evals) == (ref.accepted, ref.rejected,
                                                           ref.rhs_evals)
    assert res.rhs_evals == 2 + S * (res.accepted + res.rejected)
    assert np.array_equal(bits(res.times), bits(ref.times))
    assert np.array_equal(bits(res.x_final), bits(ref.x_final))
    assert np.array_equal(bits(res.trace[:, 0]), bits(ref.trace_err))
    assert np.array_equal(bits(res.trace[:, 1]), bits(ref.trace_dt))


def test_user_tableau_host_driven_bitwise(exe):
    """The host-driven loop (one call per reference kernel, host pow) with a
    registered tableau: every accepted state bit for bit as the oracle in the same
    execution (canonical reduction, the C library's pow)."""
    import ode_prototyping_amd as oh
    tab = user(CASH_KARP, "CASHKARP")
    x0, m = oh.scenarios.uniform_cube(48, seed=9)
    out = oh.RawOutput()
    acc, rej, rhs = oh.integrate_host_driven(tab, x0, 0.0, 10.0, 1e-10, 1e-10,
                                             oh.NBodyODE(1e-3, m, "exact"), out, exe,
                                             max_accepted=20)
    ref = O.integrate("CASHKARP", O.RHS_NBODY, x0, 0.0, 10.0, 1e-10, 1e-10, softening=1e-3,
                      masses=m, reduce=O.HIP_ORDER, pow_mode=O.POW_LIBM, max_accepted=20)
    assert (acc, rej, rhs) == (ref.accepted, ref.rejected, ref.rhs_evals)
    assert np.array_equal(bits(out.times()), bits(ref.times))
    for k in range(len(ref.times)):
        assert np.array_equal(bits(out.states()[k]), bits(ref.state(k)))


def test_user_tableau_ensemble_and_fsal(exe):
    """An ensemble with a registered FSAL tableau (BS23): every system bit for bit as
    its oracle solve, and with k1 reuse the same steps for fewer RHS evaluations."""
    import ode_prototyping_amd as oh
    from ode_prototyping_amd.ensemble import pack, solve_ensemble
    tab = user(BS23, "BS23")
    sizes = [16, 77, 130]
    xs, ms = [], []
    for k, n in enumerate(sizes):
        x, m = oh.scenarios.uniform_cube(n, seed=30 + k)
        xs.append(x)
        ms.append(m)
    res = solve_ensemble(tab, sizes, pack(xs), 0.0, 0.01, 1e-10, 1e-10, 1e-3, pack(ms),
                         exe=exe, trace_cap=4096)
    fs = solve_ensemble(tab, sizes, pack(xs), 0.0, 0.01, 1e-10, 1e-10, 1e-3, pack(ms),
                        exe=exe, fsal=True)
    finals = res.states()
    for k in range(len(sizes)):
        ref = O.integrate("BS23", O.RHS_NBODY, xs[k], 0.0, 0.01, 1e-10, 1e-10, softening=1e-3,
                          masses=ms[k], reduce=O.HIP_ORDER, pow_mode=O.POW_CORRECT,
                          store_states=False, threads=8)
        assert (res.accepted[k], res.rejected[k], res.rhs_evals[k]) == (
            ref.accepted, ref.rejected, ref.rhs_evals)
        assert np.array_equal(bits(finals[k]), bits(ref.x_final)), k
        assert np.array_equal(bits(fs.states()[k]), bits(ref.x_final)), k
        assert fs.rhs_evals[k] < res.rhs_evals[k]


def test_tableau_registration_rejects_bad_input(exe):
    import ode_prototyping_amd as oh
    with pytest.raises(oh.OdeHipError):
        oh.ButcherTableau([[0.5]], [0.5, float("nan")], [1.0, 0.0], 2, 1).id(exe)
    with pytest.raises(oh.OdeHipError):
        oh.ButcherTableau([[0.5]], [0.5, 0.5], [1.0, 0.0], 0, 1).id(exe)
    with pytest.raises(ValueError):
        oh.ButcherTableau([[0.5], [0.1, 0.2]], [0.5, 0.5], [1.0, 0.0], 2, 1)
    # dense output stays with the built-in DOPRI5 / DOP853
    tab = user(BS23, "BS23")
    x0, m = oh.scenarios.uniform_cube(16, seed=1)
    with pytest.raises(oh.OdeHipError):
        oh.integrate(tab, x0, 0.0, 0.01, 1e-10, 1e-10, oh.NBodyODE(1e-3, m, "exact"), exe=exe,
                     dense_times=[0.005])
