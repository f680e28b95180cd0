"""DOP853 / DOPRI5 dense output of the device-resident solve against the CPU
oracle.

The oracle's continuous extension is pinned by the reference's dense-output
goldens (test_oracle_dense.py); here the device pass (for DOP853 three extra
RHS evaluations per covering step, for DOPRI5 none; the polynomial evaluated
per element) must
reproduce it bit for bit on the same solve: the written rows, the steps that
wrote them and the solve itself, which dense output must not change.
"""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def oracle_dense(rhs_kind, x0, tf, tol, times, softening=0.0, masses=None, tableau="DOP853"):
    return O.integrate(tableau, rhs_kind, x0, 0.0, tf, tol, tol, softening=softening,
                       masses=masses, reduce=O.HIP_ORDER, pow_mode=O.POW_CORRECT,
                       store_states=False, threads=8, dense_times=times)


@pytest.mark.parametrize("tableau", ["DOP853", "DOPRI5"])
@pytest.mark.parametrize("n,kind,variant_eps", [(64, "cube", 1e-3), (300, "plummer", 1e-2),
                                                (1000, "plummer", 1e-2)])
def test_nbody_dense_bit_identical(exe, n, kind, variant_eps, tableau):
    import ode_prototyping_amd as oh
    from ode_prototyping_amd import scenarios
    x0, m = (scenarios.plummer(n, seed=4) if kind == "plummer"
             else scenarios.uniform_cube(n, seed=5))
    tf = 0.05 if n <= 300 else 0.02
    times = np.linspace(0.0, tf, 26)[1:-1]
    print(f"dense nbody {tableau} n={n} kind={kind}", flush=True)
    ref = oracle_dense(O.RHS_NBODY, x0, tf, 1e-10, times, variant_eps, m, tableau)
    res = oh.integrate(tableau, x0, 0.0, tf, 1e-10, 1e-10,
                       oh.NBodyODE(variant_eps, m, "exact"), exe=exe, dense_times=times)
    assert res.status == "reached_end"
    assert (res.accepted, res.rejected, res.rhs_evals) == (ref.accepted, ref.rejected,
                                                           ref.rhs_evals)
    assert np.array_equal(bits(res.x_final), bits(ref.x_final))
    assert res.n_dense == ref.n_dense == len(times)
    assert res.dense_steps == ref.dense_steps and res.dense_rhs_evals == ref.dense_rhs_evals
    assert np.array_equal(res.dense_times, times)
    assert np.array_equal(bits(res.dense_states), bits(ref.dense_states))


@pytest.mark.parametrize("tableau", ["DOP853", "DOPRI5"])
@pytest.mark.parametrize("rhs", ["vanderpol", "exponential"])
def test_simple_dense_bit_identical(exe, rhs, tableau):
    import ode_prototyping_amd as oh
    if rhs == "vanderpol":
        x0, ode, kind, tf = np.array([2.0, 0.0]), oh.VanDerPolODE(), O.RHS_VANDERPOL, 2.0
    else:
        x0, ode, kind, tf = np.arange(10.0), oh.ExponentialODE(), O.RHS_EXPONENTIAL, 3.0
    times = np.linspace(0.0, tf, 41)[1:]  # the last (tf) is never written
    ref = oracle_dense(kind, x0, tf, 1e-10, times, tableau=tableau)
    res = oh.integrate(tableau, x0, 0.0, tf, 1e-10, 1e-10, ode, exe=exe, dense_times=times)
    assert res.n_dense == ref.n_dense == len(times) - 1
    assert np.array_equal(bits(res.dense_states), bits(ref.dense_states))
    assert np.array_equal(bits(res.x_final), bits(ref.x_final))
    if rhs == "exponential":
        exact = np.exp(times[:-1])[:, None] * x0
        assert np.max(np.abs(res.dense_states - exact) / np.maximum(np.abs(exact), 1)) < 1e-8


def test_dense_output_policy(exe):
    """The DenseOutput policy: the initial state, then the reference's
    accumulated points strictly before tf, equal to explicit dense times."""
    import ode_prototyping_amd as oh
    from ode_prototyping_amd import scenarios
    x0, m = scenarios.plummer(128, seed=8)
    ode = oh.NBodyODE(1e-2, m, "exact")
    out = oh.DenseOutput(20)
    res = oh.integrate("DOP853", x0, 0.0, 0.04, 1e-10, 1e-10, ode, output=out, exe=exe)
    xs = out.x_values()
    width = 0.04 / 20
    acc = [0.0]
    t = width
    while t < 0.04:
        acc.append(t)
        t += width
    assert xs == acc and out.n_steps() == len(acc) == res.n_dense + 1
    y = out.y_values()
    assert y.shape == (x0.size, len(acc)) and np.array_equal(y[:, 0], x0)
    ref = oracle_dense(O.RHS_NBODY, x0, 0.04, 1e-10, acc[1:], 1e-2, m)
    assert np.array_equal(bits(y[:, 1:].T), bits(ref.dense_states))


def test_dense_output_does_not_change_the_solve(exe):
    """With k1 reuse the end-of-step RHS of a dense pass becomes the next
    attempt's k1: same steps, same states, one RHS fewer per dense step
    (except a dense pass on the very last step)."""
    import ode_prototyping_amd as oh
    from ode_prototyping_amd import scenarios
    x0, m = scenarios.uniform_cube(200, seed=11)
    ode = oh.NBodyODE(1e-3, m, "exact")
    times = np.linspace(0.0, 0.05, 11)[1:]
    a = oh.integrate("DOP853", x0, 0.0, 0.05, 1e-10, 1e-10, ode, exe=exe, fsal=True,
                     trace_cap=4096)
    b = oh.integrate("DOP853", x0, 0.0, 0.05, 1e-10, 1e-10, ode, exe=exe, fsal=True,
                     trace_cap=4096, dense_times=times)
    c = oh.integrate("DOP853", x0, 0.0, 0.05, 1e-10, 1e-10, ode, exe=exe, fsal=False,
                     dense_times=times)
    assert np.array_equal(bits(a.x_final), bits(b.x_final))
    assert np.array_equal(bits(a.trace), bits(b.trace))
    assert np.array_equal(bits(a.times), bits(b.times))
    assert a.rhs_evals - b.rhs_evals in (b.dense_steps, b.dense_steps - 1)
    assert np.array_equal(bits(b.dense_states), bits(c.dense_states))
    assert b.n_dense == 9 and b.dense_rhs_evals == 4 * b.dense_steps


def test_dense_output_argument_errors(exe):
    import ode_prototyping_amd as oh
    x0 = np.array([2.0, 0.0])
    ode = oh.VanDerPolODE()
    with pytest.raises(oh.core.OdeHipError, match="DOP853 or DOPRI5"):
        oh.integrate("RKF78", x0, 0.0, 1.0, 1e-8, 1e-8, ode, exe=exe, dense_times=[0.5])
    with pytest.raises(oh.core.OdeHipError, match="increasing"):
        oh.integrate("DOP853", x0, 0.0, 1.0, 1e-8, 1e-8, ode, exe=exe, dense_times=[0.5, 0.5])
    with pytest.raises(oh.core.OdeHipError, match="increasing"):
        oh.integrate("DOP853", x0, 0.0, 1.0, 1e-8, 1e-8, ode, exe=exe, dense_times=[0.0])
    # the context is still usable
    r = oh.integrate("DOP853", x0, 0.0, 1.0, 1e-8, 1e-8, ode, exe=exe, dense_times=[0.5])
    assert r.n_dense == 1


def test_dopri5_dense_costs_no_evaluation(exe):
    """DOPRI5's extension reads its seventh stage (f at the new state): the
    solve, with or without k1 reuse, is the solve without dense output, and
    no RHS evaluation is added."""
    import ode_prototyping_amd as oh
    from ode_prototyping_amd import scenarios
    x0, m = scenarios.uniform_cube(150, seed=12)
    ode = oh.NBodyODE(1e-3, m, "exact")
    times = np.linspace(0.0, 0.05, 21)[1:]
    runs = {}
    for fsal in (False, True):
        for dense in (False, True):
            runs[fsal, dense] = oh.integrate("DOPRI5", x0, 0.0, 0.05, 1e-10, 1e-10, ode, exe=exe,
                                             fsal=fsal, trace_cap=4096,
                                             dense_times=times if dense else None)
    for fsal in (False, True):
        a, b = runs[fsal, False], runs[fsal, True]
        assert np.array_equal(bits(a.x_final), bits(b.x_final))
        assert np.array_equal(bits(a.trace), bits(b.trace))
        assert a.rhs_evals == b.rhs_evals and b.dense_rhs_evals == 0 and b.n_dense == 19
    assert np.array_equal(bits(runs[False, True].dense_states), bits(runs[True, True].dense_states))


@pytest.mark.parametrize("tableau", ["DOP853", "DOPRI5"])
def test_dense_output_direct_launch_and_budget(exe, tableau):
    """The same rows with attempts launched directly instead of from the
    captured graph; and a solve stopped by an accepted-step budget writes
    exactly the rows strictly before the time it reached."""
    import ode_prototyping_amd as oh
    from ode_prototyping_amd import scenarios
    x0, m = scenarios.plummer(96, seed=21)
    ode = oh.NBodyODE(1e-2, m, "exact")
    times = np.linspace(0.0, 0.05, 31)[1:-1]
    g = oh.integrate(tableau, x0, 0.0, 0.05, 1e-10, 1e-10, ode, exe=exe, dense_times=times)
    d = oh.integrate(tableau, x0, 0.0, 0.05, 1e-10, 1e-10, ode, exe=exe, dense_times=times,
                     attempts_per_launch=-1)
    assert g.n_dense == d.n_dense == len(times)
    assert np.array_equal(bits(g.dense_states), bits(d.dense_states))
    assert np.array_equal(bits(g.x_final), bits(d.x_final))
    half = g.accepted // 2
    b = oh.integrate(tableau, x0, 0.0, 0.05, 1e-10, 1e-10, ode, exe=exe, dense_times=times,
                     max_accepted=half)
    assert b.status == "max_accepted" and b.t < 0.05
    k = int(np.sum(times < b.t))
    assert b.n_dense == k and 0 < k < len(times)
    assert np.array_equal(bits(b.dense_states), bits(g.dense_states[:k]))
