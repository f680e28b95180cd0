"""Pins for the oracle functions the reference's goldens do not cover (CPU).

Each check is chosen so that a plausible mistake (a dropped mass factor, a
sign or index error, a transposed tableau entry) fails it:
* per-particle masses reduce bit for bit to the reference's unit-mass RHS;
* a two-body system matches the closed form a = m r / (r^2 + eps^2)^(3/2)
  with equal and opposite forces (times masses);
* total momentum derivative sum_i m_i a_i vanishes (pairwise antisymmetry);
* the DOP853 tableau satisfies the order conditions of its 8th- and
  5th-order weights and consistent row sums;
* the energy of a two-body system matches the closed form;
* the correctly rounded pow mode agrees with glibc except in rare cases.
"""
import itertools
from fractions import Fraction

import numpy as np
import pytest

import oracle_lib as O


def test_unit_masses_reduce_bitwise_to_reference():
    rng = np.random.default_rng(3)
    n = 97
    x = np.concatenate([rng.uniform(-1, 1, 3 * n), rng.normal(size=3 * n)])
    a = O.rhs(O.RHS_NBODY, x, 1e-3, None)
    b = O.rhs(O.RHS_NBODY, x, 1e-3, np.ones(n))
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.mark.parametrize("eps", [0.0, 0.3])
def test_two_body_closed_form(eps):
    m = np.array([2.5, 0.75])
    r = np.array([0.3, -0.4, 1.2])
    x = np.concatenate([np.zeros(3), r, [0.1, 0.2, 0.3, -0.4, -0.5, -0.6]])
    d = O.rhs(O.RHS_NBODY, x, eps, m)
    dist = np.sqrt(r @ r + eps * eps)
    a0 = m[1] * r / dist ** 3
    a1 = -m[0] * r / dist ** 3
    assert np.allclose(d[6:9], a0, rtol=1e-14, atol=0)
    assert np.allclose(d[9:12], a1, rtol=1e-14, atol=0)
    assert np.array_equal(d[:6], x[6:])  # velocity half copied
    # equal and opposite forces
    assert np.allclose(m[0] * d[6:9], -m[1] * d[9:12], rtol=1e-15)


def test_mass_scaling_law():
    """Scaling every mass by c scales the acceleration by c (to rounding)."""
    rng = np.random.default_rng(5)
    n = 50
    x = np.concatenate([rng.uniform(-1, 1, 3 * n), rng.normal(size=3 * n)])
    m = rng.uniform(0.5, 2.0, n)
    a = O.rhs(O.RHS_NBODY, x, 1e-2, m)[3 * n:]
    b = O.rhs(O.RHS_NBODY, x, 1e-2, 4.0 * m)[3 * n:]  # power of two: exact
    assert np.array_equal(4.0 * a, b)


def test_momentum_derivative_vanishes():
    rng = np.random.default_rng(7)
    n = 300
    x = np.concatenate([rng.uniform(-1, 1, 3 * n), rng.normal(size=3 * n)])
    m = rng.uniform(0.1, 1.0, n)
    acc = O.rhs(O.RHS_NBODY, x, 1e-2, m)[3 * n:].reshape(n, 3)
    f = (m[:, None] * acc).sum(0)
    assert np.all(np.abs(f) < 1e-11 * np.abs(m[:, None] * acc).sum(0))


def _rational(v):
    return Fraction(v).limit_denominator(10 ** 40)


def test_dop853_order_conditions():
    """Quadrature conditions sum b_i c_i^(k-1) = 1/k for k <= 8 (b) and
    k <= 5 (bt), row sums of a equal c, and b sums to 1."""
    T = O.tableau("DOP853")
    S, a, b, bt = T["S"], T["a"], T["b"], T["bt"]
    assert (T["p"], T["pt"], S) == (8, 5, 12)
    c = np.zeros(S)
    for s in range(1, S):
        c[s] = a[s - 1, :s].sum()
    published_c = [0.0, 0.0526001519587677318785587544488, 0.0789002279381515978178381316732,
                   0.118350341907227396726757197510, 0.281649658092772603273242802490,
                   1.0 / 3.0, 0.25, 0.307692307692307692307692307692,
                   0.651282051282051282051282051282, 0.6, 0.857142857142857142857142857142, 1.0]
    assert np.allclose(c, published_c, atol=1e-14)
    for k in range(1, 9):
        assert abs((b * c ** (k - 1)).sum() - 1.0 / k) < 1e-14, k
    for k in range(1, 6):
        assert abs((bt * c ** (k - 1)).sum() - 1.0 / k) < 1e-13, k
    # a second-level condition: sum b_i a_ij c_j = 1/6
    assert abs(b @ (np.vstack([np.zeros(S - 1), a]) @ c[:-1]) - 1.0 / 6.0) < 1e-13


@pytest.mark.parametrize("name,p", [("DOPRI5", 5), ("RKF78", 7), ("DVERK", 6), ("RKF45", 4),
                                    ("HE21", 2)])
def test_reference_tableaux_quadrature(name, p):
    T = O.tableau(name)
    S, a, b, bt = T["S"], T["a"], T["b"], T["bt"]
    c = np.array([0.0] + [a[s - 1, :s].sum() for s in range(1, S)])
    for k in range(1, p + 1):
        assert abs((b * c ** (k - 1)).sum() - 1.0 / k) < 1e-13, (name, k)
    for k in range(1, T["pt"] + 1):
        assert abs((bt * c ** (k - 1)).sum() - 1.0 / k) < 1e-13, (name, k)


def test_two_body_energy_closed_form():
    m = np.array([1.5, 0.5])
    r = np.array([1.0, 2.0, -2.0])
    v0, v1 = np.array([0.1, 0.0, -0.2]), np.array([0.0, 0.3, 0.4])
    x = np.concatenate([np.zeros(3), r, v0, v1])
    eps = 0.1
    ke, pe = O.energy(x, 2, eps, m)
    assert abs(ke - 0.5 * (m[0] * v0 @ v0 + m[1] * v1 @ v1)) < 1e-15
    assert abs(pe + m[0] * m[1] / np.sqrt(r @ r + eps * eps)) < 1e-15
    p = O.momentum(x, 2, m)
    assert np.allclose(p, m[0] * v0 + m[1] * v1, rtol=0, atol=1e-16)


def test_two_body_energy_conserved_by_integration():
    """A softened two-body orbit integrated with DOP853 at tol 1e-12 keeps its
    energy to ~1e-11 and its momentum to rounding."""
    m = np.array([1.0, 1e-3])
    x = np.array([0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 1.0, 0])
    r = O.integrate("DOP853", O.RHS_NBODY, x, 0.0, 6.0, 1e-12, 1e-12, softening=0.0,
                    masses=m, store_states=False)
    e0 = sum(O.energy(x, 2, 0.0, m))
    e1 = sum(O.energy(r.x_final, 2, 0.0, m))
    assert abs((e1 - e0) / e0) < 1e-10
    assert np.allclose(O.momentum(r.x_final, 2, m), O.momentum(x, 2, m), atol=1e-15)


def test_pow_modes():
    rng = np.random.default_rng(0)
    xs = np.exp(rng.uniform(-30, 5, 20000))
    ys = [1 / 6, 1 / 5, 1 / 9, 1 / 3]
    diff = 0
    for x, y in zip(xs, itertools.cycle(ys)):
        a, b = O.pow_(x, y, O.POW_LIBM), O.pow_(x, y, O.POW_CORRECT)
        assert abs(a - b) <= np.spacing(b)
        diff += a != b
    assert diff < 60  # glibc is correctly rounded ~99.9 % of the time


def test_pi_controller_follows_its_formula():
    """The PI controller (no reference counterpart): every step size of a run
    follows dt' = dt * (0.38/err)^(0.3/k) * (err_prev/err)^(0.4/k) (k = 6 for
    DOP853; no growth after a rejection; clamps [0.33, 6]) with err_prev the
    error of the last accepted attempt, recomputed here from the run's own
    trace with numpy's pow.  A wrong history (err_prev from a rejected
    attempt), swapped exponents or a missing post-rejection cap fail it."""
    x = np.array([2.0, 0.0])
    r = O.integrate("DOP853", O.RHS_VANDERPOL, x, 0.0, 10.0, 1e-8, 1e-8,
                    store_states=False, store_trace=True, controller=O.CONTROLLER_PI)
    err, dt, acc = np.array(r.trace_err), np.array(r.trace_dt), np.array(r.trace_acc)
    assert acc.sum() == r.accepted and (1 - acc).sum() == r.rejected > 0
    k = 6.0
    err_prev = 1e-4
    t = 0.0
    for a in range(len(err) - 1):
        fac = (0.38 / err[a]) ** (0.3 / k) * (err_prev / err[a]) ** (0.4 / k)
        if not acc[a]:
            fac = min(fac, 1.0)
        want = dt[a] * min(max(fac, 0.33), 6.0)
        if acc[a]:
            t += dt[a]
            err_prev = max(err[a], 1e-4)
        if t + want > 10.0:
            want = 10.0 - t
        assert abs(dt[a + 1] - want) <= 1e-13 * want, a


def test_pi_controller_keeps_accuracy():
    """Only the step sizes change: on x' = x (closed form e^t) the PI run's
    final error stays at the tolerance level like the reference controller's,
    and every accepted attempt had an error norm <= 1."""
    x = np.array([1.0, 0.5, 2.0])
    for name in ("DOPRI5", "DOP853", "RKF78"):
        r = O.integrate(name, O.RHS_EXPONENTIAL, x, 0.0, 5.0, 1e-10, 1e-10,
                        store_states=False, store_trace=True, controller=O.CONTROLLER_PI)
        exact = x * np.exp(5.0)
        assert np.max(np.abs(r.x_final - exact) / exact) < 1e-8, name
        err, acc = np.array(r.trace_err), np.array(r.trace_acc)
        assert np.all(err[acc == 1] <= 1.0) and np.all(err[acc == 0] > 1.0)
