"""Registered tableaux on the CPU: the test pairs of tableaus_extra (typed in from the
publications) are pinned by the Butcher order conditions, and the oracle integrates a
registered tableau exactly as the same coefficients built in.

The order conditions below are all 17 rooted-tree conditions up to order 5, so a
mistyped coefficient, a swapped b / bt or a misplaced row fails them.
"""
import numpy as np
import pytest

import oracle_lib as O
from tableaus_extra import BS23, CASH_KARP


def _matrix(t):
    s = len(t["b"])
    A = np.zeros((s, s))
    for r, row in enumerate(t["a"]):
        A[r + 1, :r + 1] = row
    return A, np.asarray(t["b"], float), np.asarray(t["bt"], float)


def _conditions(A, w):
    """Residuals of the order conditions of weights w, by order (1..5)."""
    c = A.sum(axis=1)
    Ac, Ac2, Ac3 = A @ c, A @ c ** 2, A @ c ** 3
    AAc = A @ Ac
    return {
        1: [w.sum() - 1.0],
        2: [w @ c - 1 / 2],
        3: [w @ c ** 2 - 1 / 3, w @ Ac - 1 / 6],
        4: [w @ c ** 3 - 1 / 4, w @ (c * Ac) - 1 / 8, w @ Ac2 - 1 / 12, w @ AAc - 1 / 24],
        5: [w @ c ** 4 - 1 / 5, w @ (c ** 2 * Ac) - 1 / 10, w @ (c * Ac2) - 1 / 15,
            w @ (c * AAc) - 1 / 30, w @ Ac ** 2 - 1 / 20, w @ Ac3 - 1 / 20,
            w @ (A @ (c * Ac)) - 1 / 40, w @ (A @ Ac2) - 1 / 60, w @ (A @ AAc) - 1 / 120],
    }


@pytest.mark.parametrize("t", [BS23, CASH_KARP], ids=["BS23", "CashKarp"])
def test_extra_tableaux_order_conditions(t):
    A, b, bt = _matrix(t)
    for w, order in ((b, t["p"]), (bt, t["pt"])):
        res = _conditions(A, w)
        for k in range(1, order + 1):
            assert np.max(np.abs(res[k])) < 1e-15, (order, k, res[k])
        # and not of the next order (the weights are not swapped)
        if order + 1 in res:
            assert np.max(np.abs(res[order + 1])) > 1e-6


def test_bs23_first_same_as_last():
    A, b, _ = _matrix(BS23)
    assert np.array_equal(A[-1, :-1], b[:-1]) and b[-1] == 0.0


@pytest.mark.parametrize("name", ["RKF45", "DOPRI5"])
def test_registered_copy_of_a_builtin_integrates_identically(name):
    T = O.tableau(name)
    rows = [list(T["a"][r, :r + 1]) for r in range(T["S"] - 1)]
    O.register_tableau(name + "_COPY", rows, T["b"], T["bt"], T["p"], T["pt"])
    x0 = np.array([2.0, 0.0])
    a = O.integrate(name, O.RHS_VANDERPOL, x0, 0.0, 2.0, 1e-10, 1e-10)
    b = O.integrate(name + "_COPY", O.RHS_VANDERPOL, x0, 0.0, 2.0, 1e-10, 1e-10)
    assert (a.accepted, a.rejected) == (b.accepted, b.rejected)
    assert np.array_equal(a.x_final, b.x_final) and np.array_equal(a.times, b.times)


def test_registered_cash_karp_van_der_pol():
    """Cash-Karp through the oracle on the reference's Van der Pol case ends within
    1e-9 of the reference's DOPRI5 golden end state (SingleThreadedRKEmbeddedTest.cpp:
    138-160), as a 5(4) pair at tolerance 1e-10 should."""
    O.register_tableau("CASHKARP", CASH_KARP["a"], CASH_KARP["b"], CASH_KARP["bt"],
                       CASH_KARP["p"], CASH_KARP["pt"])
    r = O.integrate("CASHKARP", O.RHS_VANDERPOL, np.array([2.0, 0.0]), 0.0, 2.0, 1e-10, 1e-10)
    assert r.t == 2.0
    assert np.max(np.abs(r.x_final - [0.32331666704309109, -1.8329745679718159])) < 1e-9
