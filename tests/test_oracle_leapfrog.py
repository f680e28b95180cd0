"""Pins for the oracle's leapfrog beyond the reference build (CPU): time
reversibility, second-order convergence on a Kepler orbit, bounded energy
error (symplectic), and masses."""
import numpy as np

import oracle_lib as O


def kepler(e=0.3):
    """Two bodies, masses 1 and 1e-3, relative orbit with eccentricity e at
    pericentre distance 1 - e (semi-major axis 1)."""
    m = np.array([1.0, 1e-3])
    mu = m.sum()
    r = 1.0 - e
    v = np.sqrt(mu * (1 + e) / r)
    x = np.zeros(12)
    x[3] = r
    x[6 + 4] = v
    return x, m


def test_time_reversible():
    x, m = kepler()
    n = 200
    f = O.leapfrog(x, 0.0, n * 0.01 - 1e-9, 0.01, masses=m, store_states=False)
    assert f.accepted == n
    back = f.x_final.copy()
    back[6:] *= -1.0
    b = O.leapfrog(back, 0.0, n * 0.01 - 1e-9, 0.01, masses=m, store_states=False)
    got = b.x_final.copy()
    got[6:] *= -1.0
    assert np.max(np.abs(got - x)) < 1e-11


def test_second_order_convergence():
    x, m = kepler()
    T = 2 * np.pi  # one period (a = 1, mu ~ 1.001)
    ref = O.integrate("DOP853", O.RHS_NBODY, x, 0.0, T, 1e-13, 1e-13, masses=m,
                      store_states=False).x_final
    errs = []
    for dt in (T / 2000, T / 4000):
        n = int(round(T / dt))
        r = O.leapfrog(x, 0.0, T - 0.5 * dt, dt, masses=m, store_states=False)
        assert r.accepted == n
        errs.append(np.max(np.abs(r.x_final[:6] - ref[:6])))
    ratio = errs[0] / errs[1]
    assert 3.5 < ratio < 4.5, ratio


def test_energy_error_bounded():
    x, m = kepler(0.5)
    r = O.leapfrog(x, 0.0, 20 * 2 * np.pi, 2 * np.pi / 1000, masses=m)
    e = [sum(O.energy(r.state(k), 2, 0.0, m)) for k in range(0, len(r.times), 250)]
    e0 = e[0]
    drift = np.abs((np.array(e) - e0) / e0)
    # symplectic: the error oscillates without secular growth
    assert drift.max() < 2e-4
    assert drift[-len(drift) // 4:].max() < 2.0 * drift[:len(drift) // 4].max() + 1e-12


def test_masses_scale_bitwise():
    x, m = kepler()
    a = O.leapfrog(x, 0.0, 1.0, 0.01, masses=m, store_states=False)
    b = O.leapfrog(x, 0.0, 1.0, 0.01, masses=m, store_states=False, threads=4)
    assert np.array_equal(a.x_final, b.x_final)
