"""Leapfrog on the device against the oracle (whose leapfrog equals the
reference's ParallelLeapfrogIntegrator bit for bit, test_oracle_reference_build):
with the exact RHS every saved state and time is bit-identical."""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


@pytest.mark.parametrize("n,masses", [(64, False), (100, True), (300, True)])
def test_exact_leapfrog_matches_oracle(exe, n, masses):
    import ode_prototyping_amd as oh
    x, m = oh.scenarios.uniform_cube(n, seed=n)
    mm = m if masses else None
    out = oh.RawOutput()
    xf, st = oh.leapfrog(x, 0.0, 0.05, 1e-3, oh.NBodyODE(1e-3, mm, "exact"), out, exe=exe)
    ref = O.leapfrog(x, 0.0, 0.05, 1e-3, softening=1e-3, masses=mm, threads=8)
    assert st["steps"] == ref.accepted and len(out.times()) == len(ref.times)
    assert np.array_equal(bits(out.times()), bits(ref.times))
    for k in (0, 1, len(ref.times) // 2, len(ref.times) - 1):
        assert np.array_equal(bits(out.states()[k]), bits(ref.state(k))), k
    assert np.array_equal(bits(xf), bits(ref.x_final))


@pytest.mark.parametrize("variant,tol", [("fast", 1e-13), ("mixed", 1e-6)])
def test_leapfrog_variants_close(exe, variant, tol):
    import ode_prototyping_amd as oh
    x, m = oh.scenarios.plummer(500, seed=2)
    xf, _ = oh.leapfrog(x, 0.0, 0.1, 1e-3, oh.NBodyODE(1e-2, m, variant), exe=exe)
    ref = O.leapfrog(x, 0.0, 0.1, 1e-3, softening=1e-2, masses=m, store_states=False,
                     threads=8).x_final
    assert np.max(np.abs(xf - ref)) / np.max(np.abs(ref)) < tol


def test_leapfrog_energy_and_edge_cases(exe):
    import ode_prototyping_amd as oh
    x, m = oh.scenarios.plummer(1000, seed=4)
    ode = oh.NBodyODE(1e-2, m, "exact")
    xf, st = oh.leapfrog(x, 0.0, 0.5, 2e-3, ode, exe=exe)
    e0 = sum(O.energy(x, 1000, 1e-2, m))
    e1 = sum(O.energy(xf, 1000, 1e-2, m))
    assert st["steps"] == 250 or st["steps"] == 251
    assert abs((e1 - e0) / e0) < 1e-5
    # nothing to do when t0 >= tf: the initial state is the only saved one
    out = oh.RawOutput()
    xf, st = oh.leapfrog(x, 1.0, 1.0, 1e-3, ode, out, exe=exe)
    assert st["steps"] == 0 and np.array_equal(xf, x) and out.times() == [1.0]
    with pytest.raises(Exception):
        oh.leapfrog(x, 0.0, 1.0, 0.0, ode, exe=exe)
