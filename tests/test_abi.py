"""The C ABI library on the CPU: it loads, exports every declared symbol, and
its host-only parts (initial-condition generators) behave.  No device calls."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "odehip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(odehip_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from ode_prototyping_amd import _lib
    L = _lib.lib()
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the Python binding declares exactly these
    assert sorted(_lib.SIGNATURES) == names


def test_version_string():
    from ode_prototyping_amd import _lib
    assert b"gfx950" in _lib.lib().odehip_version()


def test_uniform_cube_properties():
    from ode_prototyping_amd import scenarios
    x, m = scenarios.uniform_cube(64, seed=0)
    n = 64
    assert x.shape == (6 * n,) and m.shape == (n,)
    pos, vel = x[:3 * n].reshape(n, 3), x[3 * n:].reshape(n, 3)
    assert np.all(np.abs(pos) <= 1.0)
    assert np.allclose(m, 1.0 / n)
    p = (m[:, None] * vel).sum(0)
    assert np.all(np.abs(p) < 1e-16)
    assert 0.05 < vel.std() < 0.15
    x2, _ = scenarios.uniform_cube(64, seed=0)
    assert np.array_equal(x, x2)
    x3, _ = scenarios.uniform_cube(64, seed=1)
    assert not np.array_equal(x, x3)


def test_plummer_virial_equilibrium():
    from ode_prototyping_amd import scenarios
    n = 4096
    x, m = scenarios.plummer(n, seed=0)
    pos, vel = x[:3 * n].reshape(n, 3), x[3 * n:].reshape(n, 3)
    assert abs(m.sum() - 1.0) < 1e-12
    assert np.all(np.abs((m[:, None] * pos).sum(0)) < 1e-12)
    assert np.all(np.abs((m[:, None] * vel).sum(0)) < 1e-12)
    T = 0.5 * (m * (vel ** 2).sum(1)).sum()
    # W = -sum_{i<j} m_i m_j / r_ij (unsoftened), computed in blocks
    W = 0.0
    for a in range(0, n, 512):
        d = pos[a:a + 512, None, :] - pos[None, :, :]
        r = np.sqrt((d ** 2).sum(-1))
        idx = np.arange(a, min(a + 512, n))
        r[np.arange(len(idx)), idx] = np.inf
        W -= 0.5 * (m[idx, None] * m[None, :] / r).sum()
    assert 0.9 < 2 * T / abs(W) < 1.1
    # the Plummer model's W = -3 pi M^2 / (32 a)
    assert abs(W / (-3 * np.pi / 32) - 1) < 0.05
    # half-mass radius of a Plummer sphere: a / sqrt(2^(2/3) - 1) = 1.3048 a
    r = np.sqrt((pos ** 2).sum(1))
    assert abs(np.median(r) / 1.3048 - 1) < 0.05


def test_colliding_plummer_layout():
    from ode_prototyping_amd import scenarios
    n = 2048
    x, m = scenarios.colliding_plummer(n, seed=0)
    pos, vel = x[:3 * n].reshape(n, 3), x[3 * n:].reshape(n, 3)
    h = n // 2
    assert abs(pos[:h, 0].mean() - 2.0) < 1e-12 and abs(pos[h:, 0].mean() + 2.0) < 1e-12
    assert abs(vel[:h, 0].mean() + 0.15) < 1e-12 and abs(vel[h:, 0].mean() - 0.15) < 1e-12
    assert abs(m.sum() - 1.0) < 1e-12


def test_golden_ic_fixture_regenerates():
    """The committed C1 initial condition is what the generator produces."""
    from ode_prototyping_amd import scenarios
    f = np.load(os.path.join(ROOT, "tests", "golden", "c1_uniform_cube_n64_seed0.npz"))
    x, m = scenarios.uniform_cube(64, seed=0)
    assert np.array_equal(f["x0"], x) and np.array_equal(f["masses"], m)


def test_log_uniform_sizes():
    from ode_prototyping_amd import scenarios
    s = scenarios.log_uniform_sizes(2048, 16, 4096, seed=0)
    assert s.min() >= 16 and s.max() <= 4096
    assert 100 < np.exp(np.log(s).mean()) < 400  # geometric mean ~ 256
