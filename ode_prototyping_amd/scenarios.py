"""Synthetic N-body systems (host generators in libodehip, see include/odehip.h).

Each returns ``(x0, masses)`` with ``x0 = [pos xyz (3N) | vel xyz (3N)]``, the
reference's state layout (src/ode/NBodyODE.hpp:12-19).
"""
import numpy as np

from ._lib import dptr, lib


def uniform_cube(n, seed=0, half_width=1.0, vel_sigma=0.1, total_mass=1.0):
    """Positions U[-w, w]^3, velocities N(0, sigma^2), masses M/N, momentum removed."""
    x = np.zeros(6 * n)
    m = np.zeros(n)
    if lib().odehip_ic_uniform_cube(n, seed, half_width, vel_sigma, total_mass, dptr(x),
                                    dptr(m)) != 0:
        raise ValueError("bad uniform_cube arguments")
    return x, m


def plummer(n, seed=0, scale_radius=1.0, total_mass=1.0):
    """Plummer sphere (Aarseth, Henon & Wielen 1974), G = 1, virial equilibrium."""
    x = np.zeros(6 * n)
    m = np.zeros(n)
    if lib().odehip_ic_plummer(n, seed, scale_radius, total_mass, dptr(x), dptr(m)) != 0:
        raise ValueError("bad plummer arguments")
    return x, m


def colliding_plummer(n, seed=0, scale_radius=1.0, total_mass=1.0, separation=4.0,
                      rel_speed=0.3):
    """Two Plummer spheres of n/2 bodies at (+-separation/2, 0, 0), closing at rel_speed."""
    x = np.zeros(6 * n)
    m = np.zeros(n)
    if lib().odehip_ic_colliding_plummer(n, seed, scale_radius, total_mass, separation,
                                         rel_speed, dptr(x), dptr(m)) != 0:
        raise ValueError("bad colliding_plummer arguments")
    return x, m


def log_uniform_sizes(count, n_min, n_max, seed=0):
    """Particle counts N_k = round(exp(U(ln n_min, ln n_max))) (ensemble sizes)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    u = rng.uniform(np.log(n_min), np.log(n_max), size=count)
    return np.clip(np.rint(np.exp(u)).astype(np.int64), n_min, n_max)
