"""Trajectory files in the reference's binary format.

Writer: ``output_to_file`` (``src/demo/nbody_io.hpp:4-22``), here the C ABI's
``odehip_write_trajectory``; reader: the layout ``scripts/nbody_io/
bin_reader.py:4-18`` reads -- little-endian u64 n_times, u64 n_var, f64
softening, then n_times rows of (t, n_var state values) as f64.
"""
import struct

import numpy as np

from ._lib import dptr, lib


def write_bin(path, times, states, softening=0.0):
    """Write saved states (``RawOutput``-like: times[k], states[k]) to ``path``."""
    times = np.ascontiguousarray(times, dtype=np.float64).ravel()
    states = np.ascontiguousarray(states, dtype=np.float64)
    if states.ndim != 2 or states.shape[0] != times.size:
        raise ValueError("states must be n_times x n_var")
    rc = lib().odehip_write_trajectory(str(path).encode(), times.size, states.shape[1],
                                       float(softening), dptr(times), dptr(states))
    if rc != 0:
        raise OSError(f"could not write {path}")


def write_output(path, output, softening=0.0):
    """output_to_file(filename, output, softening) for a RawOutput."""
    write_bin(path, output.times(), np.array([np.asarray(s) for s in output.states()]),
              softening)


def read_metadata(path):
    with open(path, "rb") as f:
        n_times = int.from_bytes(f.read(8), "little")
        n_var = int.from_bytes(f.read(8), "little")
        softening = struct.unpack("<d", f.read(8))[0]
    return n_times, n_var, softening


def read_bin(path):
    """(times, states [n_times x n_var], softening)."""
    n_times, n_var, softening = read_metadata(path)
    data = np.fromfile(path, dtype="<f8", offset=24)
    if data.size != n_times * (n_var + 1):
        raise ValueError(f"{path}: {data.size} values, header implies {n_times * (n_var + 1)}")
    rows = data.reshape(n_times, n_var + 1)
    return rows[:, 0].copy(), rows[:, 1:].copy(), softening
