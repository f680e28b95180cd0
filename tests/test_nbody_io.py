"""Trajectory files (CPU): the writer's byte layout is the reference's
output_to_file one (u64 n_times, u64 n_var, f64 softening, rows of t and the
state), checked field by field with struct; read back bit for bit."""
import struct

import numpy as np
import pytest


def test_layout_and_round_trip(tmp_path):
    from ode_prototyping_amd import nbody_io
    rng = np.random.default_rng(0)
    times = np.array([0.0, 0.25, 0.5])
    states = rng.normal(size=(3, 12))
    path = tmp_path / "traj.bin"
    nbody_io.write_bin(path, times, states, softening=1e-3)
    raw = path.read_bytes()
    assert len(raw) == 24 + 3 * 13 * 8
    assert struct.unpack("<QQd", raw[:24]) == (3, 12, 1e-3)
    row1 = struct.unpack("<13d", raw[24 + 13 * 8:24 + 26 * 8])
    assert row1[0] == 0.25 and np.array_equal(np.array(row1[1:]), states[1])
    t, s, eps = nbody_io.read_bin(path)
    assert eps == 1e-3 and np.array_equal(t, times) and np.array_equal(s, states)
    assert nbody_io.read_metadata(path) == (3, 12, 1e-3)


def test_raw_output_and_errors(tmp_path):
    import ode_prototyping_amd as oh
    from ode_prototyping_amd import nbody_io
    out = oh.RawOutput()
    out.save_state(0.0, np.arange(6.0))
    out.save_state(1.0, np.arange(6.0) * 2)
    nbody_io.write_output(tmp_path / "o.bin", out, softening=0.5)
    t, s, eps = nbody_io.read_bin(tmp_path / "o.bin")
    assert list(t) == [0.0, 1.0] and s[1, 5] == 10.0 and eps == 0.5
    with pytest.raises(OSError):
        nbody_io.write_bin(tmp_path / "missing_dir" / "x.bin", t, s)
    with pytest.raises(ValueError):
        nbody_io.write_bin(tmp_path / "y.bin", t, s[:1])
