"""Host-side logic of the Python mirror that runs without a GPU.

* DenseOutput's output points are the reference's accumulated x values
  (Output.hpp:74-81, 110-113: x1 + w, then += w), bit for bit, as listed by
  NumericalRecipes/test/StepperDopr853_test.cpp:181-196 and
  SimpleOscillatorTest.cpp:254-255 -- the points strictly before x2;
* DenseOutput rejects fewer than one interval, like set_n_intervals;
* the ensemble pack / unpack helpers round-trip and check their sizes.
"""
import numpy as np
import pytest

from ode_prototyping_amd.core import DenseOutput
from ode_prototyping_amd.ensemble import pack, unpack


def test_dense_output_points_are_the_reference_x_values():
    golden = [0.1, 0.2, 0.30000000000000004, 0.4, 0.5, 0.6, 0.7, 0.7999999999999999,
              0.8999999999999999, 0.9999999999999999, 1.0999999999999999, 1.2, 1.3,
              1.4000000000000001, 1.5000000000000002, 1.6000000000000003, 1.7000000000000004,
              1.8000000000000005, 1.9000000000000006]
    assert np.array_equal(DenseOutput(20).output_times(0.0, 2.0), golden)
    assert np.array_equal(DenseOutput().output_times(0.0, 2.0), golden)  # default 20
    # start at 1 with five intervals: the fifth point is 2 - 2^-52, before x2
    five = DenseOutput(5).output_times(1.0, 2.0)
    assert np.array_equal(five[:4], [1.2, 1.3999999999999999, 1.5999999999999999,
                                     1.7999999999999998])
    assert five.size == 5 and five[4] == np.nextafter(2.0, 0.0)


def test_dense_output_interval_count_is_checked():
    with pytest.raises(ValueError):
        DenseOutput(0)
    d = DenseOutput(3)
    with pytest.raises(ValueError):
        d.set_n_intervals(-1)
    assert d.n_intervals == 3 and d.n_steps() == 0


def test_pack_unpack_round_trip():
    rng = np.random.default_rng(0)
    sizes = [3, 1, 7]
    xs = [rng.normal(size=6 * n) for n in sizes]
    p = pack(xs)
    assert p.dtype == np.float64 and p.size == 6 * sum(sizes)
    for a, b in zip(unpack(p, sizes), xs):
        assert np.array_equal(a, b)
    ms = [rng.uniform(size=n) for n in sizes]
    for a, b in zip(unpack(pack(ms), sizes, per_particle=1), ms):
        assert np.array_equal(a, b)
    with pytest.raises(ValueError):
        unpack(p[:-1], sizes)
