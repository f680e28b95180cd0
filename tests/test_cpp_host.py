"""The C++ host mirror (ode_prototyping_amd/cpp/odehip.hpp) through the test
program tests/cpp/hip_rk_embedded_test.cpp, which is written like the
reference's CudaRKEmbeddedTest.cu / CudaStateTest.cu."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "hip_rk_embedded_test")


def test_cpp_test_program_is_built():
    assert os.path.exists(BIN), "run __graft_entry__.build() (it builds tests/cpp)"


@pytest.mark.gpu
def test_cpp_host_mirror_consistency():
    p = subprocess.run([BIN], capture_output=True, text=True, timeout=300)
    print(p.stdout[-4000:])
    assert p.returncode == 0, p.stdout[-4000:] + p.stderr[-2000:]
    assert "0 failures" in p.stdout
