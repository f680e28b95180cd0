import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


@pytest.fixture(scope="session")
def exe():
    """One HIP execution context for the GPU tests (fails loudly without a device)."""
    import ode_prototyping_amd as oh
    e = oh.HipExecutor(0)
    yield e
    e.close()
