import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "tests"))
sys.path.insert(0, os.path.join(REPO, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: long-running")


@pytest.fixture(scope="session")
def gpu_available():
    from clode_amd.cpp import clode_cpp_wrapper as w

    devices = w.query_opencl()[0].device_info
    if not devices:
        pytest.fail("no HIP device: GPU tests need an MI355X")
    return devices[0]
