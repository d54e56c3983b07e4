import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")
    config.addinivalue_line("markers", "slow: long CPU run")


@pytest.fixture(scope="session")
def pfg():
    import platelet_fiber_gpu_amd as m
    m.lib()
    return m


@pytest.fixture(scope="session")
def po():
    import pyoracle
    pyoracle.load()
    return pyoracle


@pytest.fixture(scope="session")
def gpu(pfg):
    if pfg.device_count() < 1:
        pytest.fail("GPU test collected but no HIP device is visible")
    return pfg
