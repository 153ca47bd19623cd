import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def gpu_available():
    from aretomo3_amd import device_count
    if device_count() < 1:
        pytest.fail("no HIP device visible: the gpu tests need the MI355X")
    return True
