import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (gfx950) and libwurst_amd.so")
    config.addinivalue_line("markers", "slow: long-running check")


@pytest.fixture(scope="session")
def gpu_device():
    import wurstrenderer_amd as wr
    return wr.Device(0)
