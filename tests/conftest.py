import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); calls through libtwisty_hip.so")


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle

    if not os.path.exists(pyoracle.LIB_PATH):
        pyoracle.build()
    return pyoracle


@pytest.fixture(scope="session")
def hip_device():
    """Skips when no HIP device is visible; raises when the library is missing."""
    from twisty_amd import capi

    n = capi.device_count()
    if n < 1:
        pytest.skip("no HIP device visible")
    return 0
