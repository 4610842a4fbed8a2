import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: longer CPU test")


@pytest.fixture(scope="session")
def gpu():
    """Asserts a usable GPU: the HIP path must be the one that runs (no CPU fallback)."""
    from cuda_pf_localization_amd import device_available
    assert device_available(), "pfc_device_available() == 0 on a GPU test run"
    return True
