import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")
    config.addinivalue_line("markers", "slow: long CPU-side test")


def gpu_available() -> bool:
    try:
        from damavand_amd import _native
        return _native.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def native():
    """The loaded HIP library; a GPU test fails loudly when it is missing."""
    from damavand_amd import _native
    n = _native.device_count()
    assert n > 0, "no GPU visible to the HIP library"
    return _native
