import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libwpf_hip.so)")
    config.addinivalue_line("markers", "slow: longer CPU oracle runs")


def gpu_available() -> bool:
    try:
        from mizuho_edit_amd import _lib
        return _lib.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """Skip nothing silently: a gpu-marked test on a box without a GPU or
    without the built library fails."""
    from mizuho_edit_amd import _lib
    n = _lib.device_count()
    assert n > 0, "no GPU visible to libwpf_hip.so"
    return n
