import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library on the device)")


@pytest.fixture(scope="session")
def hip():
    """The HIP library with a visible GPU; fails (does not skip) without one."""
    from ca3dmm_amd import device
    n = device.device_count()
    assert n > 0, "no HIP device visible: GPU tests must run on an MI355X"
    return device
