import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on the device)")


def gpu_available() -> bool:
    try:
        from trax_amd.api import Device
        Device(0).close()
        return True
    except Exception:
        return False


@pytest.fixture(scope="session")
def dev():
    from trax_amd.api import Device
    d = Device(0)
    yield d
    d.close()
