import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: larger CPU-side graphs")


@pytest.fixture(scope="session")
def gpu():
    """The device is required: GPU tests fail (not skip) if the native path cannot run."""
    import grgl_amd

    n = grgl_amd.device_count()
    assert n > 0, "no GPU visible to libgrgl_hip.so"
    return n
