import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the hip device)")
    config.addinivalue_line("markers", "slow: long-running")


@pytest.fixture(scope="session")
def hip_device():
    """The hip device, initialised once per test session."""
    from ospray_amd import api as A

    dev = A.init("hip")
    yield dev
