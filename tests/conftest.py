import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


@pytest.fixture(scope="session")
def ctx():
    """One gorender_hip context on device 0 for the whole GPU session."""
    from gorender_amd.device import Context
    c = Context(0)
    yield c
    c.close()
