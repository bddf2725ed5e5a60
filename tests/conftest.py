import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library through the C-ABI)")


@pytest.fixture(scope="session")
def ctx():
    """One library context on device 0 for the whole GPU test session."""
    import pfem_amd

    c = pfem_amd.Context(0)
    yield c
    c.close()
