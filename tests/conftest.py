import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def smg():
    """The HIP library's Python mirror, initialised on cuda:0 (GPU tests only)."""
    import schwinger_mg_amd as m
    m.lib()
    return m
