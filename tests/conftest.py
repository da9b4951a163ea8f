import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs libbmx_hip kernels)")


@pytest.fixture(scope="session")
def eng():
    """libbmx_hip context on device 0 (GPU tests only)."""
    from biomedisa_amd import _hip
    e = _hip.engine(0)
    yield e
