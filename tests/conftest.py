import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libpcr.so on the device)")


@pytest.fixture(scope="session")
def ctx():
    """One device context for all GPU tests; fails loudly without a GPU."""
    import compute_rasterizer_amd as pcr
    c = pcr.Context(0)
    yield c
    c.close()
