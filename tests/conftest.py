import os
import sys

import pytest

try:  # torch first: libvgm.so must bind to the HIP runtime torch loads
    import torch  # noqa: F401
except ImportError:
    pass

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libvgm.so on cuda:0)")


@pytest.fixture(scope="session")
def ctx():
    """One library context on device 0 for the whole GPU session."""
    import vecgeom_amd as vg

    c = vg.Context(0)
    yield c
    c.close()
