import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built liboc_asm.so")


@pytest.fixture(scope="session")
def gpu():
    """Initialises the library on device 0; fails loudly if the HIP library is missing."""
    import cm_amd
    cm_amd.initialise(0)
    yield cm_amd
