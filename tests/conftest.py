import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the HIP library")


@pytest.fixture(scope="session")
def gpu():
    """GPU tests run only where a device exists (``-m gpu``); they fail loudly,
    never skip, when the HIP library is missing."""
    import caviar_amd
    caviar_amd.load_library()
    return True
