import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


@pytest.fixture(scope="session")
def gpu_lib():
    """The product HIP library, loaded through the package loader (fails loudly if absent)."""
    from pyalign_amd import _lib
    return _lib.lib()
