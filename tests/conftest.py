import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs libam kernels)")


@pytest.fixture(scope="session")
def gpu_lib():
    """libam.so loaded in-tree; a GPU test fails (not skips) if it is absent."""
    from adaptivemerging_amd import _lib

    return _lib.load()
