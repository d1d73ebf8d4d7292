import os
import sys

import pytest

TESTS = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(TESTS)
sys.path.insert(0, REPO)
sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X (gfx950) GPU and libfhe_amd.so")
    config.addinivalue_line(
        "markers", "reference: needs oracle/_ref (the compiled reference)")


@pytest.fixture(scope="session")
def gpu_lib():
    """The product library.  On a GPU run a missing build is an error, not a
    skip: there is no CPU fallback to test instead."""
    import fhe_amd
    return fhe_amd.library()
