import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C-ABI on the GPU)")


@pytest.fixture(scope="session")
def H():
    import hwyblas_amd
    return hwyblas_amd


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.build()
    return oracle
