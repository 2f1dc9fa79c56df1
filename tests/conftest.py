import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def oracle64():
    from oracle import Oracle
    return Oracle(64)


@pytest.fixture(scope="session")
def oracle32():
    from oracle import Oracle
    return Oracle(32)


@pytest.fixture(scope="session")
def lib():
    import naga_amd
    return naga_amd.load()
