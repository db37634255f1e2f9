import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def hip64():
    from noaidi_amd import prima
    k = prima.HipKernels(0, 64)
    yield k
    k.close()


@pytest.fixture(scope="session")
def hip32():
    from noaidi_amd import prima
    k = prima.HipKernels(0, 32)
    yield k
    k.close()
