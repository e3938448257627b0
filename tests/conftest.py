import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libconanpm.so)")


def gpu_available() -> bool:
    return os.path.exists("/dev/kfd")


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.load()
    return oracle


@pytest.fixture(scope="session")
def pm():
    if not gpu_available():
        pytest.fail("GPU test selected but no GPU is visible")
    from conan_amd import pm as pmmod
    pmmod._lib.load()
    return pmmod
