import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (gfx950); runs the HIP path")


@pytest.fixture(scope="session")
def orc():
    import oracle

    return oracle.Oracle()


@pytest.fixture(scope="session")
def ref():
    import oracle

    if not oracle.reference_available():
        pytest.skip("reference CPU build (oracle/_ref) not present")
    return oracle.Reference()


@pytest.fixture(scope="session")
def cpu_threads():
    return min(16, os.cpu_count() or 1)
