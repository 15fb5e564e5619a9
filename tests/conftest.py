import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and the built libsinv_hip.so")


@pytest.fixture(scope="session")
def built_lib():
    from sinv_amd import _native
    if not os.path.exists(_native.LIB_PATH):
        from sinv_amd import build
        build.build()
    return _native.lib()
