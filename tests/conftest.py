import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C ABI)")


@pytest.fixture(scope="session")
def gfx():
    import cudagfx_amd

    cudagfx_amd.load()
    cudagfx_amd.set_verbose(0)
    return cudagfx_amd


@pytest.fixture(scope="session")
def orc():
    import oracle

    return oracle
