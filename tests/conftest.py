import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built libcudaDecon.so")


@pytest.fixture(scope="session")
def cdn():
    import cudadecon_amd
    cudadecon_amd.lib()
    return cudadecon_amd


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle
