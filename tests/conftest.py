import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.build()
    return O


@pytest.fixture(scope="session")
def gpu_mod():
    from lustrine_amd import gpu as G
    G.load()
    n = G.device_count()
    if n < 1:
        pytest.fail("no GPU visible: the gpu tests need an MI355X")
    return G


@pytest.fixture()
def ctx(gpu_mod):
    c = gpu_mod.Gpu(0)
    yield c
    c.close()
