import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI library)")


@pytest.fixture(scope="session")
def hip_ctx():
    from nimbus_amd import lib as L
    ctx = L.Context(0)
    yield ctx
    ctx.close()


@pytest.fixture
def rng():
    return np.random.default_rng(12345)
