import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def engine():
    import desi_amd
    e = desi_amd.Engine(0)
    yield e
    e.close()


def catalog(name):
    import numpy as np
    return np.loadtxt(os.path.join(GOLDEN, "catalogs", name), dtype=np.float32)
