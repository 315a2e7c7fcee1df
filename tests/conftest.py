import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def golden(name):
    return dict(np.load(os.path.join(GOLDEN, name), allow_pickle=False))


def relerr_vec(a, b):
    """(rms, max) over bodies of |a-b| / |b| for (n,3) arrays."""
    e = np.sqrt(np.sum((a - b) ** 2, axis=1) / np.sum(b ** 2, axis=1))
    return float(np.sqrt(np.mean(e ** 2))), float(e.max())


def relerr_scalar(a, b):
    e = np.abs(a - b) / np.abs(b)
    return float(np.sqrt(np.mean(e ** 2))), float(e.max())


@pytest.fixture(scope="session")
def oracle_lib():
    import pyoracle
    return pyoracle
