import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


def relerr(a, b):
    """Normwise relative error ||a - b|| / ||b|| (||a - b|| when b == 0)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    d = np.linalg.norm(a - b)
    return d / nb if nb > 0 else d


class GoldenState:
    """Initial state stored in a golden fixture."""

    def __init__(self, d):
        self.species = d["in/species"].astype(np.int32)
        self.ids = d["in/ids"].astype(np.int32)
        self.x = d["in/x"]
        self.p = d["in/p"]
        self.E = d["in/E"]
        self.B = d["in/B"]


@pytest.fixture(scope="session")
def mini_golden():
    return np.load(os.path.join(ROOT, "tests", "golden", "mini_ref.npz"))
