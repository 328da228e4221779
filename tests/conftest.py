import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests"), os.path.join(ROOT, "tests", "golden")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


@pytest.fixture(scope="session")
def golden_q15():
    return np.load(os.path.join(ROOT, "tests", "golden", "golden_q15.npz"))


@pytest.fixture(scope="session")
def golden_q31():
    return np.load(os.path.join(ROOT, "tests", "golden", "golden_q31.npz"))


@pytest.fixture(scope="session")
def golden_float():
    return np.load(os.path.join(ROOT, "tests", "golden", "golden_float.npz"))
