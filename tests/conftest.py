import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests"), os.path.join(ROOT, "tests", "golden")):
    if p not in sys.path:
        sys.path.insert(0, p)


# Tests race first-use tuning between threads and force variants by name: their timings
# must not reach the persisted table of tuned choices that later processes (a benchmark)
# reuse.  The tests of the table point it at a temporary file of their own.
os.environ["KISSFFT_AMD_TUNE_TABLE"] = "0"


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
