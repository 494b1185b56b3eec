import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C-ABI on cuda:0)")


def ans_check(result, ans, tol):
    """Element-wise check of the reference's test harness
    (test/test_utils.hpp:52-84): absolute error where ans <= tol, relative
    error otherwise; every element must be below tol."""
    result = np.asarray(result, dtype=np.float64)
    ans = np.asarray(ans, dtype=np.float64)
    err = np.where(ans <= tol, np.abs(result - ans), np.abs((result - ans) / np.where(ans == 0, 1, ans)))
    return bool(np.all(err < tol))


@pytest.fixture(scope="session")
def oracle():
    import oracle as o
    o.lib()
    return o
