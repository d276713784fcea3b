import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


@pytest.fixture(scope="session")
def tables():
    from mapp_amd import workload as W
    return W.johnson_tables()


@pytest.fixture(scope="session")
def small_tables():
    """coarser grid for finite-difference and closed-form tests"""
    from mapp_amd import workload as W
    return W.johnson_tables(nr=2001, nrho=2001)


def rel_inf(a, b):
    """max |a-b| / max |b| (infinity-norm relative error)"""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / (den if den > 0 else 1.0))
