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


def tile(lat, reps):
    """the lattice repeated reps times along each box vector (replica-major
    site order); with the list cut below half the small box, every site of the
    result sees the same periodic environment as its image in lat"""
    from mapp_amd.workload import Lattice
    nt = lat.c.shape[1]
    g = np.arange(reps)
    shifts = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3).astype(np.float64)
    s = lat.x[:, :3]
    s_big = ((s[None, :, :] + shifts[:, None, :]) / reps).reshape(-1, 3)
    alpha = np.tile(lat.x[:, 3:3 + nt], (len(shifts), 1))
    c = np.tile(lat.c, (len(shifts), 1))
    H = lat.H * reps
    return Lattice(H=H, x=np.ascontiguousarray(np.concatenate([s_big, alpha], axis=1)), c=np.ascontiguousarray(c),
                   cart=s_big @ H)
