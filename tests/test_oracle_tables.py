"""Oracle pins for the tables: the Gauss-Hermite rule against the literals the
reference hard-codes (src/ff_dmd_I.cpp:1841-2134, fixture tests/golden/
gh_reference.json) and against the mathematics of the rule; the cubic and
quartic spline rows (src/ff_dmd_I.cpp:1760-1836) against exact polynomials."""
import json
import math
import os

import numpy as np
import pytest

from mapp_amd import workload as W
from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "gh_reference.json")


def _orc_with(no_i, tables):
    o = O.Oracle(nt=len(tables.elements), variant="DMD_I")
    o.set_fs(tables)
    o.set_params(no_i=no_i, masses=tables.masses)
    return o


@pytest.fixture(scope="module")
def tiny_tables():
    return W.johnson_tables(nr=201, nrho=201)


@pytest.mark.parametrize("n", range(1, 15))
def test_gh_nodes_weights_match_reference_literals(tiny_tables, n):
    ref = json.load(open(GOLDEN))["rules"][str(n)]
    o = _orc_with(n, tiny_tables)
    xi, w0, w1, w2 = o.gh(n)
    rx = np.array([float(v) for v in ref["nodes"]])
    rw = np.array([float(v) for v in ref["weights"]])
    assert np.array_equal(xi, rx), (xi, rx)
    assert np.array_equal(w0, rw), (w0, rw)
    # wi_1 = wi_0 xi, wi_2 = wi_0 xi xi (src/ff_dmd_I.cpp:2129-2133)
    assert np.array_equal(w1, rw * rx)
    assert np.array_equal(w2, (rw * rx) * rx)


@pytest.mark.parametrize("n", range(1, 15))
def test_gh_rule_mathematics(tiny_tables, n):
    o = _orc_with(n, tiny_tables)
    xi, w0, _, _ = o.gh(n)
    hx, hw = np.polynomial.hermite.hermgauss(n)
    assert np.max(np.abs(xi - hx)) < 1e-13
    assert np.max(np.abs(w0 - hw) / hw) < 1e-12
    # exact for x^(2k) up to degree 2n-1: int x^2k e^-x^2 = Gamma(k + 1/2)
    for k in range(0, n):
        if 2 * k > 2 * n - 1:
            break
        exact = math.gamma(k + 0.5)
        assert abs(np.sum(w0 * xi ** (2 * k)) - exact) <= 1e-12 * max(1.0, exact)
        assert abs(np.sum(w0 * xi ** (2 * k + 1))) <= 1e-12 * max(1.0, exact)


def _poly_tables(nr, dr, rho_fn, rphi_fn, F_fn, drho):
    r = np.arange(nr) * dr
    rg = np.arange(nr) * drho
    return W.FSTables(["X"], nr, drho, nr, dr, F_fn(rg)[None, :], rho_fn(r)[None, None, :],
                      rphi_fn(r)[None, :], [50.0])


def test_cubic_spline_reproduces_cubics():
    """interpolate(): the 5-point derivative is exact for quartics, so interior
    rows of a cubic reproduce it, and c4..c6 are its derivative / delta"""
    nr, dr = 401, 0.0125
    f = lambda x: 0.3 * x ** 3 - 1.1 * x ** 2 + 0.7 * x + 2.0
    df = lambda x: 0.9 * x ** 2 - 2.2 * x + 0.7
    t = _poly_tables(nr, dr, f, f, lambda x: 0 * x, 0.01)
    o = O.Oracle(nt=1, variant="DMD_I")
    o.set_fs(t)
    o.set_params(no_i=3, masses=[50.0])
    rows = o.table(1, 0)
    for m in (2, 50, 200, 397):
        for p in (0.0, 0.25, 0.5, 0.9):
            x = (m + p) * dr
            c = rows[m]
            val = c[3] * p ** 3 + c[2] * p ** 2 + c[1] * p + c[0]
            der = c[6] * p ** 2 + c[5] * p + c[4]
            assert abs(val - f(x)) < 1e-12 * max(1, abs(f(x)))
            assert abs(der - df(x)) < 1e-9 * max(1, abs(df(x)))


def test_spline_c1_continuity_and_nodes(tables):
    o = O.Oracle(nt=2, variant="DMD_I")
    o.set_fs(tables)
    o.set_params(no_i=3, masses=tables.masses)
    for kind, slot in ((1, 0), (1, 3), (2, 2)):
        rows = o.table(kind, slot)
        # node interpolation c0 = f_i (slot 3 = type2rho[1][1] holds file rho[1][1])
        fi = {(1, 0): tables.rho[0, 0], (1, 3): tables.rho[1, 1], (2, 2): tables.rphi[2]}[(kind, slot)]
        assert np.array_equal(rows[:, 0], fi)
        # value and derivative continuous at interior nodes
        end_val = rows[:-1, 3] + rows[:-1, 2] + rows[:-1, 1] + rows[:-1, 0]
        assert np.max(np.abs(end_val[2:-2] - rows[3:-2, 0])) < 1e-12 * np.max(np.abs(rows[:, 0]))
        end_der = rows[:-1, 6] + rows[:-1, 5] + rows[:-1, 4]
        assert np.max(np.abs(end_der[2:-2] - rows[3:-2, 4])) < 1e-9 * np.max(np.abs(rows[:, 4]))


def test_quartic_F_rows_derivative_consistency(tables):
    """interpolate_m(): c5..c8 are the p-derivative of c0..c4 divided by delta"""
    o = O.Oracle(nt=2, variant="DMD_I")
    o.set_fs(tables)
    o.set_params(no_i=3, masses=tables.masses)
    rows = o.table(0, 1)
    drho = tables.drho
    m = np.arange(5, tables.nrho - 5)
    c = rows[m]
    assert np.allclose(c[:, 5], c[:, 1] / drho, rtol=1e-15, atol=0)
    assert np.allclose(c[:, 6], 2 * c[:, 2] / drho, rtol=1e-15, atol=0)
    assert np.allclose(c[:, 7], 3 * c[:, 3] / drho, rtol=1e-15, atol=0)
    assert np.allclose(c[:, 8], 4 * c[:, 4] / drho, rtol=1e-15, atol=0)
    # quartic term: c4 = (c2[i+1] - c2[i]) / 6 - c3[i] / 2
    assert np.allclose(c[:, 4], (rows[m + 1, 2] - c[:, 2]) / 6.0 - 0.5 * c[:, 3], rtol=1e-13, atol=1e-300)


def test_fs_file_roundtrip(tmp_path, tables):
    """write_fs emits %.17g, so the file parses back to the same doubles"""
    path = str(tmp_path / "NiCu.fs")
    W.write_fs(path, tables)
    lines = open(path).read().split("\n")
    assert lines[3].split() == ["2", "Ni", "Cu"]
    hdr = lines[4].split()
    assert int(hdr[0]) == tables.nrho and float(hdr[1]) == tables.drho
    assert int(hdr[2]) == tables.nr and float(hdr[3]) == tables.dr
    vals = np.array([float(v) for l in lines[6:] for v in l.split()])
    nt, nr, nrho = 2, tables.nr, tables.nrho
    assert vals.size == nt * nrho + nt * nt * nr + 3 * nr
    assert np.array_equal(vals[:nrho], tables.F[0])
    assert np.array_equal(vals[nrho:nrho + nr], tables.rho[0, 0])
    assert np.array_equal(vals[-nr:], tables.rphi[2])
