"""The oracle's compensated energy and virial sums (orc_set_compensated).

The reference sums the DMD energy and virial as one running double sum over
every pair and site term (src/ff_dmd_I.cpp:250-257, :343-348, :403-430).  At
1,048,576 sites that is ~1e8 additions rounded at the scale of a 5.5e6 eV
total.  The compensated mode adds the same terms in the same order with
Neumaier's summation.

Pins:
  * on a small system (500 sites, ~1e5 terms) the two modes agree to 1e-13;
  * exactness: one type (every slot and index quirk vanishes), a 6^3-cell
    crystal repeated 4x4x4 times has exactly 64 times the energy and virial of
    the cell in exact arithmetic; the compensated sums reproduce that to
    2e-15, the plain running sum of the big system does not do better;
  * the size of the drift at the full workload (64^3 cells, two types): the
    plain sum is off the compensated one by 1e-12 to 1e-10 relative, so a
    1e-12 bar at that size needs the compensated yardstick.
"""
import numpy as np
import pytest

from tests.conftest import rel_inf, tile


def _orc(lat, tables, comp):
    from oracle import oracle as O
    o = O.make(lat, tables, no_i=3)
    o.set_compensated(comp)
    return o


def test_modes_agree_on_a_small_system(small_tables):
    from mapp_amd import workload as W
    lat = W.fcc_alloy(5, seed=3)
    a, b = _orc(lat, small_tables, False), _orc(lat, small_tables, True)
    ea, eb = a.force_calc(1), b.force_calc(1)
    assert abs(ea[0] - eb[0]) <= 1e-13 * abs(eb[0])
    assert rel_inf(ea[1:], eb[1:]) <= 1e-13
    assert abs(a.energy_calc() - b.energy_calc()) <= 1e-13 * abs(eb[0])
    # the per-site arrays do not depend on the mode
    assert np.array_equal(a.get("f"), b.get("f"))


def test_compensated_sum_is_exact_under_tiling():
    from mapp_amd import workload as W
    t1 = W.johnson_tables(elements=("Ni",), nr=2000, nrho=2000)
    cell = W.fcc_alloy(6, nt=1, seed=8)
    big = tile(cell, 4)
    e_cell = _orc(cell, t1, True).force_calc(1)
    e_big_c = _orc(big, t1, True).force_calc(1)
    e_big_p = _orc(big, t1, False).force_calc(1)
    exact = 64.0 * e_cell
    err_c = [abs(e_big_c[0] - exact[0]) / abs(exact[0]), rel_inf(e_big_c[1:], exact[1:])]
    err_p = [abs(e_big_p[0] - exact[0]) / abs(exact[0]), rel_inf(e_big_p[1:], exact[1:])]
    print("tiled 4^3: compensated", err_c, "plain", err_p, flush=True)
    assert max(err_c) <= 2e-15
    assert err_p[0] >= err_c[0]


def test_reference_sum_drifts_at_full_size():
    from mapp_amd import workload as W
    tables = W.johnson_tables()
    lat = W.fcc_alloy(64, seed=1)
    o = _orc(lat, tables, False)
    ep = o.force_calc(1)
    o.set_compensated(True)
    ec = o.force_calc(1)
    d = abs(ep[0] - ec[0]) / abs(ec[0])
    print("64^3: plain", ep[0], "compensated", ec[0], "rel", d, "abs", abs(ep[0] - ec[0]), flush=True)
    assert 1e-12 <= d <= 1e-10
