"""Oracle pins for the Gauss-Hermite averaged primitives (src/ff_dmd_I.cpp:180-242)
by closed forms of Gaussian moments:

  rho(r) = r^2 (odd extension u^3)  and  r phi(r) = r^3 (odd extension u^3)
      -> rho_bar = phi_bar = r^2 + 3/(2 alpha), d/dr = 2r, d/dalpha = -3/(2 alpha^2)
  rho(r) = C, r phi(r) = C r -> rho_bar = phi_bar = C, d/dr = 0, d/dalpha = 0

for every rule with n >= 2 (exact through degree 2n-1); the derivatives use
the moments xi g and xi^2 g (degrees 4 and 5 here, src/ff_dmd_I.cpp:233-234),
so their closed forms hold from n = 3 on.  Also d/dr, d/dalpha
against central differences of the oracle's own values on the Ni-Cu tables."""
import numpy as np
import pytest

from mapp_amd import workload as W
from oracle import oracle as O


def _orc(tab, no_i):
    o = O.Oracle(nt=1, variant="DMD_I")
    o.set_fs(tab)
    o.set_params(no_i=no_i, alpha_min=1.0, alpha_max=1e9, masses=[50.0])
    return o


def _tab(rho_fn, rphi_fn, nr=4001, rc=5.0):
    dr = rc / (nr - 1)
    r = np.arange(nr) * dr
    return W.FSTables(["X"], nr, 0.01, nr, dr, np.zeros((1, nr)), rho_fn(r)[None, None, :],
                      rphi_fn(r)[None, :], [50.0])


@pytest.mark.parametrize("no_i", [2, 3, 5, 8, 14])
@pytest.mark.parametrize("r,alpha", [(2.5, 40.0), (3.1, 90.0), (1.9, 25.5)])
def test_cubic_moment_closed_form(no_i, r, alpha):
    o = _orc(_tab(lambda x: x * x, lambda x: x ** 3), no_i)
    out = o.pair_quadrature(r, alpha, 0, 0)
    want = r * r + 1.5 / alpha
    for v in (out[0], out[3]):
        assert abs(v - want) < 1e-11 * want
    if no_i < 3:
        return
    for v in (out[1], out[4]):
        assert abs(v - 2 * r) < 1e-9 * 2 * r
    for v in (out[2], out[5]):
        assert abs(v + 1.5 / alpha ** 2) < 1e-8 * 1.5 / alpha ** 2


@pytest.mark.parametrize("no_i", [1, 3, 5])
def test_constant_closed_form(no_i):
    C = 0.731
    o = _orc(_tab(lambda x: C + 0 * x, lambda x: C * x), no_i)
    out = o.pair_quadrature(2.7, 55.0, 0, 0)
    assert abs(out[0] - C) < 1e-13 and abs(out[3] - C) < 1e-13
    if no_i < 2:  # xi*g is degree 2 here; the one-point rule is exact through degree 1
        return
    assert abs(out[1]) < 1e-11 and abs(out[4]) < 1e-11
    assert abs(out[2]) < 1e-12 and abs(out[5]) < 1e-12


def test_single_point_rule_is_plain_value():
    """n = 1 (node 0, weight sqrt(pi)): the average is the function at r"""
    o = _orc(_tab(lambda x: x * x, lambda x: x ** 3), 1)
    out = o.pair_quadrature(2.2, 70.0, 0, 0)
    assert abs(out[0] - 2.2 ** 2) < 1e-12 and abs(out[3] - 2.2 ** 2) < 1e-12


def test_derivatives_match_finite_differences(tables):
    """With a rich rule (n = 14) the moment formulas for d/dr and d/dalpha agree
    with central differences of the averaged values; a sign slip or a dropped
    2 sqrt(alpha) f_1 term would not."""
    o = O.Oracle(nt=2, variant="DMD_I")
    o.set_fs(tables)
    o.set_params(no_i=14, masses=tables.masses)
    # the last case straddles the cutoff, where the truncated window converges slowly
    for (r, a, rs, ps, tol) in [(2.6, 41.0, 0, 0, 1e-6), (3.55, 33.0, 1, 1, 1e-6), (4.4, 52.0, 3, 2, 1e-6),
                                (5.05, 47.0, 2, 1, 1e-2)]:
        out = o.pair_quadrature(r, a, rs, ps)
        h, ha = 1e-5, 1e-4
        up, dn = o.pair_quadrature(r + h, a, rs, ps), o.pair_quadrature(r - h, a, rs, ps)
        ua, da = o.pair_quadrature(r, a + ha, rs, ps), o.pair_quadrature(r, a - ha, rs, ps)
        for k0 in (0, 3):
            fd_r = (up[k0] - dn[k0]) / (2 * h)
            fd_a = (ua[k0] - da[k0]) / (2 * ha)
            assert abs(out[k0 + 1] - fd_r) < tol * (abs(fd_r) + 1e-6), (r, a, k0, out[k0 + 1], fd_r)
            assert abs(out[k0 + 2] - fd_a) < 10 * tol * (abs(fd_a) + 1e-9), (r, a, k0, out[k0 + 2], fd_a)


def test_three_point_derivatives_are_the_rules_moments(tables):
    """With n = 3 the reference's derivative formulas differ from the derivative
    of its own 3-point values (a property of the method, not of this restatement):
    the gap is O(1e-3) relative in the nearest-neighbour shell and vanishes as n grows"""
    gaps = []
    for n in (3, 7, 14):
        o = O.Oracle(nt=2, variant="DMD_I")
        o.set_fs(tables)
        o.set_params(no_i=n, masses=tables.masses)
        out = o.pair_quadrature(2.6, 41.0, 0, 0)
        h = 1e-5
        fd = (o.pair_quadrature(2.6 + h, 41.0, 0, 0)[3] - o.pair_quadrature(2.6 - h, 41.0, 0, 0)[3]) / (2 * h)
        gaps.append(abs(out[4] - fd) / abs(fd))
    assert 1e-5 < gaps[0] < 1e-2 and gaps[1] < 1e-8 and gaps[2] < 1e-8


def test_rule_beyond_cutoff_is_skipped(tables):
    """lower = (r - rc) sqrt(alpha) >= xi_max -> nothing computed (src/ff_dmd_I.cpp:187)"""
    o = O.Oracle(nt=2, variant="DMD_I")
    o.set_fs(tables)
    o.set_params(no_i=3, masses=tables.masses)
    xi_max = 1.2247448713915890491
    a = 30.0
    r_edge = 5.0 + xi_max / np.sqrt(a)
    assert o.pair_quadrature(r_edge + 1e-6, a, 0, 0) is None
    assert o.pair_quadrature(r_edge - 1e-3, a, 0, 0) is not None
