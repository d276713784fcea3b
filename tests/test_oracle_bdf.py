"""Pins of the oracle's BDF clock coefficients (Clock_BDF::interpolate restated).

The predictor interpolates y at t[0..q-1] and dy at t[0] with a polynomial of
degree q: it must reproduce every polynomial of degree <= q at t_new, and the
corrector's derivative weights must give that polynomial's derivative at
t_new.  Order 1 is implicit Euler (a = -y, beta = h); order 2 on uniform
steps is textbook BDF-2 (c - 4/3 y_n + 1/3 y_{n-1} = 2h/3 f(c)).

The reference adds the (t_new - t[0]) c1 c2 term of dalpha_y[0]
(src/clock_bdf.cpp:327) where the derivative of its own alpha_y[0] subtracts
it; with the correction (fix=True) the derivative is exact, without it the
derivative weights do not even sum to zero for q >= 2.  The device clock
replicates the reference by default."""
import numpy as np
import pytest

from oracle.oracle import bdf_coef


def _times(q, rng):
    h = rng.uniform(0.5, 2.0, size=q - 1)
    return np.concatenate([[5.0], 5.0 - np.cumsum(h)])


def test_order_one_is_implicit_euler():
    cf = bdf_coef(1, 0.3, [2.0])
    assert cf["beta"] == 0.3 and abs(cf["alpha_dy_0"] - 0.3) < 1e-15
    assert cf["alpha_y"][0] == 1.0 and cf["dalpha_y"][0] == 0.0 and cf["dalpha_dy_0"] == 1.0


def test_order_two_uniform_is_textbook_bdf2():
    h = 0.1
    cf = bdf_coef(2, h, [1.0, 1.0 - h], fix=True)
    b = cf["beta"]
    assert abs(b - 2 * h / 3) < 1e-15
    # a = beta (dalpha_dy_0 dy + sum dalpha_y y) - (alpha_dy_0 dy + sum alpha_y y)
    coef_y = b * cf["dalpha_y"] - cf["alpha_y"]
    coef_dy = b * cf["dalpha_dy_0"] - cf["alpha_dy_0"]
    np.testing.assert_allclose(coef_y, [-4 / 3, 1 / 3], rtol=1e-13)
    assert abs(coef_dy) < 1e-14
    # the reference's sign drops y_n from the corrector
    ref = bdf_coef(2, h, [1.0, 1.0 - h])
    assert abs(b * ref["dalpha_y"][0] - ref["alpha_y"][0]) < 1e-13


@pytest.mark.parametrize("q", [2, 3, 4, 5])
def test_predictor_and_derivative_exact_for_polynomials(q):
    rng = np.random.default_rng(q)
    t = _times(q, rng)
    dt = 0.7
    tn = t[0] + dt
    cf = bdf_coef(q, dt, t, fix=True)
    ref = bdf_coef(q, dt, t)
    np.testing.assert_allclose(cf["alpha_y"], ref["alpha_y"], rtol=0, atol=0)
    for deg in range(q + 1):
        p = np.polynomial.Polynomial(rng.normal(size=deg + 1))
        dp = p.deriv()
        pred = cf["alpha_dy_0"] * dp(t[0]) + np.dot(cf["alpha_y"], p(t))
        assert abs(pred - p(tn)) <= 1e-9 * max(1.0, abs(p(tn)))
        der = cf["dalpha_dy_0"] * dp(t[0]) + np.dot(cf["dalpha_y"], p(t))
        assert abs(der - dp(tn)) <= 1e-9 * max(1.0, abs(dp(tn)))
    assert abs(np.sum(cf["dalpha_y"])) < 1e-9
    assert abs(np.sum(ref["dalpha_y"])) > 1e-3
