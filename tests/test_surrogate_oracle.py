"""Pins of the FP64 surrogate oracle (forward, loss, back-propagation)."""
import numpy as np
import pytest

from oracle import mlp_oracle as M


def _norm(rng):
    y_mean = rng.normal(size=(4, M.NOUT))
    y_std = rng.uniform(0.5, 2.0, size=(4, M.NOUT))
    return np.concatenate([[1.8, 6.0, 60.0, 120.0, 3.9, 2.1, 90.0, 30.0], y_mean.ravel(), y_std.ravel()])


def _samples(rng, n):
    X = np.stack([rng.uniform(1.8, 6.0, n), rng.uniform(60, 120, n), rng.uniform(60, 120, n)], axis=1)
    T = rng.normal(size=(n, M.NOUT))
    return X, T


def test_parameter_count_matches_layout():
    assert M.NPARAM == 8966 and M.NNORM == 56


def test_gradient_matches_central_differences():
    rng = np.random.default_rng(3)
    flat = M.glorot(rng)
    norm = _norm(rng)
    X, T = _samples(rng, 16)
    inv = 1.0 / (16 * M.NOUT)
    _, g = M.loss_and_grad(flat, norm, 2, X, T, inv)
    # one coordinate in every parameter block, plus random ones
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in M.SHAPES])
    idx = list(offs[:-1] + 1) + list(rng.integers(0, M.NPARAM, 24))
    h = 1e-6
    for k in idx:
        e = np.zeros_like(flat)
        e[k] = h
        lp, _ = M.loss_and_grad(flat + e, norm, 2, X, T, inv)
        lm, _ = M.loss_and_grad(flat - e, norm, 2, X, T, inv)
        fd = (lp - lm) * inv / (2 * h)
        assert abs(fd - g[k]) <= 1e-7 + 1e-6 * abs(fd), (k, fd, g[k])


def test_zero_hidden_network_closed_form():
    rng = np.random.default_rng(4)
    d = {name: np.zeros(shp) for name, shp in M.SHAPES}
    d["b4"] = rng.normal(size=M.NOUT)
    flat = M.pack(d)
    norm = _norm(rng)
    X, T = _samples(rng, 8)
    tp = 1
    y = M.forward(flat, norm, tp, X)
    _, _, _, _, y_mean, y_std = M.norm_parts(norm)
    np.testing.assert_allclose(y, np.broadcast_to(d["b4"] * y_std[tp] + y_mean[tp], y.shape), rtol=1e-14)
    inv = 1.0 / (8 * M.NOUT)
    _, g = M.loss_and_grad(flat, norm, tp, X, T, inv)
    t = (T - y_mean[tp]) / y_std[tp]
    gb4 = (2.0 * (d["b4"] - t) * inv).sum(0)
    np.testing.assert_allclose(g[-M.NOUT:], gb4, rtol=1e-13)
    # with zero weights downstream, nothing flows to W1..W3 and b1..b3 except through W4 = 0
    assert np.all(g[:M.NPARAM - M.NOUT - M.HID * M.NOUT] == 0.0)


def test_silu_derivative():
    z = np.linspace(-8, 8, 101)
    fd = (M.silu(z + 1e-6) - M.silu(z - 1e-6)) / 2e-6
    np.testing.assert_allclose(M.dsilu(z), fd, rtol=1e-7, atol=1e-9)
