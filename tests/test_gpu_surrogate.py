"""GPU tests of the learned surrogate of the pair quadrature (two types).

* the training kernel's gradient against the FP64 back-propagation of
  oracle/mlp_oracle.py (FP32 kernel: <= 2e-4 of each network's largest
  gradient component);
* with every cell of (r, alpha_i, alpha_j) marked for fallback, the surrogate
  pass reproduces the mixed-precision quadrature: forces within 1e-5 and
  energy within 1e-6 of the FP64 oracle, the mixed path's own bar;
* the enforced bound: trained networks, mask derived for a relative force
  bound of 5e-6 (the FP32 fallback adds up to ~3e-6), on 6^3 and 8^3 cells:
  forces within 1e-5 of the FP64 oracle (5-point rule) with the networks
  serving a measured share of the (pair, type pair) evaluations; at looser
  bounds the share grows and the force error stays inside bound + 3e-6;
* training gain: at the same bound the trained networks serve more of the
  evaluations than the same networks left at their initial weights, and with
  every cell served their validation error (standardised units) is below 0.05
  and a tenth of the untrained one, force error under a fifth and energy error
  under a tenth of the untrained ones.
"""
import numpy as np
import pytest

from tests.conftest import rel_inf

pytestmark = pytest.mark.gpu


def _setup(lat, tables, no_i=5):
    from mapp_amd.dmd import setup
    from oracle import oracle as O
    ff = setup(lat, tables, no_i=no_i, precision="mixed")
    orc = O.make(lat, tables, no_i=no_i, variant="DMD_I")
    return ff, orc


def test_training_gradient_matches_oracle(small_tables):
    from mapp_amd import workload as W
    from oracle import mlp_oracle as M
    lat = W.fcc_alloy(3, seed=2)
    ff, _ = _setup(lat, small_tables)
    rng = np.random.default_rng(11)
    params = np.stack([M.glorot(rng) for _ in range(4)]).astype(np.float32)
    y_mean = rng.normal(size=(4, 6))
    y_std = rng.uniform(0.5, 2.0, size=(4, 6))
    norm = np.concatenate([[1.8, 6.0, 60.0, 120.0, 3.9, 2.1, 90.0, 30.0], y_mean.ravel(), y_std.ravel()]).astype(
        np.float32)
    n = 128
    X = np.stack([rng.uniform(1.8, 6.0, 4 * n), rng.uniform(60, 120, 4 * n), rng.uniform(60, 120, 4 * n)],
                 axis=1).astype(np.float32)
    T = rng.normal(size=(4 * n, 6)).astype(np.float32)
    grad, loss = ff.surrogate_gradient(params, norm, X, T)
    inv = 1.0 / (n * 6)
    for tp in range(4):
        l_ref, g_ref = M.loss_and_grad(params[tp].astype(np.float64), norm.astype(np.float64), tp,
                                       X[tp * n:(tp + 1) * n], T[tp * n:(tp + 1) * n], inv)
        err = np.max(np.abs(grad[tp] - g_ref)) / np.max(np.abs(g_ref))
        print(f"tp {tp}: grad rel {err:.2e} loss {loss[tp]:.6f} ref {l_ref * inv:.6f}", flush=True)
        assert err <= 2e-4
        assert abs(loss[tp] - l_ref * inv) <= 1e-5 * l_ref * inv
    ff.close()


def test_all_fallback_reproduces_quadrature(tables):
    from mapp_amd import workload as W
    lat = W.fcc_alloy(6, seed=3)
    ff, orc = _setup(lat, tables)
    rep = ff.surrogate_train(n_per_tp=4096, epochs=1, batch_per_tp=256, seed=5, tol=-1.0)
    assert rep["masked_frac"] == 1.0
    ff.set_precision("surrogate")
    en_g = ff.force_calc(1)
    en_o = orc.force_calc(1)
    ef = rel_inf(ff.get("f"), orc.get("f"))
    ee = abs(en_g[0] - en_o[0]) / abs(en_o[0])
    nf = ff.surrogate_fallbacks()
    print("all-fallback: energy rel", ee, "force rel", ef, "fallbacks", nf, flush=True)
    assert nf > 0
    assert ee <= 1e-6
    assert ef <= 1e-5
    ff.close()


@pytest.fixture(scope="module")
def trained_pair(tables):
    from mapp_amd import workload as W
    lat = W.fcc_alloy(6, seed=3)
    ff, orc = _setup(lat, tables)
    en_o = orc.force_calc(1)
    return lat, ff, orc, en_o


def _surrogate_errors(ff, f_o, en_o, bound=None, **train):
    rep = ff.surrogate_train(n_per_tp=1 << 18, epochs=8, batch_per_tp=512, seed=7, **train)
    if bound is not None:
        ff.surrogate_set_bound(bound)
    ff.set_precision("surrogate")
    en_g = ff.force_calc(1)
    out = (rep, rel_inf(ff.get("f"), f_o), abs(en_g[0] - en_o[0]) / abs(en_o[0]), ff.surrogate_fallbacks())
    ff.set_precision("mixed")
    return out


def _served_share(ff, bound):
    """fraction of the quadrature-branch (pair, type pair) evaluations the networks serve at bound"""
    ff.surrogate_set_bound(-1.0)
    ff.set_precision("surrogate")
    ff.force_calc(1)
    total = ff.surrogate_fallbacks()
    info = ff.surrogate_set_bound(bound)
    ff.force_calc(1)
    share = 1.0 - ff.surrogate_fallbacks() / total
    return share, info


@pytest.mark.parametrize("n", [6, 8, 10])
def test_trained_surrogate_meets_bound(tables, n):
    from mapp_amd import workload as W
    lat = W.fcc_alloy(n, seed=3 + n)
    ff, orc = _setup(lat, tables)
    en_o = orc.force_calc(1)
    f_o = orc.get("f")
    ff.surrogate_train(n_per_tp=1 << 18, epochs=8, batch_per_tp=512, seed=7, tol=5e-6)
    cb = ff.surrogate_cell_bounds()
    prof = np.min(cb, axis=(0, 2, 3))  # per r cell: the smallest bound any of its cells serves at
    print("smallest servable bound per r cell (r from 1.8 to the quadrature cut):",
          np.array2string(prof, precision=1), flush=True)
    rows = []
    for bound in (5e-6, 1e-4, 1e-3, 1e-2):
        share, info = _served_share(ff, bound)
        en_g = ff.force_calc(1)
        ef = rel_inf(ff.get("f"), f_o)
        ee = abs(en_g[0] - en_o[0]) / abs(en_o[0])
        rows.append((bound, share, info["served"], ef, ee))
        print(f"{lat.natms} sites, bound {bound:.0e}: served {share:.4f} of the evaluations "
              f"({info['served']:.4f} of the cells), force rel {ef:.3e}, energy rel {ee:.3e}", flush=True)
        assert ef <= bound + 3e-6
    print("tolerances at the last bound", info["eps"], "scales", info["scales"], flush=True)
    assert rows[0][3] <= 1e-5
    shares = [r[1] for r in rows]
    assert all(b >= a for a, b in zip(shares, shares[1:]))
    assert shares[-1] > 0.0
    ff.close()


def test_trained_surrogate_gain(trained_pair):
    """what training buys: the same networks with (nearly) no training against trained ones,
    both on the same baseline and envelope.  With every cell served: force and energy errors
    vs the FP64 oracle; at a bound of 1e-2: the share of the quadrature evaluations each
    replaces; and the median over cells of the smallest bound a cell can serve at."""
    lat, ff, orc, en_o = trained_pair
    f_o = orc.get("f")
    raw = _surrogate_errors(ff, f_o, en_o, lr=1e-12, tol=1e30)
    raw_share, _ = _served_share(ff, 1e-2)
    raw_cb = float(np.median(ff.surrogate_cell_bounds()))
    fit = _surrogate_errors(ff, f_o, en_o, lr=2e-3, tol=1e30)
    fit_share, _ = _served_share(ff, 1e-2)
    fit_cb = float(np.median(ff.surrogate_cell_bounds()))
    for name, (rep, ef, ee, nf), sh, cb in (("untrained", raw, raw_share, raw_cb), ("trained", fit, fit_share, fit_cb)):
        print(f"{name}: force rel {ef:.3e} energy rel {ee:.3e} (every cell served, fallbacks {nf}); "
              f"share served at 1e-2: {sh:.4f}; median cell bound {cb:.3e}; "
              f"network error {rep['val_mean_err']:.3e} mean {rep['val_max_err']:.3e} max (std units)", flush=True)
    assert fit[3] == 0
    assert fit[1] < raw[1] and fit[2] < raw[2]
    assert fit_share > raw_share
    assert fit_cb < raw_cb
