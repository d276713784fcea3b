"""GPU tests of the learned surrogate of the pair quadrature (two types).

* the training kernel's gradient against the FP64 back-propagation of
  oracle/mlp_oracle.py (FP32 kernel: <= 2e-4 of each network's largest
  gradient component);
* with every cell of (r, alpha_i, alpha_j) marked for fallback, the surrogate
  pass reproduces the mixed-precision quadrature: forces within 1e-5 and
  energy within 1e-6 of the FP64 oracle, the mixed path's own bar;
* trained networks: forces within 5e-2 of the FP64 oracle (infinity-norm
  relative), and the error-driven fallback never makes them worse.
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


def test_trained_surrogate_forces(trained_pair):
    lat, ff, orc, en_o = trained_pair
    f_o = orc.get("f")
    errs = {}
    for tol in (1e30, 0.05):
        rep = ff.surrogate_train(n_per_tp=1 << 18, epochs=4, batch_per_tp=512, lr=2e-3, seed=7, tol=tol)
        ff.set_precision("surrogate")
        en_g = ff.force_calc(1)
        errs[tol] = (rel_inf(ff.get("f"), f_o), abs(en_g[0] - en_o[0]) / abs(en_o[0]), ff.surrogate_fallbacks())
        print(f"tol {tol:g}: report {rep}", flush=True)
        print(f"tol {tol:g}: force rel {errs[tol][0]:.3e} energy rel {errs[tol][1]:.3e} fallbacks {errs[tol][2]}",
              flush=True)
        ff.set_precision("mixed")
    assert errs[1e30][2] == 0
    assert errs[1e30][0] <= 5e-2
    assert errs[0.05][0] <= errs[1e30][0] * 1.0001
