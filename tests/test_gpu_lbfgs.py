"""GPU minimiser (mapp_dmd_min_lbfgs) against the oracle's restatement of
Min_LBFGS + LineSearch_BackTrack on the same seeded inputs.

Bars: same exit code, step count and force/energy call counts; energy after
every step within 1e-10 relative (FP64 path, deterministic device dot
products against the oracle's sequential sums)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(lat, tables, consistent, max_iter, e_tol, m_it=2):
    from mapp_amd.dmd import setup
    from oracle import oracle as O
    ff = setup(lat, tables, no_i=3)
    orc = O.make(lat, tables, no_i=3)
    ff.set_energy_mode(consistent)
    orc.set_energy_full(consistent)
    g = ff.min_lbfgs(max_iter, m_it, e_tol)
    o = orc.min_lbfgs(max_iter, m_it, e_tol)
    x_g = ff.get("x")
    x_o = orc.x_all()[: lat.natms]
    ff.close()
    return g, o, x_g, x_o


def _check(g, o, x_g, x_o, tol=1e-10):
    code_g, en_g, nf_g, ne_g = g
    code_o, en_o, nf_o, ne_o = o
    print("gpu", code_g, len(en_g), nf_g, ne_g, "oracle", code_o, len(en_o), nf_o, ne_o, flush=True)
    assert (code_g, len(en_g), nf_g, ne_g) == (code_o, len(en_o), nf_o, ne_o)
    rel = np.max(np.abs(en_g - en_o) / np.abs(en_o))
    print("energy rel", rel, "x diff", np.max(np.abs(x_g - x_o)), flush=True)
    assert rel <= tol
    assert np.max(np.abs(x_g - x_o)) <= 1e-9 * max(1.0, np.max(np.abs(x_o)))


def test_lbfgs_reference_energy_stops_like_oracle(small_tables):
    from mapp_amd import workload as W
    lat = W.fcc_alloy(3, seed=4)
    _check(*_run(lat, small_tables, False, 20, 1e-8))


def test_lbfgs_consistent_energy_matches_oracle():
    """the setting where f = -grad U (symmetric densities, one alpha per site):
    a well-posed descent whose line-search decisions are not near ties"""
    from mapp_amd import workload as W
    tables = W.johnson_tables(symmetric_density=True)
    lat = W.fcc_alloy(5, seed=4, alpha_per_site=True)
    _check(*_run(lat, tables, True, 12, -1.0))


def test_lbfgs_one_type_matches_oracle(tables):
    from mapp_amd import workload as W
    t1 = W.johnson_tables(elements=("Ni",))
    lat = W.fcc_alloy(5, nt=1, seed=4)
    _check(*_run(lat, t1, False, 15, 1e-12, m_it=3))


def test_consistent_energy_equals_force_energy(tables):
    from mapp_amd import workload as W
    from mapp_amd.dmd import setup
    from oracle import oracle as O
    lat = W.fcc_alloy(6, seed=2)
    ff = setup(lat, tables, no_i=3)
    orc = O.make(lat, tables, no_i=3)
    ff.set_energy_mode(True)
    orc.set_energy_full(True)
    e_g, f_g = ff.energy_calc(), ff.force_calc(1)[0]
    e_o, f_o = orc.energy_calc(), orc.force_calc(1)[0]
    assert abs(e_g - f_g) <= 1e-12 * abs(f_g)
    assert abs(e_g - e_o) <= 1e-12 * abs(e_o)
    assert abs(f_g - f_o) <= 1e-12 * abs(f_o)
    ff.close()
