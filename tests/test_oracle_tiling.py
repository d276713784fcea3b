"""The tiling argument behind the full-size GPU checks, pinned on the oracle
(FP64, 1e-12 relative): a 4^3 FCC cell tiled 2 x 2 x 2 reproduces the cell.

Tiling turns pairs that both ends own into pairs with a ghost end, so it is
an invariance of the reference only where its slot and index quirks vanish:
  - two types with per-site widths and symmetric density tables: force_calc
    (energy, virial, f, E, dE) and the consistent energy; not energy_calc's
    triangle (src/ff_dmd_I.cpp:508) nor c_d (dE[jatm+jtype], :2336-2339);
  - one type: every quantity, c_d and calc_g included.
The last test pins that with the reference's quirks (per-type widths,
asymmetric density tables) the energy is not tile-invariant."""
import numpy as np

from tests.conftest import rel_inf, tile


def _check_force(os_, ob, n):
    es, eb = os_.force_calc(1), ob.force_calc(1)
    assert abs(eb[0] - n * es[0]) <= 1e-12 * abs(n * es[0])
    assert rel_inf(np.asarray(eb[1:7]), n * np.asarray(es[1:7])) <= 1e-12
    for name in ("f", "E", "dE"):
        assert rel_inf(ob.get(name), np.tile(os_.get(name), (n, 1))) <= 1e-12, name


def test_two_types_quirk_free_tiling():
    from mapp_amd import workload as W
    from oracle import oracle as O
    tables = W.johnson_tables(symmetric_density=True)
    small = W.fcc_alloy(4, seed=12, alpha_per_site=True)
    big = tile(small, 2)
    os_, ob = O.make(small, tables, no_i=3), O.make(big, tables, no_i=3)
    _check_force(os_, ob, 8)
    for o in (os_, ob):
        o.set_energy_full(True)
    assert abs(ob.energy_calc() - 8 * os_.energy_calc()) <= 1e-12 * abs(8 * os_.energy_calc())
    ob.create_2nd_neigh_lst()
    ob.c_d_calc()
    cd = ob.get("c_d")
    assert np.abs(cd.sum(axis=0)).max() <= 1e-12 * np.abs(cd).sum()


def test_one_type_tiling():
    from mapp_amd import workload as W
    from oracle import oracle as O
    tables = W.johnson_tables(elements=("Ni",))
    small = W.fcc_alloy(4, nt=1, seed=13)
    big = tile(small, 2)
    os_, ob = O.make(small, tables, no_i=3), O.make(big, tables, no_i=3)
    _check_force(os_, ob, 8)
    assert abs(ob.energy_calc() - 8 * os_.energy_calc()) <= 1e-12 * abs(8 * os_.energy_calc())
    for o in (os_, ob):
        o.create_2nd_neigh_lst()
        o.c_d_calc()
    assert rel_inf(ob.get("c_d"), np.tile(os_.get("c_d"), (8, 1))) <= 1e-12
    a = np.random.default_rng(3).uniform(-0.2, 0.2, size=small.c.shape)
    cs, gs = os_.calc_g(0, 0.37, a)
    cb, gb = ob.calc_g(0, 0.37, np.tile(a, (8, 1)))
    assert abs(cb - 8 * cs) <= 1e-12 * abs(8 * cs)
    assert rel_inf(gb, np.tile(gs, (8, 1))) <= 1e-12


def test_reference_quirks_break_tiling(tables):
    from mapp_amd import workload as W
    from oracle import oracle as O
    small = W.fcc_alloy(4, seed=12)
    big = tile(small, 2)
    es = O.make(small, tables, no_i=3).force_calc(1)
    eb = O.make(big, tables, no_i=3).force_calc(1)
    assert abs(eb[0] - 8 * es[0]) > 1e-6 * abs(8 * es[0])


def test_tiling_detects_a_displaced_site():
    from mapp_amd import workload as W
    from oracle import oracle as O
    tables = W.johnson_tables(elements=("Ni",))
    small = W.fcc_alloy(4, nt=1, seed=13)
    big = tile(small, 2)
    big.x[5, 0] += 0.01 / big.H[0, 0]
    os_, ob = O.make(small, tables, no_i=3), O.make(big, tables, no_i=3)
    os_.force_calc(1)
    ob.force_calc(1)
    assert rel_inf(ob.get("f"), np.tile(os_.get("f"), (8, 1))) > 1e-6
