"""GPU checks at the full size of the 1M-site workloads (64^3 FCC cells,
1,048,576 sites), where the oracle would take minutes per call, through
properties that do not depend on the size.

Each 64^3 system is an 8^3 cell (2048 sites, checked against the FP64 oracle
here) tiled 8 x 8 x 8 times, so every site sees the same periodic environment
as its image in the cell: per-site results (forces, alpha forces, E, dE, c_d,
g) equal the cell's and extensive ones (energy, virial, calc_g cost) are 512
times the cell's, wherever the reference's slot and index quirks vanish
(two types with per-site widths and symmetric density tables for force_calc
and the consistent energy; one type for everything; the oracle pins which,
test_oracle_tiling.py).  The big runs go through the full-size neighbour
build, ghosts, lists of 40M pairs, packing and every kernel at its real grid.
Bars: 1e-12 relative (infinity norm over the array) for FP64, as for the
small parity tests; sums that vanish by symmetry (total force, total c_d) to
1e-12 of the summed magnitudes; repeated calls bitwise equal; FP32 quadrature
sums within 1e-5 of FP64 on forces."""
import numpy as np
import pytest

from tests.conftest import rel_inf, tile

pytestmark = pytest.mark.gpu

REPS = 8  # 8^3 cell tiled 8 x 8 x 8 = 64^3 cells
NREP = REPS ** 3


def _build(small, tables):
    from mapp_amd.dmd import setup
    from oracle import oracle as O
    big = tile(small, REPS)
    assert big.natms == 1048576
    return big, setup(small, tables, no_i=3), setup(big, tables, no_i=3), O.make(small, tables, no_i=3)


@pytest.fixture(scope="module")
def two_types():
    from mapp_amd import workload as W
    tables = W.johnson_tables(symmetric_density=True)
    small = W.fcc_alloy(8, seed=31, alpha_per_site=True)
    big, fs, fb, orc = _build(small, tables)
    yield tables, small, big, fs, fb, orc
    fs.close()
    fb.close()


@pytest.fixture(scope="module")
def one_type():
    from mapp_amd import workload as W
    tables = W.johnson_tables(elements=("Ni",))
    small = W.fcc_alloy(8, nt=1, seed=32)
    big, fs, fb, orc = _build(small, tables)
    yield tables, small, big, fs, fb, orc
    fs.close()
    fb.close()


def _per_site(a_small, a_big):
    return rel_inf(a_big, np.tile(a_small, (NREP, 1)))


def _force_checks(fs, fb, orc):
    es = fs.force_calc(1)
    eo = orc.force_calc(1)
    assert abs(es[0] - eo[0]) <= 1e-12 * abs(eo[0])
    assert rel_inf(fs.get("f"), orc.get("f")) <= 1e-12
    eb = fb.force_calc(1)
    print("energy", eb[0], NREP * es[0], flush=True)
    assert abs(eb[0] - NREP * es[0]) <= 1e-12 * abs(NREP * es[0])
    assert rel_inf(np.asarray(eb[1:7]), NREP * np.asarray(es[1:7])) <= 1e-12
    for name in ("f", "E", "dE"):
        err = _per_site(fs.get(name), fb.get(name))
        print(name, err, flush=True)
        assert err <= 1e-12, name
    f = fb.get("f")[:, :3]
    tot = np.abs(f.sum(axis=0)).max()
    print("total force", tot, "sum |f|", np.abs(f).sum(), flush=True)
    assert tot <= 1e-12 * np.abs(f).sum()


def test_two_types_force_calc_tiled(two_types):
    tables, small, big, fs, fb, orc = two_types
    _force_checks(fs, fb, orc)


def test_two_types_repeat_bitwise(two_types):
    tables, small, big, fs, fb, orc = two_types
    e1 = fb.force_calc(1)
    f1 = fb.get("f")
    e2 = fb.force_calc(1)
    f2 = fb.get("f")
    assert np.array_equal(np.asarray(e1), np.asarray(e2))
    assert np.array_equal(f1.view(np.uint64), f2.view(np.uint64))


def test_two_types_consistent_energy_tiled(two_types):
    tables, small, big, fs, fb, orc = two_types
    for o in (fs, fb):
        o.set_energy_mode(True)
    orc.set_energy_full(True)
    es, eb = fs.energy_calc(), fb.energy_calc()
    for o in (fs, fb):
        o.set_energy_mode(False)
    assert abs(es - orc.energy_calc()) <= 1e-12 * abs(es)
    assert abs(eb - NREP * es) <= 1e-12 * abs(NREP * es)


def test_two_types_c_d_sums_to_zero(two_types):
    tables, small, big, fs, fb, orc = two_types
    fb.force_calc(1)
    fb.create_2nd_neigh_lst()
    fb.c_d_calc()
    cd = fb.get("c_d")
    tot = np.abs(cd.sum(axis=0)).max()
    print("total c_d", tot, "sum |c_d|", np.abs(cd).sum(), flush=True)
    assert tot <= 1e-12 * np.abs(cd).sum()


def test_two_types_mixed_precision(two_types):
    from mapp_amd.dmd import setup
    tables, small, big, fs, fb, orc = two_types
    fb.force_calc(1)
    f64 = fb.get("f")
    fm = setup(big, tables, no_i=3, precision="mixed")
    fm.force_calc(1)
    err = rel_inf(fm.get("f"), f64)
    fm.close()
    print("mixed force rel", err, flush=True)
    assert err <= 1e-5


def test_one_type_everything_tiled(one_type):
    tables, small, big, fs, fb, orc = one_type
    _force_checks(fs, fb, orc)
    es, eb = fs.energy_calc(), fb.energy_calc()
    assert abs(es - orc.energy_calc()) <= 1e-12 * abs(es)
    assert abs(eb - NREP * es) <= 1e-12 * abs(NREP * es)
    for o in (fs, fb, orc):
        o.create_2nd_neigh_lst()
        o.c_d_calc()
    assert rel_inf(fs.get("c_d"), orc.get("c_d")) <= 1e-12
    for name in ("c_d", "n"):
        err = _per_site(fs.get(name), fb.get(name))
        print(name, err, flush=True)
        assert err <= 1e-12, name
    a = np.random.default_rng(7).uniform(-0.2, 0.2, size=small.c.shape)
    cost_o, g_o = orc.calc_g(0, 0.37, a)
    fs.set_vec("a", a)
    fb.set_vec("a", np.tile(a, (NREP, 1)))
    cost_s, cost_b = fs.calc_g(0, 0.37), fb.calc_g(0, 0.37)
    assert abs(cost_s - cost_o) <= 1e-12 * abs(cost_o)
    assert abs(cost_b - NREP * cost_s) <= 1e-12 * abs(NREP * cost_s)
    assert rel_inf(fs.get("g"), g_o) <= 1e-12
    assert _per_site(fs.get("g"), fb.get("g")) <= 1e-12
