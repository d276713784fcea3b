"""GPU parity of MEAM (ForceField_meam, src/ff_meam.cpp) through the C ABI
against the oracle on the same seeded inputs.

Bars: the full neighbour list bit-exact; energy, virial and forces <= 1e-12
relative (infinity norm over the array; the GPU sums each site's terms in
another order than the reference's sequential loops); energy_calc <= 1e-12.
Cases: one element, two elements (L12 cross lattice, ialloy 0, 1, 2),
partially screened pairs with the reference's force loop and the corrected
one, every lattice / ibar / erose form / second-neighbour series / background
option (bcc with nn2, diamond, B1, B2, C11), and a perfect fcc crystal whose
energy is the Rose energy (the pair potential's construction, through the
product's own tables)."""
import numpy as np
import pytest

from tests.conftest import rel_inf

pytestmark = pytest.mark.gpu


def _pair(state, H, s, t, fix=0, skin=0.3):
    from mapp_amd.dmd import setup_meam
    from oracle import oracle as O
    ff = setup_meam(H, s, t, state, skin=skin, fix=fix)
    orc = O.Oracle(nt=int(state["nt"]), variant="MEAM")
    orc.meam_coef(state)
    orc.meam_set_fix(fix)
    orc.set_system(H, s.copy(), np.asarray(t, dtype=np.float64).reshape(-1, 1))
    orc.init(skin)
    return ff, orc


def _compare(ff, orc, label):
    goff, gnl = ff.neighbor_list()
    ooff, onl = orc.neighbor_list()
    assert np.array_equal(goff, ooff) and np.array_equal(gnl, onl), label
    eg, eo = ff.force_calc(1), orc.force_calc(1)
    print(label, "energy", eg[0], eo[0], "entries", len(gnl), flush=True)
    assert abs(eg[0] - eo[0]) <= 1e-12 * abs(eo[0])
    assert rel_inf(eg[1:], eo[1:]) <= 1e-12
    err = rel_inf(ff.get("f"), orc.get("f"))
    print(label, "force rel", err, flush=True)
    assert err <= 1e-12
    assert abs(ff.energy_calc() - orc.energy_calc()) <= 1e-12 * abs(eo[0])


def _screening_local(ialloy=0):
    from mapp_amd import workload as W
    return W.meam_local(2, ialloy=ialloy) + [("Cmin", 1, 1, 1, 0.5), ("Cmin", 2, 2, 2, 0.6), ("Cmin", 1, 2, 1, 0.7),
                                             ("Cmin", 1, 2, 2, 0.8), ("Cmin", 1, 1, 2, 0.55), ("Cmin", 2, 2, 1, 0.65)]


@pytest.mark.parametrize("n,seed", [(4, 1), (6, 2)])
def test_one_element(n, seed):
    from mapp_amd import workload as W
    st = W.meam_state(["Aa"], local=[("rc", 4.5)])
    H, s, t = W.md_lattice(n, 3.55, "fcc", nt=1, seed=seed, jitter=0.05)
    ff, orc = _pair(st, H, s, t)
    _compare(ff, orc, f"one {n}")
    ff.close()


@pytest.mark.parametrize("ialloy", [0, 1, 2])
def test_two_elements(ialloy):
    from mapp_amd import workload as W
    st = W.meam_state(["Aa", "Bb"], local=W.meam_local(2, ialloy=ialloy))
    H, s, t = W.md_lattice(5, 3.58, "fcc", nt=2, seed=10 + ialloy, jitter=0.05)
    ff, orc = _pair(st, H, s, t)
    _compare(ff, orc, f"two ialloy {ialloy}")
    ff.close()


@pytest.mark.parametrize("fix", [0, 1])
@pytest.mark.parametrize("ialloy", [0, 1])
def test_partial_screening_periodic(fix, ialloy):
    """second neighbours partially screened across the periodic boundaries
    (three-body forces on screening atoms, dropped for ghosts as in the
    reference)"""
    from mapp_amd import workload as W
    st = W.meam_state(["Aa", "Bb"], local=_screening_local(ialloy))
    H, s, t = W.md_lattice(4, 3.58, "fcc", nt=2, seed=20 + ialloy, jitter=0.06)
    ff, orc = _pair(st, H, s, t, fix=fix)
    _compare(ff, orc, f"screened fix {fix} ialloy {ialloy}")
    ff.close()


@pytest.mark.parametrize("name", ["bcc_nn2", "dia", "ibar0", "ibar4_gsmooth", "ibar_m5", "erose1",
                                  "erose2_emb_lin_neg", "b2_nn2", "c11", "b1_bkgd_dyn"])
def test_variants(name):
    """every lattice, ibar, erose form, nn2 series and background option"""
    from mapp_amd import workload as W
    st, kind, a, nt = W.meam_variants()[name]
    n = {"bcc": 4, "dia": 2, "fcc": 4}[kind]
    H, s, t = W.md_lattice(n, a, kind, nt=nt, seed=7, jitter=0.05)
    ff, orc = _pair(st, H, s, t)
    _compare(ff, orc, name)
    ff.close()


def test_partial_screening_cluster():
    from mapp_amd import workload as W
    st = W.meam_state(["Aa", "Bb"], local=_screening_local())
    Hs, s0, t = W.md_lattice(2, 3.58, "fcc", nt=2, seed=6, jitter=0.08)
    H = np.diag([30.0] * 3)
    s = (s0 @ Hs + 9.0) / 30.0
    for fix in (0, 1):
        ff, orc = _pair(st, H, s, t, fix=fix)
        _compare(ff, orc, f"cluster fix {fix}")
        f = ff.get("f")
        assert np.abs(f.sum(axis=0)).max() <= 1e-12 * np.abs(f).sum()
        ff.close()


def test_rose_energy_on_gpu():
    from mapp_amd import workload as W
    from mapp_amd.dmd import setup_meam
    st = W.meam_state(["Aa"], local=[("rc", 4.5)])
    dr = 1.1 * 4.5 / 1000
    r1 = 507 * dr
    H, s, t = W.md_lattice(5, r1 * np.sqrt(2.0), "fcc", nt=1, jitter=0.0)
    ff = setup_meam(H, s, t, st)
    en = ff.force_calc(1)
    astar = st["alpha"][0, 0] * (r1 / st["re"][0, 0] - 1.0)
    want = -st["Ec"][0, 0] * (1.0 + astar) * np.exp(-astar)
    assert abs(en[0] / len(t) - want) <= 1e-12 * abs(want)
    assert np.abs(ff.get("f")).max() <= 1e-12 * abs(want)
    ff.close()


def test_reader_setup_matches_coef(tmp_path):
    """ff_coef MEAM <library> <local> gives the same force field as the coefficients"""
    from mapp_amd import workload as W
    from mapp_amd.dmd import ForceFieldDMD, setup_meam
    lib, loc = str(tmp_path / "lib.meam"), str(tmp_path / "loc.meam")
    W.write_meam_lib(lib, ["Bb", "Aa"])
    local = W.meam_local(2)
    W.write_meam_local(loc, local)
    st = W.meam_state(["Aa", "Bb"], lib_order=["Bb", "Aa"], local=local)
    H, s, t = W.md_lattice(4, 3.58, "fcc", nt=2, seed=31, jitter=0.05)
    a = setup_meam(H, s, t, st)
    b = ForceFieldDMD(no_types=2, variant="MEAM")
    b.read_meam_params(lib, loc, ["Aa", "Bb"])
    b.set_system(H, s, t.reshape(-1, 1))
    b.init(0.3)
    ea, eb = a.force_calc(1), b.force_calc(1)
    assert np.array_equal(ea, eb)
    assert np.array_equal(a.get("f"), b.get("f"))
    a.close()
    b.close()
