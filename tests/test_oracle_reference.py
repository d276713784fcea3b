"""The oracle against the reference's own force field, bit for bit.

The golden fixtures (tests/golden/ref_dmd_*.npz) hold what ForceField_DMD,
_DMD_I and _DMD_II, compiled from the reference's sources, produced for two
500-site binary systems and four 256-site systems with one, three and four types
(one of them with widths in all three regimes and concentrations on the margins):
per-(site, type) widths (the
density-slot transposition of src/ff_dmd_I.cpp:206 vs :376 and the shared phi
slot do not cancel), asymmetric density tables rho_ab != rho_ba, the
dE[jatm+jtype] index of :2336 / :2515, energy_calc's type triangle (:508), and
in the second system widths in all three regimes of :178-180 / :261-310 with
concentrations on the edges of calc_ent's margin (:2230-2237).

Ghost positions, both neighbour lists and crd must be identical, and so must
every array of the force field: the oracle restates the reference's operation
order and both are built with -ffp-contract=off against the same libm, so
they agree bit for bit.  A dropped term, wrong index or transposed slot moves
them by far more than rounding (checked below by perturbing the oracle's
inputs in ways those mistakes would).

Where the reference build is present (this container), the last test also
runs it live on a fresh system and compares again."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from tests import ref_fixture as R
from tests.conftest import rel_inf

PATHS = R.fixture_paths()


def _oracle(d, q, variant):
    from oracle import oracle as O
    t = R.tables(d)
    lat = R.lattice(d)
    kw, skin = R.coef(d)
    o = O.Oracle(nt=len(t.elements), variant=variant)
    o.set_fs(t)
    o.set_params(no_i=q, masses=t.masses, **kw)
    o.set_system(lat.H, lat.x, lat.c)
    o.init(skin)
    o.set_fresh_caches(True)
    return o


def _run_sequence(o, d):
    """the harness's sequence: 2nd list, force_calc(1), energy_calc, c_d_calc, calc_g(0), calc_g(1)"""
    out = {}
    o.create_2nd_neigh_lst()
    out["en_st"] = o.force_calc(1)
    out["f"] = o.get("f")
    out["E_force"] = o.get("E")
    out["dE_force"] = o.get("dE")
    out["energy_calc"] = o.energy_calc()
    o.c_d_calc()
    for k, v in (("c_d", "c_d"), ("n_cd", "n"), ("E_cd", "E"), ("dE_cd", "dE")):
        out[k] = o.get(v)
    beta = float(d["beta"])
    out["cost0"], out["g"] = o.calc_g(0, beta, d["a"])
    out["cost1"], _ = o.calc_g(1, beta, d["a"])
    return out


COMPARED = ("en_st", "f", "E_force", "dE_force", "energy_calc", "c_d", "n_cd", "E_cd", "dE_cd", "cost0", "g",
            "cost1")


@pytest.mark.parametrize("path", PATHS, ids=[os.path.basename(p) for p in PATHS])
@pytest.mark.parametrize("q,variant", R.RUNS)
def test_oracle_matches_reference(path, q, variant):
    d = R.load(path)
    o = _oracle(d, q, variant)
    n = o.natms
    # geometry: ghosts, half list, 2nd list, crd
    xa = o.x_all()
    assert np.array_equal(xa.view(np.uint64), d[f"q{q}_x_all"].view(np.uint64))
    off, nl = o.neighbor_list()
    roff, rnl = R.csr(d, q, "list")
    assert np.array_equal(off, roff) and np.array_equal(nl, rnl)
    out = _run_sequence(o, d)
    off2, nl2 = o.neighbor_list2()
    roff2, rnl2 = R.csr(d, q, "list2")
    assert np.array_equal(off2, roff2) and np.array_equal(nl2, rnl2)
    assert np.array_equal(o.crd(), d[f"q{q}_crd"])
    worst = {}
    for k in COMPARED:
        ref = np.asarray(d[f"{variant}_q{q}_{k}"], dtype=np.float64).reshape(-1)
        got = np.asarray(out[k], dtype=np.float64).reshape(-1)
        assert got.shape == ref.shape, k
        worst[k] = (rel_inf(got, ref), int(np.count_nonzero(got != ref)))
    print(os.path.basename(path), variant, q, {k: f"{e:.1e}/{m}" for k, (e, m) in worst.items()}, "natms", n,
          flush=True)
    for k, (e, m) in worst.items():
        assert m == 0, (k, e, m)


@pytest.mark.parametrize("path", PATHS[:1], ids=[os.path.basename(p) for p in PATHS[:1]])
def test_plausible_mistakes_fail_the_bar(path):
    """what a plausible restatement mistake produces is far from rounding:
    swapping the density slots (a transposed index), one width per site
    instead of one per (site, type) (what would hide the slot transposition),
    another variant's mobility or the corrected dE index of the chemical
    potential moves forces or rates by more than 1e-6"""
    d = R.load(path)
    ref_f = d["DMD_I_q3_f"]
    ref_cd = d["DMD_I_q3_c_d"]
    # transposed density slots: rho[a][b] <-> rho[b][a]
    t = R.tables(d)
    t.rho = np.ascontiguousarray(np.swapaxes(t.rho, 0, 1))
    from oracle import oracle as O
    lat = R.lattice(d)
    kw, skin = R.coef(d)
    o = O.Oracle(nt=2, variant="DMD_I")
    o.set_fs(t)
    o.set_params(no_i=3, masses=t.masses, **kw)
    o.set_system(lat.H, lat.x, lat.c)
    o.init(skin)
    o.force_calc(1)
    assert rel_inf(o.get("f"), ref_f) > 1e-6
    # the DMD_II mobility on DMD_I's data
    o2 = _oracle(d, 3, "DMD_II")
    o2.create_2nd_neigh_lst()
    o2.force_calc(1)
    o2.c_d_calc()
    assert rel_inf(o2.get("c_d"), ref_cd) > 1e-3
    # one width per site instead of one per (site, type)
    lat2 = R.lattice(d)
    lat2.x[:, 4] = lat2.x[:, 3]
    o3 = O.Oracle(nt=2, variant="DMD_I")
    o3.set_fs(R.tables(d))
    o3.set_params(no_i=3, masses=t.masses, **kw)
    o3.set_system(lat2.H, lat2.x, lat2.c)
    o3.init(skin)
    o3.force_calc(1)
    assert rel_inf(o3.get("f"), ref_f) > 1e-6
    # the corrected dE[atom*nt+type] index in place of the reference's dE[jatm+jtype]
    o4 = _oracle(d, 3, "DMD_I")
    o4.set_fix_dE_index(True)
    o4.create_2nd_neigh_lst()
    o4.force_calc(1)
    o4.c_d_calc()
    assert rel_inf(o4.get("c_d"), ref_cd) > 1e-6


HARNESS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref", "ref_harness")


@pytest.mark.skipif(not os.path.exists(HARNESS), reason="reference build absent (make -f oracle/Makefile.ref)")
@pytest.mark.parametrize("variant", ["DMD", "DMD_I", "DMD_II"])
def test_oracle_matches_live_reference(variant):
    """a system the fixtures do not hold: 4^3 cells, another seed, the 5-point
    rule for DMD_I and DMD"""
    import importlib.util
    from mapp_amd import workload as W
    spec = importlib.util.spec_from_file_location("mk", os.path.join(R.GOLDEN, "make_ref_fixtures.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    t = W.johnson_tables(nr=1000, nrho=1000)
    lat = W.fcc_alloy(4, seed=913)
    q = 3 if variant == "DMD_II" else 5
    with tempfile.TemporaryDirectory() as tmp:
        ref = mk.run("live", lat, t, variant, q, tmp)
    d = dict(F=t.F, rho=t.rho, rphi=t.rphi, grid=np.array([t.nrho, t.drho, t.nr, t.dr]), masses=np.array(t.masses),
             beta=np.array(0.37), coef=np.array([3.0, 25.0, 1000.0, 600.0, 0.5]), H=ref["H"], x_read=ref["x_read"],
             c_read=ref["c_read"], a=ref["a"])
    o = _oracle(d, q, variant)
    off, nl = o.neighbor_list()
    assert np.array_equal(off, ref["list_off"].reshape(-1).astype(np.int32))
    assert np.array_equal(nl, ref["list"].reshape(-1).astype(np.int32))
    out = _run_sequence(o, d)
    for k in COMPARED:
        got = np.asarray(out[k], dtype=np.float64).reshape(-1)
        assert np.array_equal(got.view(np.uint64), ref[k].reshape(-1).view(np.uint64)), (variant, k)


@pytest.mark.skipif(not os.path.exists(HARNESS), reason="reference build absent (make -f oracle/Makefile.ref)")
def test_reference_timing_mode():
    """oracle/ref_run.time_force (the CPU baseline of bench.py): two processes
    with a common start, each timing its calls of the reference's force_calc
    on its own copy; their energy is the one the harness dumps in its normal
    sequence for the same system"""
    from mapp_amd import workload as W
    from oracle import ref_run as RR
    t = W.johnson_tables(nr=1000, nrho=1000)
    lat = W.fcc_alloy(3, seed=17)
    with tempfile.TemporaryDirectory() as tmp:
        ts = RR.time_force(tmp, lat, t, 2, 3)
        timed = RR.read_dump(os.path.join(tmp, "rank1.bin"))["en_st"]
    assert ts.shape == (2, 3) and np.all(ts > 0)
    with tempfile.TemporaryDirectory() as tmp:
        cfg, fs, setfl = RR.write_inputs(tmp, "sys", lat, t)
        a_path = os.path.join(tmp, "a.txt")
        np.savetxt(a_path, np.zeros(lat.c.size), fmt="%.17g")
        out = os.path.join(tmp, "full.bin")
        job = os.path.join(tmp, "job")
        RR.write_job(job, cfg, fs, setfl, "DMD_I", 3, out, a_path=a_path)
        subprocess.run([HARNESS, job], check=True, stdout=subprocess.DEVNULL)
        full = RR.read_dump(out)["en_st"]
    assert np.array_equal(timed.view(np.uint64), full.view(np.uint64))


def _md_matches(o, d):
    off, nl = o.neighbor_list()
    assert np.array_equal(off, d["list_off"].reshape(-1).astype(np.int32))
    assert np.array_equal(nl, d["list"].reshape(-1).astype(np.int32))
    assert np.array_equal(o.x_all()[:, :3].view(np.uint64), d["x_all"].view(np.uint64))
    en = np.asarray(o.force_calc(1), dtype=np.float64)
    assert np.array_equal(en.view(np.uint64), d["en_st"].reshape(-1).view(np.uint64))
    f = np.asarray(o.get("f"), dtype=np.float64).reshape(d["f"].shape)
    assert np.array_equal(f.view(np.uint64), d["f"].view(np.uint64))
    u = np.array([o.energy_calc()])
    assert np.array_equal(u.view(np.uint64), d["energy_calc"].reshape(-1).view(np.uint64))


MD_PATHS = R.md_fixture_paths()


@pytest.mark.parametrize("path", MD_PATHS, ids=[os.path.basename(p) for p in MD_PATHS])
def test_oracle_md_matches_reference(path):
    """the oracle's MD-mode loops (ForceField_eam src/ff_eam.cpp:44-356, _lj
    src/ff_lj.cpp:160-341, _fs src/ff_fs.cpp:379-686, _fsm src/ff_fsm.cpp,
    _meam src/ff_meam.cpp) against the reference's build, bit for bit: ghosts,
    half (MEAM: full) list, energy, virial, forces, energy_calc"""
    d = R.load(path)
    _md_matches(R.md_oracle(d), d)


@pytest.mark.skipif(not os.path.exists(HARNESS), reason="reference build absent (make -f oracle/Makefile.ref)")
@pytest.mark.parametrize("variant", ["EAM", "LJ", "FS", "MEAM"])
def test_oracle_md_matches_live_reference(variant):
    """systems the fixtures do not hold: 3^3 cells, other seeds"""
    import importlib.util
    from mapp_amd import workload as W
    spec = importlib.util.spec_from_file_location("mk", os.path.join(R.GOLDEN, "make_ref_fixtures.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    kind, a = {"EAM": ("fcc", 3.6), "LJ": ("fcc", 3.6), "FS": ("bcc", 2.87), "MEAM": ("fcc", 3.58)}[variant]
    H, s, t = W.md_lattice(3, a, kind, seed=915, frac=0.4)
    t = np.asarray(t).astype(np.int64)
    with tempfile.TemporaryDirectory() as tmp:
        d = mk.run_md("live", variant, H, s, t, tmp)
    d.update(variant=np.array(variant), skin=np.array(mk.MD_SKIN), H=H)
    if variant == "EAM":
        tb = mk.md_coefficients("EAM")[2]
        d.update(F=tb.F, rho=tb.rho, rphi=tb.rphi, grid=np.array([tb.nrho, tb.drho, tb.nr, tb.dr]),
                 masses=np.array(tb.masses))
    assert np.array_equal(d["type"].reshape(-1), t)
    _md_matches(R.md_oracle(d), d)
