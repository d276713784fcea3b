"""GPU parity of the MD-mode pair potentials through the C ABI against the
oracle on the same seeded inputs: Lennard-Jones (ForceField_lj,
src/ff_lj.cpp) with a different cutoff per type pair on a two-type FCC
lattice, and Finnis-Sinclair (ForceField_fs, src/ff_fs.cpp, and its
squared-cutoff form ForceField_fsm, src/ff_fsm.cpp) on a two-type BCC
lattice.  Bars: neighbour lists (per-type-pair cutoffs) bit-exact; energy,
virial and forces <= 1e-12 relative (infinity norm); energy_calc <= 1e-12."""
import numpy as np
import pytest

from tests.conftest import rel_inf

pytestmark = pytest.mark.gpu


def _compare(ff, orc, label):
    goff, gnl = ff.neighbor_list()
    ooff, onl = orc.neighbor_list()
    assert np.array_equal(goff, ooff) and np.array_equal(gnl, onl), label
    eg, eo = ff.force_calc(1), orc.force_calc(1)
    print(label, "energy", eg[0], eo[0], "pairs", len(gnl), flush=True)
    assert abs(eg[0] - eo[0]) <= 1e-12 * abs(eo[0])
    assert rel_inf(eg[1:], eo[1:]) <= 1e-12
    err = rel_inf(ff.get("f"), orc.get("f"))
    print(label, "force rel", err, flush=True)
    assert err <= 1e-12
    assert abs(ff.energy_calc() - orc.energy_calc()) <= 1e-12 * abs(eo[0])


@pytest.mark.parametrize("n,seed", [(6, 3), (8, 4)])
def test_lj_parity(n, seed):
    from mapp_amd import workload as W
    from mapp_amd.dmd import setup_lj
    from oracle import oracle as O
    H, s, t = W.md_lattice(n, 3.6, "fcc", seed=seed)
    ff = setup_lj(H, s, t, W.LJ_COEFS, nt=2)
    orc = O.Oracle(nt=2, variant="LJ")
    for c in W.LJ_COEFS:
        orc.lj_coef(*c)
    orc.set_system(H, s.copy(), t.reshape(-1, 1))
    orc.init(0.5)
    _compare(ff, orc, f"lj {n}")
    ff.close()


@pytest.mark.parametrize("variant,n,seed", [("FS", 6, 5), ("FS", 8, 6), ("FSM", 6, 9), ("FSM", 8, 10)])
def test_fs_parity(variant, n, seed):
    from mapp_amd import workload as W
    from mapp_amd.dmd import setup_fs
    from oracle import oracle as O
    p = W.fs_params()
    H, s, t = W.md_lattice(n, 2.87, "bcc", seed=seed)
    ff = setup_fs(H, s, t, p, variant=variant)
    orc = O.Oracle(nt=2, variant=variant)
    orc.fs_coef(p)
    orc.set_system(H, s.copy(), t.reshape(-1, 1))
    orc.init(0.5)
    _compare(ff, orc, f"{variant} {n}")
    ff.close()


def test_fs_file_reader_matches_coef(tmp_path):
    """ff_coef FS <file> (elements listed in another order) gives the same force field"""
    from mapp_amd import workload as W
    from mapp_amd.dmd import ForceFieldDMD, setup_fs
    p = W.fs_params(("Fe", "C"))
    path = str(tmp_path / "fs.txt")
    W.write_fs_params(path, p, order=[1, 0])
    H, s, t = W.md_lattice(5, 2.87, "bcc", seed=8)
    a = setup_fs(H, s, t, p)
    b = ForceFieldDMD(no_types=2, variant="FS")
    b.read_fs_params(path, ["Fe", "C"])
    b.set_system(H, s, t.reshape(-1, 1))
    b.init(0.5)
    ea, eb = a.force_calc(1), b.force_calc(1)
    assert np.array_equal(ea, eb)
    assert np.array_equal(a.get("f"), b.get("f"))
    a.close()
    b.close()
