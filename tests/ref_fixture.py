"""Loading the golden outputs of the reference's own force field
(tests/golden/ref_dmd_*.npz, written by tests/golden/make_ref_fixtures.py from
the reference compiled by oracle/Makefile.ref) and rebuilding their inputs.

Each fixture holds the tables, box, positions / widths / concentrations as the
reference read them, calc_g's a and beta, then per rule q (3, 5) the geometry
(q{q}_x_all, q{q}_list_off, q{q}_list, q{q}_list2_off, q{q}_list2, q{q}_crd)
and per variant and rule the results ({variant}_q{q}_en_st, _f, _E_force,
_dE_force, _energy_calc, _c_d, _n_cd, _E_cd, _dE_cd, _cost0, _g, _cost1)."""
import glob
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
# (rule, variant) pairs each fixture holds
RUNS = ((3, "DMD"), (3, "DMD_I"), (3, "DMD_II"), (5, "DMD_II"))


def fixture_paths():
    return sorted(glob.glob(os.path.join(GOLDEN, "ref_dmd_*.npz")))


def md_fixture_paths():
    """the MD-mode force fields (ref_md_*.npz: plain EAM, LJ, FS, FSM, MEAM):
    variant, skin, box, positions and types as read, x_all, list_off, list,
    en_st, f, energy_calc; the EAM case also its tables"""
    return sorted(glob.glob(os.path.join(GOLDEN, "ref_md_*.npz")))


def eam_tables(d):
    from mapp_amd.workload import FSTables
    nrho, drho, nr, dr = d["grid"]
    return FSTables(elements=["Ni", "Cu", "Pd", "Ag"][: d["F"].shape[0]], nrho=int(nrho), drho=float(drho),
                    nr=int(nr), dr=float(dr), F=d["F"], rho=d["rho"], rphi=d["rphi"], masses=list(d["masses"]))


def md_oracle(d):
    """the oracle set up as the fixture's force field (coefficients rebuilt from
    mapp_amd.workload as tests/golden/make_ref_fixtures.py wrote them)"""
    from mapp_amd import workload as W
    from oracle import oracle as O
    variant = str(d["variant"])
    o = O.Oracle(nt=2, variant=variant)
    if variant == "EAM":
        t = eam_tables(d)
        o.set_fs(t)
        o.set_params(masses=t.masses)
    elif variant == "LJ":
        for c in W.LJ_COEFS:
            o.lj_coef(*c)
    elif variant in ("FS", "FSM"):
        o.fs_coef(W.fs_params())
    else:
        o.meam_coef(W.meam_state(["Aa", "Bb"], local=W.meam_local(2)))
    o.set_system(d["H"], np.ascontiguousarray(d["x_read"]).copy(), d["type"].reshape(-1, 1).astype(np.float64))
    o.init(float(d["skin"]))
    return o


def md_gpu(d):
    """the HIP path set up as the fixture's force field"""
    from mapp_amd import workload as W
    from mapp_amd.dmd import setup_eam, setup_fs, setup_lj, setup_meam
    variant, skin = str(d["variant"]), float(d["skin"])
    H, s, t = d["H"], np.ascontiguousarray(d["x_read"]), d["type"].reshape(-1)
    if variant == "EAM":
        return setup_eam(H, s, t, eam_tables(d), skin=skin)
    if variant == "LJ":
        return setup_lj(H, s, t, W.LJ_COEFS, nt=2, skin=skin)
    if variant in ("FS", "FSM"):
        return setup_fs(H, s, t, W.fs_params(), skin=skin, variant=variant)
    return setup_meam(H, s, t, W.meam_state(["Aa", "Bb"], local=W.meam_local(2)), skin=skin)


def load(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def elements(d):
    """the species of a fixture, in type order (the count is the width of c)"""
    return ["Ni", "Cu", "Pd", "Ag"][: d["c_read"].shape[1]]


def tables(d):
    from mapp_amd.workload import FSTables
    nrho, drho, nr, dr = d["grid"]
    return FSTables(elements=elements(d), nrho=int(nrho), drho=float(drho), nr=int(nr), dr=float(dr),
                    F=d["F"], rho=d["rho"], rphi=d["rphi"], masses=list(d["masses"]))


def lattice(d):
    from mapp_amd.workload import Lattice
    x = np.ascontiguousarray(d["x_read"])
    return Lattice(H=np.ascontiguousarray(d["H"]), x=x, c=np.ascontiguousarray(d["c_read"]), cart=None)


def coef(d):
    r_crd, alpha_min, alpha_max, T, skin = (float(v) for v in d["coef"])
    return dict(r_crd=r_crd, alpha_min=alpha_min, alpha_max=alpha_max, T=T), skin


def csr(d, q, name):
    off = d[f"q{q}_{name}_off"]
    return off, d[f"q{q}_{name}"]
