"""Writes tests/golden/ref_dmd_*.npz: inputs and outputs of the reference's own
force field classes (ForceField_DMD, _DMD_I, _DMD_II), compiled from
/root/reference/src by oracle/Makefile.ref and driven by oracle/ref_harness.cpp.

Runs only in the build container (where the reference sources and the image's
MPI are); the fixtures it writes are data -- inputs and expected outputs -- and
travel with the tests:

    make -f oracle/Makefile.ref && python tests/golden/make_ref_fixtures.py [case ...]

ref_md_*.npz hold the same for the MD-mode force fields (plain EAM reading the
FS file, LJ, FS, FSM, MEAM): variant, box, positions and types as read, ghosts,
the half (MEAM: full) list, force_calc's energy / virial / f and energy_calc;
the EAM case also its tables (the others are rebuilt from workload.py).

Each DMD case stores the tables it ran on (Johnson-style Ni-Cu on 2000-point grids
for the two binary cases; Ni-Cu-Pd and Ni-Cu-Pd-Ag on 1000-point grids for the
three- and four-type cases; the FS file's values), the box, the positions / widths / concentrations as
the reference read them from the extended CFG, calc_g's a and beta, and per
variant every array the harness dumps: ghosts' positions, the half list and
the 2nd list, crd, force_calc's energy / virial / f / E / dE, energy_calc,
c_d_calc's c_d / n / E / dE, calc_g's cost and g (chk 0) and cost (chk 1).
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from oracle.ref_run import COEF, HARNESS, read_dump, write_cfg_md, write_inputs, write_job, write_md_job  # noqa: E402

# arrays that depend on the rule (through the list cut) but not on the variant
GEOMETRY = ("x_all", "list_off", "list", "list2_off", "list2", "crd")


def lattice_cases():
    """case name -> (lattice, element names, table grid points)"""
    from mapp_amd import workload as W
    cases = {}
    # 5^3 FCC Ni-Cu (500 sites): alpha per (site, type) ~ U[60, 120]
    cases["fcc5"] = (W.fcc_alloy(5, seed=101), ("Ni", "Cu"), 2000)
    # widths from three bands so that pair widths fall in every regime of
    # src/ff_dmd_I.cpp:178-180, :261-310, and concentrations on the edges of
    # calc_ent's margin
    lat = W.fcc_alloy(5, seed=102)
    rng = np.random.default_rng(202)
    band = rng.integers(0, 3, size=lat.c.shape)
    lat.x[:, 3:] = rng.uniform(np.array([12.0, 60.0, 2000.0])[band], np.array([40.0, 120.0, 4000.0])[band])
    edge = np.array([0.0, 1.0, 5e-9, 1.0 - 5e-9, 1e-8, 1.0 - 1e-8, 2e-8])
    pick = rng.integers(0, 3, size=lat.natms) == 0
    cA = lat.c[:, 0].copy()
    cA[pick] = edge[np.arange(pick.sum()) % len(edge)]
    lat.c[:, 0] = cA
    lat.c[:, 1] = 1.0 - cA
    cases["regimes5"] = (lat, ("Ni", "Cu"), 2000)
    # 4^3 FCC with three and four types (256 sites): concentrations drawn flat
    # over the simplex, every density slot rho_ab distinct
    cases["ternary4"] = (W.fcc_alloy(4, nt=3, seed=103), ("Ni", "Cu", "Pd"), 1000)
    cases["quaternary4"] = (W.fcc_alloy(4, nt=4, seed=104), ("Ni", "Cu", "Pd", "Ag"), 1000)
    # one type (pure Ni, c ~ U[0.3, 1]: the vacancy concentration is 1 - c)
    cases["unary4"] = (W.fcc_alloy(4, nt=1, seed=106), ("Ni",), 1000)
    # three types with widths in the three bands and concentration vectors on
    # calc_ent's margins (types that are absent, nearly absent or nearly pure)
    lat = W.fcc_alloy(4, nt=3, seed=105)
    rng = np.random.default_rng(205)
    band = rng.integers(0, 3, size=lat.c.shape)
    lat.x[:, 3:] = rng.uniform(np.array([12.0, 60.0, 2000.0])[band], np.array([40.0, 120.0, 4000.0])[band])
    edges = np.array([[1.0, 0.0, 0.0], [0.0, 0.0, 1.0], [1.0 - 5e-9, 5e-9, 0.0], [1e-8, 1.0 - 2e-8, 1e-8],
                      [2e-8, 0.5, 0.5 - 2e-8], [0.0, 1.0 - 1e-8, 1e-8], [1.0 / 3, 1.0 / 3, 1.0 / 3]])
    pick = np.flatnonzero(rng.integers(0, 3, size=lat.natms) == 0)
    lat.c[pick] = edges[np.arange(pick.size) % len(edges)]
    cases["regimes_ternary4"] = (lat, ("Ni", "Cu", "Pd"), 1000)
    return cases


def run(case, lat, tables, variant, no_i, tmp):
    cfg, fs, setfl = write_inputs(tmp, case, lat, tables)
    rng = np.random.default_rng(sum(case.encode()))  # the same a for every variant and rule of a case
    a = rng.uniform(-0.2, 0.2, size=lat.c.shape)
    a_path = os.path.join(tmp, "a.txt")
    np.savetxt(a_path, a.reshape(-1), fmt="%.17g")
    out = os.path.join(tmp, f"{case}_{variant}_{no_i}.bin")
    job = os.path.join(tmp, "job")
    write_job(job, cfg, fs, setfl, variant, no_i, out, a_path=a_path)
    subprocess.run([HARNESS, job], check=True, stdout=subprocess.DEVNULL)
    d = read_dump(out)
    d["a"] = a
    return d


MD_SKIN = 0.5


def md_cases():
    """case name -> (variant, box, reduced positions, integer types) for the
    MD-mode force fields"""
    from mapp_amd import workload as W
    cases = {}
    lat = W.fcc_alloy(4, seed=21)
    cases["eam_fcc4"] = ("EAM", lat.H, np.ascontiguousarray(lat.x[:, :3]), np.arange(lat.natms) % 3 == 0)
    cases["lj_fcc4"] = ("LJ",) + W.md_lattice(4, 3.6, "fcc", seed=3)
    H, s, t = W.md_lattice(5, 2.87, "bcc", seed=5)
    cases["fs_bcc5"] = ("FS", H, s, t)
    cases["fsm_bcc5"] = ("FSM", H, s, t)
    cases["meam_fcc4"] = ("MEAM",) + W.md_lattice(4, 3.58, "fcc", nt=2, seed=41, jitter=0.05)
    return {k: (v, H, s, np.asarray(t).astype(np.int64)) for k, (v, H, s, t) in cases.items()}


def md_coefficients(variant):
    """element names, masses and the coefficient inputs of each MD variant, as
    the tests rebuild them: EAM the Johnson tables on 2000-point grids, LJ
    workload.LJ_COEFS, FS / FSM workload.fs_params(), MEAM workload.meam_state
    of Aa and Bb with workload.meam_local(2)"""
    from mapp_amd import workload as W
    if variant == "EAM":
        t = W.johnson_tables(nr=2000, nrho=2000)
        return list(t.elements), list(t.masses), t
    if variant == "LJ":
        return ["A", "B"], [1.0, 1.0], W.LJ_COEFS
    if variant in ("FS", "FSM"):
        p = W.fs_params()
        return list(p.elements), [1.0, 1.0], p
    names = ["Aa", "Bb"]
    return names, [W.MEAM_LIB[e]["atwt"] for e in names], W.meam_local(2)


def run_md(case, variant, H, s, types, tmp):
    from mapp_amd import workload as W
    names, masses, co = md_coefficients(variant)
    cfg = os.path.join(tmp, f"{case}.cfg")
    write_cfg_md(cfg, H, s, types, masses, names)
    kw = {}
    if variant == "EAM":
        kw["fs"] = os.path.join(tmp, "tables.fs")
        W.write_fs(kw["fs"], co)
    elif variant == "LJ":
        kw["lj"] = [v for a, b, eps, sig, cut in co for v in (names[a], names[b], repr(eps), repr(sig), repr(cut))]
    elif variant in ("FS", "FSM"):
        kw["params"] = os.path.join(tmp, "fs.params")
        W.write_fs_params(kw["params"], co)
    else:
        kw["lib"] = os.path.join(tmp, "meam.library")
        W.write_meam_lib(kw["lib"], names)
        kw["local"] = os.path.join(tmp, "meam.local")
        W.write_meam_local(kw["local"], co)
    out = os.path.join(tmp, f"{case}.bin")
    job = os.path.join(tmp, "job")
    write_md_job(job, cfg, variant, out, MD_SKIN, **kw)
    subprocess.run([HARNESS, job], check=True, stdout=subprocess.DEVNULL)
    return read_dump(out)


def main():
    from mapp_amd import workload as W
    if not os.path.exists(HARNESS):
        sys.exit("build the harness first: make -f oracle/Makefile.ref")
    wanted = set(sys.argv[1:])
    for case, (lat, elements, ngrid) in lattice_cases().items():
        if wanted and case not in wanted:
            continue
        tables = W.johnson_tables(elements=elements, nr=ngrid, nrho=ngrid)
        arrays = dict(F=tables.F, rho=tables.rho, rphi=tables.rphi,
                      grid=np.array([tables.nrho, tables.drho, tables.nr, tables.dr]),
                      masses=np.array(tables.masses), beta=np.array(0.37),
                      coef=np.array([COEF["r_crd"], COEF["alpha_min"], COEF["alpha_max"], COEF["T"], COEF["skin"]]))
        # the three mobility variants with the 3-point rule; the 5-point rule
        # (a different list cut and rule) with DMD_II's
        for no_i, variant in ((3, "DMD"), (3, "DMD_I"), (3, "DMD_II"), (5, "DMD_II")):
            with tempfile.TemporaryDirectory() as tmp:
                d = run(case, lat, tables, variant, no_i, tmp)
            for k in ("H", "x_read", "c_read", "a"):
                if k in arrays:
                    assert np.array_equal(arrays[k], d[k]), (case, k)
                else:
                    arrays[k] = d[k]
            for k, v in d.items():
                if k in ("H", "x_read", "c_read", "a"):
                    continue
                if k in GEOMETRY:  # the same for every variant of a rule: stored once
                    v = v.astype(np.int32).reshape(-1) if k != "x_all" else v
                    key = f"q{no_i}_{k}"
                    if key in arrays:
                        assert np.array_equal(arrays[key], v), (case, variant, key)
                    arrays[key] = v
                else:
                    arrays[f"{variant}_q{no_i}_{k}"] = v
            print(case, variant, no_i, "energy", d["en_st"][0, 0], "natms", d["x_read"].shape[0], flush=True)
        path = os.path.join(HERE, f"ref_dmd_{case}.npz")
        np.savez_compressed(path, **arrays)
        print("wrote", path, os.path.getsize(path), "bytes")
    for case, (variant, H, s, types) in md_cases().items():
        if wanted and case not in wanted:
            continue
        with tempfile.TemporaryDirectory() as tmp:
            d = run_md(case, variant, H, s, types, tmp)
        assert np.array_equal(d["x_read"], s) and np.array_equal(d["type"].reshape(-1), types), case
        arrays = dict(variant=np.array(variant), skin=np.array(MD_SKIN), **d)
        if variant == "EAM":
            t = md_coefficients(variant)[2]
            arrays.update(F=t.F, rho=t.rho, rphi=t.rphi, grid=np.array([t.nrho, t.drho, t.nr, t.dr]),
                          masses=np.array(t.masses))
        path = os.path.join(HERE, f"ref_md_{case}.npz")
        np.savez_compressed(path, **arrays)
        print(case, variant, "energy", d["en_st"][0, 0], "natms", d["x_read"].shape[0], flush=True)
        print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
