"""Writes tests/golden/ref_dmd_*.npz: inputs and outputs of the reference's own
force field classes (ForceField_DMD, _DMD_I, _DMD_II), compiled from
/root/reference/src by oracle/Makefile.ref and driven by oracle/ref_harness.cpp.

Runs only in the build container (where the reference sources and the image's
MPI are); the fixtures it writes are data -- inputs and expected outputs -- and
travel with the tests:

    make -f oracle/Makefile.ref && python tests/golden/make_ref_fixtures.py [case ...]

ref_eam_*.npz hold the same for the plain EAM force field in MD mode
(ForceField_eam reading the FS file): tables, box, positions and types as
read, ghosts, the half list, force_calc's energy / virial / f and energy_calc.

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
from oracle.ref_run import COEF, HARNESS, read_dump, write_cfg_md, write_inputs, write_job  # noqa: E402

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


def eam_cases():
    """case name -> (box, reduced positions, integer types, table grid points)
    for the plain EAM force field in MD mode"""
    from mapp_amd import workload as W
    lat = W.fcc_alloy(4, seed=21)
    types = (np.arange(lat.natms) % 3 == 0).astype(np.int64)
    return {"eam_fcc4": (lat.H, np.ascontiguousarray(lat.x[:, :3]), types, 2000)}


def run_eam(case, H, s, types, tables, tmp):
    import mapp_amd.workload as W
    cfg = os.path.join(tmp, f"{case}.cfg")
    write_cfg_md(cfg, H, s, types, tables.masses, tables.elements)
    fs = os.path.join(tmp, "tables.fs")
    W.write_fs(fs, tables)
    out = os.path.join(tmp, f"{case}.bin")
    job = os.path.join(tmp, "job")
    write_job(job, cfg, fs, fs, "EAM", 1, out)
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
    for case, (H, s, types, ngrid) in eam_cases().items():
        if wanted and case not in wanted:
            continue
        tables = W.johnson_tables(nr=ngrid, nrho=ngrid)
        with tempfile.TemporaryDirectory() as tmp:
            d = run_eam(case, H, s, types, tables, tmp)
        assert np.array_equal(d["x_read"], s) and np.array_equal(d["type"].reshape(-1), types), case
        arrays = dict(F=tables.F, rho=tables.rho, rphi=tables.rphi,
                      grid=np.array([tables.nrho, tables.drho, tables.nr, tables.dr]),
                      masses=np.array(tables.masses), skin=np.array(COEF["skin"]), **d)
        path = os.path.join(HERE, f"ref_{case}.npz")
        np.savez_compressed(path, **arrays)
        print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
