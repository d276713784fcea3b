"""Writes tests/golden/ref_dmd_*.npz: inputs and outputs of the reference's own
force field classes (ForceField_DMD, _DMD_I, _DMD_II), compiled from
/root/reference/src by oracle/Makefile.ref and driven by oracle/ref_harness.cpp.

Runs only in the build container (where the reference sources and the image's
MPI are); the fixtures it writes are data -- inputs and expected outputs -- and
travel with the tests:

    make -f oracle/Makefile.ref && python tests/golden/make_ref_fixtures.py [case ...]

Each case stores the tables it ran on (Johnson-style Ni-Cu on 2000-point grids
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
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")

COEF = dict(r_crd=3.0, alpha_min=25.0, alpha_max=1000.0, T=600.0, skin=0.5)
# arrays that depend on the rule (through the list cut) but not on the variant
GEOMETRY = ("x_all", "list_off", "list", "list2_off", "list2", "crd")


def write_setfl_style(path, t):
    """the layout ForceField_DMD_I / _DMD read for "FS" (src/ff_dmd_I.cpp:1069-1298):
    3 comments, element line, grid line, one header line, then per element F
    and its density, then r*phi for i >= j, five values per line.  Only its
    grid matters here: the harness replaces the tables it produces."""
    nt = len(t.elements)
    with open(path, "w") as fh:
        fh.write("# setfl-style copy of the same tables\n#\n#\n")
        fh.write(f"{nt} " + " ".join(t.elements) + "\n")
        fh.write(f"{t.nrho} {t.drho:.17g} {t.nr} {t.dr:.17g} {t.rc:.17g}\n")
        fh.write("0 0.0 0.0 fcc\n")
        blocks = []
        for a in range(nt):
            blocks += [t.F[a], t.rho[a, a]]
        blocks += [t.rphi[k] for k in range(nt * (nt + 1) // 2)]
        for b in blocks:
            for s in range(0, b.size, 5):
                fh.write(" ".join(f"{v:.17g}" for v in b[s:s + 5]) + "\n")


def read_dump(path):
    out = {}
    with open(path, "rb") as fh:
        while True:
            head = fh.readline()
            if not head:
                break
            name, rows, cols = head.decode().split()
            n = int(rows) * int(cols)
            out[name] = np.frombuffer(fh.read(8 * n), dtype="<f8").reshape(int(rows), int(cols)).copy()
            fh.read(1)
    return out


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
    from mapp_amd.dmd import write_cfg
    from mapp_amd import workload as W
    cfg = os.path.join(tmp, f"{case}.cfg")
    write_cfg(cfg, lat.H, lat.x, lat.c, tables.masses, tables.elements)
    fs = os.path.join(tmp, "tables.fs")
    W.write_fs(fs, tables)
    setfl = os.path.join(tmp, "tables.setfl")
    write_setfl_style(setfl, tables)
    rng = np.random.default_rng(sum(case.encode()))  # the same a for every variant and rule of a case
    a = rng.uniform(-0.2, 0.2, size=lat.c.shape)
    np.savetxt(os.path.join(tmp, "a.txt"), a.reshape(-1), fmt="%.17g")
    out = os.path.join(tmp, f"{case}_{variant}_{no_i}.bin")
    with open(os.path.join(tmp, "job"), "w") as fh:
        fh.write(f"cfg {cfg}\nfs {fs}\nsetfl {setfl}\nvariant {variant}\n")
        fh.write(f"coef {COEF['r_crd']!r} {no_i} {COEF['alpha_min']!r} {COEF['alpha_max']!r} {COEF['T']!r}\n")
        fh.write(f"skin {COEF['skin']!r}\na {os.path.join(tmp, 'a.txt')}\nbeta 0.37\nout {out}\n")
    subprocess.run([HARNESS, os.path.join(tmp, "job")], check=True, stdout=subprocess.DEVNULL)
    d = read_dump(out)
    d["a"] = a
    return d


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


if __name__ == "__main__":
    main()
