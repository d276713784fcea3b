"""Synthetic DMD inputs: a jittered binary FCC alloy and a Johnson-style
analytic EAM for Ni-Cu tabulated in the Finnis-Sinclair file layout that
ForceField_DMD_II::set_fs reads (src/ff_dmd_II.cpp:1253-1480).

Everything is seeded and vectorised with numpy; the same arrays feed the GPU
library and the CPU oracle.
"""
from dataclasses import dataclass, field

import numpy as np

# Johnson-style parameters (r_e in Angstrom, energies in eV); densities are
# f_e exp(-beta (r/r_e - 1)), pairs phi_e exp(-gamma (r/r_e - 1)), embedding
# F = -E_c (1 - (a/b) ln x) x^(a/b) - 6 phi_e x^(g/b), x = rho / rho_e.
ELEMENTS = {
    "Ni": dict(mass=58.6934, r_e=2.49, f_e=1.25, beta=6.10, phi_e=0.62, gamma=8.80, E_c=4.45, a=4.98),
    "Cu": dict(mass=63.546, r_e=2.556, f_e=1.00, beta=5.85, phi_e=0.59, gamma=8.00, E_c=3.54, a=5.09),
    # third and fourth species for the three- and four-type alloys (same functional form)
    "Pd": dict(mass=106.42, r_e=2.52, f_e=1.10, beta=6.00, phi_e=0.60, gamma=8.40, E_c=3.91, a=5.20),
    "Ag": dict(mass=107.8682, r_e=2.58, f_e=0.90, beta=5.95, phi_e=0.55, gamma=8.60, E_c=2.85, a=5.40),
}


def _taper(r, rc, h=0.5):
    x = np.clip((rc - r) / h, 0.0, None)
    x4 = x ** 4
    return x4 / (1.0 + x4)


@dataclass
class FSTables:
    elements: list
    nrho: int
    drho: float
    nr: int
    dr: float
    F: np.ndarray     # [nt][nrho]
    rho: np.ndarray   # [nt][nt][nr]: density at an a site from a b neighbour
    rphi: np.ndarray  # [nt(nt+1)/2][nr]: r*phi_ij, i >= j
    masses: list = field(default_factory=list)

    @property
    def rc(self):
        return (self.nr - 1.0) * self.dr


def johnson_tables(elements=("Ni", "Cu"), nr=10000, nrho=10000, rc=5.0, rho_max=40.0, symmetric_density=False):
    """Tabulate the alloy on nr/nrho-point grids with a smooth cutoff at rc.

    rho[a][b] (density at a site of type a due to a neighbour of type b) is
    the density function of b; symmetric_density=True uses the mean of the
    two functions instead, so rho[a][b] == rho[b][a]."""
    nt = len(elements)
    dr = rc / (nr - 1)
    drho = rho_max / (nrho - 1)
    r = np.arange(nr, dtype=np.float64) * dr
    rho_g = np.arange(nrho, dtype=np.float64) * drho
    tp = _taper(r, rc)
    par = [ELEMENTS[e] for e in elements]
    dens = [p["f_e"] * np.exp(-p["beta"] * (r / p["r_e"] - 1.0)) * tp for p in par]
    pair = [p["phi_e"] * np.exp(-p["gamma"] * (r / p["r_e"] - 1.0)) * tp for p in par]
    F = np.zeros((nt, nrho))
    for a, p in enumerate(par):
        rho_e = 12.0 * p["f_e"]
        x = rho_g / rho_e
        with np.errstate(divide="ignore", invalid="ignore"):
            k = p["a"] / p["beta"]
            val = -p["E_c"] * (1.0 - k * np.log(x)) * x ** k - 6.0 * p["phi_e"] * x ** (p["gamma"] / p["beta"])
        val[0] = 0.0
        F[a] = val
    rho = np.zeros((nt, nt, nr))
    for a in range(nt):
        for b in range(nt):
            rho[a, b] = 0.5 * (dens[a] + dens[b]) if symmetric_density else dens[b]
    rphi = np.zeros((nt * (nt + 1) // 2, nr))
    k = 0
    for i in range(nt):
        for j in range(i + 1):
            if i == j:
                phi = pair[i]
            else:
                with np.errstate(divide="ignore", invalid="ignore"):
                    fi, fj = dens[i], dens[j]
                    phi = 0.5 * (np.where(fi > 0, fj / fi, 0.0) * pair[i] + np.where(fj > 0, fi / fj, 0.0) * pair[j])
            rphi[k] = r * phi
            k += 1
    return FSTables(list(elements), nrho, drho, nr, dr, F, rho, rphi, [p["mass"] for p in par])


def write_fs(path, t: FSTables):
    """FS layout as read by the reference: 3 comments, element line, grid line,
    one element header line, then a flat stream of values (%.17g)."""
    nt = len(t.elements)
    with open(path, "w") as fh:
        fh.write("# Johnson-style analytic EAM, tabulated\n# synthetic Ni-Cu alloy\n# FS layout\n")
        fh.write(f"{nt} " + " ".join(t.elements) + "\n")
        fh.write(f"{t.nrho} {t.drho:.17g} {t.nr} {t.dr:.17g} {t.rc:.17g}\n")
        fh.write("0 0.0 0.0 fcc\n")
        vals = []
        for a in range(nt):
            vals.append(t.F[a])
            for b in range(nt):
                vals.append(t.rho[a, b])
        vals.append(t.rphi.reshape(-1))
        flat = np.concatenate([np.asarray(v).reshape(-1) for v in vals])
        for s in range(0, flat.size, 5):
            fh.write(" ".join(f"{v:.17g}" for v in flat[s:s + 5]) + "\n")


@dataclass
class Lattice:
    H: np.ndarray        # 3x3 lower triangular, rows are box vectors
    x: np.ndarray        # [natms][3+nt]: reduced coords then alpha
    c: np.ndarray        # [natms][nt]
    cart: np.ndarray     # [natms][3] jittered cartesian positions

    @property
    def natms(self):
        return self.x.shape[0]


def m3inv(H):
    """B = H^-1 by the cofactor formula of M3INV (src/init.h:72-85)."""
    A = H
    B = np.zeros((3, 3))
    B[0, 0] = A[1, 1] * A[2, 2] - A[1, 2] * A[2, 1]
    B[1, 1] = A[2, 2] * A[0, 0] - A[2, 0] * A[0, 2]
    B[2, 2] = A[0, 0] * A[1, 1] - A[0, 1] * A[1, 0]
    B[1, 0] = A[1, 2] * A[2, 0] - A[1, 0] * A[2, 2]
    B[2, 1] = A[2, 0] * A[0, 1] - A[2, 1] * A[0, 0]
    B[0, 2] = A[0, 1] * A[1, 2] - A[0, 2] * A[1, 1]
    B[2, 0] = A[1, 0] * A[2, 1] - A[2, 0] * A[1, 1]
    B[0, 1] = A[2, 1] * A[0, 2] - A[0, 1] * A[2, 2]
    B[1, 2] = A[0, 2] * A[1, 0] - A[1, 2] * A[0, 0]
    det = A[0, 0] * B[0, 0] + A[0, 1] * B[1, 0] + A[0, 2] * B[2, 0]
    return B / det


def x2s(cart, B):
    """Atoms::x2s with the wrap into [0,1) (src/atoms.cpp:1869-1897)."""
    s = np.empty_like(cart)
    s[:, 0] = cart[:, 0] * B[0, 0]
    s[:, 0] += B[1, 0] * cart[:, 1]
    s[:, 0] += B[2, 0] * cart[:, 2]
    s[:, 1] = cart[:, 1] * B[1, 1]
    s[:, 1] += B[2, 1] * cart[:, 2]
    s[:, 2] = cart[:, 2] * B[2, 2]
    for _ in range(4):
        s = np.where(s < 0.0, s + 1.0, s)
        s = np.where(s >= 1.0, s - 1.0, s)
    return s


def fcc_alloy(n, a=3.6, nt=2, seed=1, jitter=0.03, alpha_lo=60.0, alpha_hi=120.0, c_per_site=None,
              alpha_per_site=False):
    """n^3 FCC cells of lattice constant a; positions jittered N(0, jitter);
    alpha ~ U[alpha_lo, alpha_hi] per (site, type) (one value per site, shared
    by the types, when alpha_per_site); c_A ~ U[0,1], c_B = 1 - c_A for two
    types, a flat Dirichlet draw over the types for three or four."""
    rng = np.random.default_rng(seed)
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    g = np.arange(n)
    cells = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3)
    s0 = ((cells[:, None, :] + basis[None, :, :]) / n).reshape(-1, 3)
    H = np.diag([a * n] * 3).astype(np.float64)
    cart = s0 @ H + rng.normal(0.0, jitter, size=s0.shape)
    B = m3inv(H)
    s = x2s(cart, B)
    natms = s.shape[0]
    alpha = rng.uniform(alpha_lo, alpha_hi, size=(natms, nt))
    if alpha_per_site:
        alpha[:, 1:] = alpha[:, :1]
    if nt == 1:
        c = rng.uniform(0.3, 1.0, size=(natms, 1)) if c_per_site is None else np.full((natms, 1), c_per_site)
    elif nt == 2:
        cA = rng.uniform(0.0, 1.0, size=natms)
        c = np.zeros((natms, nt))
        c[:, 0] = cA
        c[:, 1] = 1.0 - cA
    else:
        c = rng.dirichlet(np.ones(nt), size=natms)
    x = np.concatenate([s, alpha], axis=1)
    return Lattice(H=H, x=np.ascontiguousarray(x), c=np.ascontiguousarray(c), cart=cart)


@dataclass
class FSParams:
    """Finnis-Sinclair coefficients of ForceField_fs (type order): A[nt],
    t1[nt][nt], t2[nt][nt] (mat_t_1[i][j]), and per type pair in COMP order
    (i <= j: j(j+1)/2 + i) cut_rho, cut_phi, k1, k2, k3."""
    elements: list
    A: np.ndarray
    t1: np.ndarray
    t2: np.ndarray
    cut_rho: np.ndarray
    cut_phi: np.ndarray
    k1: np.ndarray
    k2: np.ndarray
    k3: np.ndarray


def comp(i, j):
    """COMP(i, j): the unordered type pair index"""
    return (j + 1) * j // 2 + i if i < j else (i + 1) * i // 2 + j


def fs_params(elements=("A", "B")):
    """synthetic two-type Finnis-Sinclair coefficients for a BCC lattice of
    a = 2.87: densities d^2 (t1 + t2 d), d = r - cut_rho, positive over the
    first two shells; pair terms d^2 (k1 + k2 r + k3 r^2) with a minimum near
    the first shell"""
    nt = len(elements)
    A = np.array([1.8, 2.3, 2.0, 1.6][:nt])
    t1 = np.array([[1.0, 0.8, 0.9, 0.7], [0.6, 1.2, 0.5, 0.8], [0.9, 0.7, 1.1, 0.6], [0.8, 0.6, 0.7, 1.0]])[:nt, :nt]
    t2 = np.array([[0.3, -0.2, 0.1, 0.2], [0.25, -0.1, 0.2, 0.1], [0.2, 0.1, -0.3, 0.1], [0.1, 0.2, 0.1, -0.2]])[:nt, :nt]
    nc = nt * (nt + 1) // 2
    cut_rho = np.array([3.5, 3.3, 3.1, 3.4, 3.2, 3.0, 3.3, 3.2, 3.1, 3.0])[:nc]
    cut_phi = np.array([3.4, 3.2, 3.0, 3.3, 3.1, 2.9, 3.2, 3.1, 3.0, 2.9])[:nc]
    k1 = np.array([1.3, 1.1, 0.9, 1.2, 1.0, 0.8, 1.1, 1.0, 0.9, 0.8])[:nc]
    k2 = np.array([-0.45, -0.38, -0.30, -0.40, -0.35, -0.28, -0.4, -0.36, -0.3, -0.27])[:nc]
    k3 = np.array([0.035, 0.03, 0.025, 0.032, 0.028, 0.022, 0.03, 0.028, 0.025, 0.022])[:nc]
    return FSParams(elements=list(elements), A=A, t1=t1, t2=t2, cut_rho=cut_rho, cut_phi=cut_phi, k1=k1, k2=k2, k3=k3)


def write_fs_params(path, p: FSParams, order=None):
    """the parameter file of ForceField_fs::coef (element line, A lines, t1 t2
    lines, cut_rho cut_phi k1 k2 k3 lines) with %.17g, in the element order
    `order` (indices into p.elements; default the type order)"""
    nt = len(p.elements)
    order = list(range(nt)) if order is None else list(order)
    with open(path, "w") as fh:
        fh.write("# Finnis-Sinclair parameters\n")
        fh.write(" ".join(p.elements[k] for k in order) + "\n")
        fh.write("# A\n")
        for k in order:
            fh.write("%.17g\n" % p.A[k])
        fh.write("# t1 t2\n")
        for a in order:
            for b in order:
                fh.write("%.17g %.17g\n" % (p.t1[a][b], p.t2[a][b]))
        fh.write("# cut_rho cut_phi k1 k2 k3\n")
        for ia in range(nt):
            for ib in range(ia, nt):
                c = comp(order[ia], order[ib])
                fh.write("%.17g %.17g %.17g %.17g %.17g\n" % (p.cut_rho[c], p.cut_phi[c], p.k1[c], p.k2[c], p.k3[c]))


# Lennard-Jones coefficients of a two-type test alloy on the FCC a = 3.6
# lattice: (itype, jtype, epsilon, sigma, cutoff), one cutoff per type pair,
# each between two neighbour shells (3.6, 4.41, 5.09, 5.69)
LJ_COEFS = [(0, 0, 0.50, 2.30, 4.75), (0, 1, 0.40, 2.35, 4.0), (1, 1, 0.60, 2.40, 5.4)]


def md_lattice(n, a, kind="fcc", nt=2, frac=0.3, seed=1, jitter=0.03):
    """n^3 cubic cells of an FCC, BCC or diamond lattice, jittered N(0, jitter); type 1
    on a fraction frac of the sites (nt = 2).  Returns H, reduced positions
    [natms][3] and the types."""
    rng = np.random.default_rng(seed)
    fcc = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    basis = {"fcc": fcc, "bcc": np.array([[0, 0, 0], [0.5, 0.5, 0.5]]),
             "dia": np.concatenate([fcc, fcc + 0.25])}[kind]
    g = np.arange(n)
    cells = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3)
    s0 = ((cells[:, None, :] + basis[None, :, :]) / n).reshape(-1, 3)
    H = np.diag([a * n] * 3).astype(np.float64)
    cart = s0 @ H + rng.normal(0.0, jitter, size=s0.shape)
    s = x2s(cart, m3inv(H))
    types = np.zeros(s.shape[0])
    if nt > 1:
        types[rng.random(s.shape[0]) < frac] = 1.0
    return H, np.ascontiguousarray(s), types


# defaults of the DMD workloads (no_i = 3 or 5 points, alpha_min/max, T, skin)
DEFAULTS = dict(r_crd=3.0, alpha_min=25.0, alpha_max=1000.0, T=600.0, skin=0.5)


# ---------------------------------------------------------------- MEAM
# Synthetic two-element MEAM parameters (values of the usual magnitude for fcc
# metals, made up for the tests): the library-file entry of each element in
# the order of ForceField_meam::read_global (src/ff_meam.cpp:2745-2790) and
# the local-file settings of read_local (:2858-3191).
MEAM_LIB_KEYS = ("lat", "z", "ielt", "atwt", "alpha", "b0", "b1", "b2", "b3", "alat", "esub", "asub",
                 "t0", "t1", "t2", "t3", "rozero", "ibar")
MEAM_LATTICES = ("fcc", "bcc", "hcp", "dim", "dia", "b1", "c11", "l12", "b2")

MEAM_LIB = {
    "Aa": dict(lat="fcc", z=12.0, ielt=28, atwt=58.7, alpha=5.0, b0=2.5, b1=2.1, b2=5.0, b3=2.5, alat=3.55,
               esub=4.4, asub=1.0, t0=1.0, t1=3.0, t2=1.5, t3=3.0, rozero=1.0, ibar=3),
    "Bb": dict(lat="fcc", z=12.0, ielt=29, atwt=63.5, alpha=5.1, b0=3.4, b1=2.0, b2=4.2, b3=3.0, alat=3.62,
               esub=3.5, asub=0.95, t0=1.0, t1=2.6, t2=2.8, t3=2.1, rozero=1.1, ibar=3),
    "Cc": dict(lat="bcc", z=8.0, ielt=26, atwt=55.8, alpha=5.2, b0=3.6, b1=1.0, b2=1.0, b3=1.0, alat=2.86,
               esub=4.3, asub=0.56, t0=1.0, t1=1.5, t2=2.0, t3=-1.2, rozero=1.0, ibar=3),
    "Dd": dict(lat="dia", z=4.0, ielt=14, atwt=28.1, alpha=4.9, b0=4.4, b1=5.5, b2=5.5, b3=5.5, alat=5.43,
               esub=4.6, asub=1.0, t0=1.0, t1=3.3, t2=5.1, t3=-0.8, rozero=1.0, ibar=1),
}


def meam_local(nt, **over):
    """local-file settings as (keyword, indices..., value) tuples, 1-based
    element indices of the library order"""
    loc = [("rc", 4.5), ("delr", 0.1), ("augt1", 1), ("ialloy", 0), ("emb_lin_neg", 0), ("erose_form", 0)]
    if nt > 1:
        loc += [("lattce", 1, 2, "l12"), ("Ec", 1, 2, 4.15), ("re", 1, 2, 2.52), ("alpha", 1, 2, 5.05),
                ("Cmin", 1, 1, 2, 1.8), ("Cmin", 2, 2, 1, 1.9), ("Cmax", 1, 2, 1, 2.75)]
    keyed = {}
    for e in loc:
        keyed[e[:-1]] = e
    for k, v in over.items():
        keyed[(k,)] = (k, v)
    return list(keyed.values())


def write_meam_lib(path, elements, lib=None):
    """a library file: per element the quoted name and lattice, then 16 numbers"""
    lib = lib or MEAM_LIB
    with open(path, "w") as fh:
        fh.write("# MEAM library (synthetic parameters)\n")
        fh.write("# elt lat z ielement atwt\n# alpha b0 b1 b2 b3 alat esub asub\n# t0 t1 t2 t3 rozero ibar\n")
        for el in elements:
            e = lib[el]
            fh.write(f"'{el}' '{e['lat']}' {e['z']!r} {e['ielt']} {e['atwt']!r}\n")
            fh.write(" ".join(repr(float(e[k])) for k in ("alpha", "b0", "b1", "b2", "b3", "alat", "esub",
                                                          "asub")) + "\n")
            fh.write(" ".join(repr(float(e[k])) for k in ("t0", "t1", "t2", "t3", "rozero")) + f" {e['ibar']}\n")


def write_meam_local(path, local):
    """a local parameter file in the keyword(i,j) = value form"""
    with open(path, "w") as fh:
        fh.write("# local MEAM settings\n")
        for e in local:
            if len(e) == 2:
                fh.write(f"{e[0]} = {e[1]}\n")
            else:
                idx = ",".join(str(i) for i in e[1:-1])
                v = e[-1]
                fh.write(f"{e[0]}({idx}) = " + (f"'{v}'" if isinstance(v, str) else repr(v)) + "\n")


def meam_state(types, lib_order=None, local=(), lib=None):
    """The members ForceField_meam holds after read_global and read_local for
    atom types `types` (element names), a library listing `lib_order` (default:
    types) and local settings `local`: a dict of numpy arrays in type order,
    the layout of the parameter structs of the product and the oracle."""
    lib = lib or MEAM_LIB
    lib_order = list(lib_order or types)
    nt = len(types)
    M = 4
    st = dict(nt=nt)
    for k in ("Z", "A", "rho0", "beta0", "beta1", "beta2", "beta3", "t0", "t1", "t2", "t3"):
        st[k] = np.zeros(M)
    st["ielt"] = np.zeros(M, dtype=np.int32)
    st["ibar"] = np.zeros(M, dtype=np.int32)
    st["lattice"] = -np.ones((M, M), dtype=np.int32)
    for k in ("nn2", "zbl"):
        st[k] = np.zeros((M, M), dtype=np.int32)
    for k in ("Ec", "alpha", "re", "delta", "attrac", "repuls", "ebound"):
        st[k] = np.zeros((M, M))
    st["cmin"] = np.zeros((M, M, M))
    st["cmax"] = np.zeros((M, M, M))
    type_ref = []  # library position -> type
    for el in lib_order:
        if el not in types:
            continue
        t = types.index(el)
        type_ref.append(t)
        e = lib[el]
        lat = MEAM_LATTICES.index(e["lat"])
        st["lattice"][t, t] = lat
        st["Z"][t] = e["z"]
        st["ielt"][t] = e["ielt"]
        st["alpha"][t, t] = e["alpha"]
        for q in range(4):
            st[f"beta{q}"][t] = e[f"b{q}"]
        al = float(e["alat"])  # nearest-neighbour distance of the element's lattice (:2766-2776)
        st["re"][t, t] = {0: lambda: al / np.sqrt(2.0), 1: lambda: al * 0.5 * np.sqrt(3.0), 2: lambda: al,
                          3: lambda: al, 4: lambda: al * 0.25 * np.sqrt(3.0)}[lat]()
        st["Ec"][t, t] = e["esub"]
        st["A"][t] = e["asub"]
        for q in range(4):
            st[f"t{q}"][t] = e[f"t{q}"]
        st["rho0"][t] = e["rozero"]
        st["ibar"][t] = e["ibar"]
    st["cmin"][:] = 2.0
    st["cmax"][:] = 2.8
    st["ebound"][:] = 1.96 / 1.8
    st["zbl"][:] = 1
    sc = dict(rc=4.0, delr=0.1, delr_inv=10.0, gsmooth=99.0, augt1=1, ialloy=0, mix_ref_t=0, erose_form=0,
              emb_lin_neg=0, bkgd_dyn=0)
    for e in local:
        key, vals = e[0], e[1:]
        ti = [type_ref[int(v) - 1] for v in vals[:-1]]
        v = vals[-1]
        if key in ("Ec", "alpha", "re"):
            st[key][ti[0], ti[1]] = st[key][ti[1], ti[0]] = v
        elif key == "rho0":
            st["rho0"][ti[0]] = v
        elif key in ("delta", "attrac", "repuls"):
            st[key][ti[0], ti[1]] = v
        elif key == "lattce":
            st["lattice"][ti[0], ti[1]] = st["lattice"][ti[1], ti[0]] = MEAM_LATTICES.index(v)
        elif key == "nn2":
            st["nn2"][ti[0], ti[1]] = st["nn2"][ti[1], ti[0]] = int(v)
        elif key in ("Cmin", "Cmax"):
            a = st["cmin" if key == "Cmin" else "cmax"]
            a[ti[0], ti[1], ti[2]] = a[ti[1], ti[0], ti[2]] = v
        elif key in ("rc", "delr", "gsmooth_factor"):
            sc["gsmooth" if key == "gsmooth_factor" else key] = float(v)
        elif key in ("augt1", "ialloy", "erose_form", "emb_lin_neg"):
            sc[key] = int(v)
        elif key == "mixture_ref_t":
            sc["mix_ref_t"] = int(v)
        elif key == "bkgd_dyn":  # the reader stores it in emb_lin_neg
            sc["emb_lin_neg"] = int(v)
        else:
            raise ValueError(key)
    st.update(sc)
    return st


def meam_variants():
    """MEAM settings that reach every branch of the set-up and force loop:
    name -> (state, lattice kind, lattice constant, number of types)"""
    A = MEAM_LIB["Aa"]
    return {
        "bcc_nn2": (meam_state(["Cc"], local=[("rc", 4.0), ("nn2", 1, 1, 1), ("Cmin", 1, 1, 1, 0.36)]), "bcc", 2.86, 1),
        "dia": (meam_state(["Dd"], local=[("rc", 4.5)]), "dia", 5.43, 1),
        "ibar0": (meam_state(["Aa"], local=[("rc", 4.5)], lib={"Aa": dict(A, ibar=0)}), "fcc", 3.55, 1),
        "ibar4_gsmooth": (meam_state(["Aa"], local=[("rc", 4.5), ("gsmooth_factor", 10.0)],
                                     lib={"Aa": dict(A, ibar=4)}), "fcc", 3.55, 1),
        "ibar_m5": (meam_state(["Aa"], local=[("rc", 4.5)], lib={"Aa": dict(A, ibar=-5)}), "fcc", 3.55, 1),
        "erose1": (meam_state(["Aa", "Bb"], local=meam_local(2, erose_form=1) +
                              [("attrac", 1, 1, 0.05), ("repuls", 1, 1, 0.03), ("attrac", 1, 2, 0.02),
                               ("repuls", 2, 2, 0.04)]), "fcc", 3.58, 2),
        "erose2_emb_lin_neg": (meam_state(["Aa", "Bb"], local=meam_local(2, erose_form=2, emb_lin_neg=1) +
                                          [("attrac", 1, 1, 0.05), ("repuls", 1, 2, 0.03)]), "fcc", 3.58, 2),
        "b2_nn2": (meam_state(["Aa", "Cc"], local=[("rc", 4.5), ("lattce", 1, 2, "b2"), ("nn2", 1, 2, 1),
                                                   ("Cmin", 1, 1, 2, 0.8), ("Cmin", 2, 2, 1, 0.9)]), "fcc", 3.5, 2),
        "c11": (meam_state(["Dd", "Aa"], local=[("rc", 4.5), ("lattce", 1, 2, "c11")]), "fcc", 3.6, 2),
        "b1_bkgd_dyn": (meam_state(["Aa", "Bb"], local=[("rc", 4.5), ("lattce", 1, 2, "b1"), ("bkgd_dyn", 1),
                                                         ("augt1", 0)]), "fcc", 3.58, 2),
    }
