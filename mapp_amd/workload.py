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
    by the types, when alpha_per_site); c_A ~ U[0,1], c_B = 1 - c_A."""
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
    else:
        cA = rng.uniform(0.0, 1.0, size=natms)
        c = np.zeros((natms, nt))
        c[:, 0] = cA
        c[:, 1] = 1.0 - cA
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
    """n^3 cubic cells of an FCC or BCC lattice, jittered N(0, jitter); type 1
    on a fraction frac of the sites (nt = 2).  Returns H, reduced positions
    [natms][3] and the types."""
    rng = np.random.default_rng(seed)
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]]) if kind == "fcc" else \
        np.array([[0, 0, 0], [0.5, 0.5, 0.5]])
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
