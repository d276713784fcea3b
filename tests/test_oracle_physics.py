"""Oracle pins by physics and geometry:
  - the half list plus the lists of lower-index owners, mapped to (owner,
    lattice image), equals the brute-force set of periodic neighbours
  - one type (where the reference's nt = 2 slot quirks vanish) with a 14-point
    rule: f = -grad U by central differences in x and alpha; sum f = 0;
    energy_calc = force_calc's energy; translation invariance
  - c_d conserves each species (antisymmetric exchange) and vanishes for a
    uniform perfect crystal; calc_g's g = -d cost / d c (one type)
  - the classical branch (alpha >= alpha_max) equals plain EAM"""
import numpy as np
import pytest

from mapp_amd import workload as W
from oracle import oracle as O


def _make(lat, tables, **kw):
    return O.make(lat, tables, **kw)


@pytest.fixture(scope="module")
def t1(small_tables):
    """single-species tables (Ni) on the coarse grid"""
    t = small_tables
    return W.FSTables(["Ni"], t.nrho, t.drho, t.nr, t.dr, t.F[:1], t.rho[:1, :1], t.rphi[:1], t.masses[:1])


@pytest.mark.parametrize("n,seed", [(4, 1), (6, 2)])
def test_neighbor_list_equals_brute_force(tables, n, seed):
    lat = W.fcc_alloy(n, seed=seed)
    o = _make(lat, tables)
    off, nl = o.neighbor_list()
    natms = o.natms
    X = o.x_all()
    gsrc = o.ghosts()[1]
    owner = np.arange(X.shape[0])
    for g in range(len(gsrc)):
        s = gsrc[g]
        owner[natms + g] = owner[s] if s >= natms else s
    L = lat.H[0, 0]
    cut = np.sqrt(o.scalars()["cut_sk_sq"])
    found = set()
    for i in range(natms):
        for j in nl[off[i]:off[i + 1]]:
            assert j > i
            k = owner[j]
            shift = np.round((X[j, :3] - X[k, :3]) / L).astype(int)
            key = (i, k, *shift)
            assert key not in found
            found.add(key)
            # mirror entry seen from k (owned partners only)
            if j < natms:
                found.add((k, i, *(-shift)))
    brute = set()
    P = X[:natms, :3]
    for i in range(natms):
        d = P[i] - P
        for sh in np.array(np.meshgrid([-1, 0, 1], [-1, 0, 1], [-1, 0, 1])).reshape(3, -1).T:
            dd = d - sh * L
            r2 = (dd ** 2).sum(1)
            for k in np.nonzero(r2 < cut * cut)[0]:
                if k == i and not sh.any():
                    continue
                brute.add((i, int(k), *sh))
    # every brute-force neighbour of i is either in i's list or i is in its owner's list
    assert brute == found


def test_ghosts_are_periodic_images(tables):
    lat = W.fcc_alloy(5, seed=3)
    o = _make(lat, tables)
    X = o.x_all()
    natms = o.natms
    _, src = o.ghosts()
    L = lat.H[0, 0]
    for g, s in enumerate(src):
        d = (X[natms + g, :3] - X[s, :3]) / L
        assert np.allclose(d, np.round(d), atol=1e-12)
        assert np.abs(np.round(d)).sum() == 1
        assert np.array_equal(X[natms + g, 3:], X[s, 3:])


@pytest.fixture(scope="module")
def one_type(t1):
    lat = W.fcc_alloy(4, nt=1, seed=5)
    o = _make(lat, t1, no_i=14)
    en = o.force_calc(1)
    return lat, o, en


def test_forces_are_minus_gradient(one_type):
    lat, o, en = one_type
    f = o.get("f").copy()
    x0 = o.x_all()[: o.natms].copy()
    rng = np.random.default_rng(0)
    picks = [(int(k), int(d)) for k, d in zip(rng.integers(0, o.natms, 6), rng.integers(0, 4, 6))]
    fmax = np.abs(f[:, :3]).max()
    amax = np.abs(f[:, 3]).max()
    for k, d in picks:
        h = 1e-5 if d < 3 else 2e-3
        xp = x0.copy()
        xp[k, d] += h
        o.set_x(xp)
        ep = o.force_calc(0)[0]
        xm = x0.copy()
        xm[k, d] -= h
        o.set_x(xm)
        em = o.force_calc(0)[0]
        fd = -(ep - em) / (2 * h)
        scale = fmax if d < 3 else amax
        assert abs(fd - f[k, d]) < 2e-6 * scale + 1e-10, (k, d, fd, f[k, d])
    o.set_x(x0)


def test_sum_of_forces_and_energy_calc(one_type):
    lat, o, en = one_type
    o.set_x(o.x_all()[: o.natms].copy())
    en = o.force_calc(1)
    f = o.get("f")
    assert np.abs(f[:, :3].sum(0)).max() < 1e-10 * np.abs(f[:, :3]).max() * o.natms
    assert abs(o.energy_calc() - en[0]) < 1e-12 * abs(en[0])


def test_translation_invariance(t1):
    lat = W.fcc_alloy(4, nt=1, seed=6)
    o1 = _make(lat, t1, no_i=5)
    e1 = o1.force_calc(1)[0]
    lat2 = W.fcc_alloy(4, nt=1, seed=6)
    shift = np.array([0.123, -0.211, 0.377])
    cart = lat2.cart + shift
    lat2.x[:, :3] = W.x2s(cart, W.m3inv(lat2.H))
    o2 = _make(lat2, t1, no_i=5)
    e2 = o2.force_calc(1)[0]
    assert abs(e1 - e2) < 1e-11 * abs(e1)


@pytest.mark.parametrize("variant", ["DMD", "DMD_I", "DMD_II"])
def test_c_d_conserves_species(tables, variant):
    lat = W.fcc_alloy(5, seed=7)
    o = _make(lat, tables, variant=variant)
    o.force_calc(1)
    o.create_2nd_neigh_lst()
    o.c_d_calc()
    cd = o.get("c_d")
    assert np.abs(cd).max() > 0
    assert np.abs(cd.sum(0)).max() < 1e-12 * np.abs(cd).sum()


def _uniform(nt, seed=8):
    lat = W.fcc_alloy(4, nt=nt, seed=seed, jitter=0.0)
    lat.x[:, 3:] = 80.0
    if nt == 2:
        lat.c[:, 0] = 0.3
        lat.c[:, 1] = 0.7
    else:
        lat.c[:] = 0.6
    return lat


@pytest.mark.parametrize("variant", ["DMD", "DMD_I", "DMD_II"])
def test_c_d_zero_for_uniform_crystal(t1, variant):
    o = _make(_uniform(1), t1, variant=variant)
    o.force_calc(1)
    o.create_2nd_neigh_lst()
    assert np.all(o.crd() == 12)
    o.c_d_calc()
    assert np.abs(o.get("c_d")).max() < 1e-9


@pytest.mark.parametrize("variant", ["DMD", "DMD_I", "DMD_II"])
def test_c_d_index_quirk_two_types(tables, variant):
    """nt = 2: the reference reads dE[jatm+jtype] (src/ff_dmd_I.cpp:2336), which
    mixes atoms and types, so a uniform crystal gets a nonzero dc/dt; with the
    index as [atom*nt+type] it vanishes, so the quirk is the whole effect"""
    lat = _uniform(2)
    o = _make(lat, tables, variant=variant)
    o.force_calc(1)
    o.create_2nd_neigh_lst()
    o.c_d_calc()
    assert np.abs(o.get("c_d")).max() > 1e-6
    o.set_fix_dE_index(True)
    o.c_d_calc()
    assert np.abs(o.get("c_d")).max() < 1e-9


def _fd_gradient(o, lat, a, beta_t, ks):
    c0 = lat.c.copy()
    out = []
    for k in ks:
        h = 1e-6
        cp = c0.copy(); cp[k, 0] += h
        o.set_c(cp)
        ep = o.calc_g(1, beta_t, a)[0]
        cm = c0.copy(); cm[k, 0] -= h
        o.set_c(cm)
        em = o.calc_g(1, beta_t, a)[0]
        out.append((ep - em) / (2 * h))
    o.set_c(c0)
    return np.array(out)


def _calc_g_system(tab, variant, seed=9):
    lat = W.fcc_alloy(3, nt=1, seed=seed)
    lat.c[:] = np.random.default_rng(3).uniform(0.2, 0.8, size=lat.c.shape)
    o = _make(lat, tab, variant=variant, no_i=3)
    o.force_calc(1)
    o.create_2nd_neigh_lst()
    a = np.random.default_rng(4).uniform(-0.3, 0.3, size=lat.c.shape)
    return lat, o, a


@pytest.mark.parametrize("variant", ["DMD", "DMD_I", "DMD_II"])
def test_calc_g_direct_term_is_minus_cost_gradient(t1, variant):
    """With a null potential n does not depend on c, and calc_g's g is exactly
    -d cost / d c for every variant (the s and t sweeps)"""
    z = W.FSTables(["Ni"], t1.nrho, t1.drho, t1.nr, t1.dr, 0 * t1.F, 0 * t1.rho, 0 * t1.rphi, t1.masses)
    lat, o, a = _calc_g_system(z, variant)
    ks = [0, 7, 19, 33]
    cost, g = o.calc_g(0, 0.05, a)
    fd = _fd_gradient(o, lat, a, 0.05, ks)
    assert np.max(np.abs(-fd - g[ks, 0])) < 1e-6 * np.abs(g).max()


def test_calc_g_dmd_ii_is_minus_cost_gradient(t1):
    """DMD_II with the full potential: g = -d cost / d c including the terms
    through n (the t, v and ddE sweeps shared by all variants), to the 1e-4 the
    reference's formulas reach (its F'' row, src/ff_dmd_I.cpp:2491, also drops
    the factor 2 of the c7 term; a missing or sign-flipped term would be O(1))"""
    lat, o, a = _calc_g_system(t1, "DMD_II")
    ks = [0, 7, 19, 33]
    cost, g = o.calc_g(0, 0.05, a)
    fd = _fd_gradient(o, lat, a, 0.05, ks)
    assert np.max(np.abs(-fd - g[ks, 0])) < 2e-4 * np.abs(g).max()


@pytest.mark.parametrize("variant", ["DMD", "DMD_I"])
def test_calc_g_n_term_sign_quirk(t1, variant):
    """DMD and DMD_I add the chain-rule term through n with the opposite sign to
    the direct term (src/ff_dmd_I.cpp:2576-2670), so with the full potential their
    g is not -d cost / d c; kept as in the reference"""
    lat, o, a = _calc_g_system(t1, variant)
    ks = [0, 7, 19]
    cost, g = o.calc_g(0, 0.05, a)
    fd = _fd_gradient(o, lat, a, 0.05, ks)
    assert np.max(np.abs(-fd - g[ks, 0])) > 0.5 * np.abs(fd).max()


def test_classical_branch_equals_plain_eam(small_tables):
    """alpha >= alpha_max everywhere: the DMD energy minus its per-site alpha terms
    is plain EAM's (F linear, so the quartic and cubic F rows agree exactly)"""
    t = small_tables
    F = -0.5 * np.arange(t.nrho)[None, :] * t.drho
    t1 = W.FSTables(["Ni"], t.nrho, t.drho, t.nr, t.dr, F, t.rho[:1, :1], t.rphi[:1], t.masses[:1])
    lat = W.fcc_alloy(4, nt=1, seed=10, c_per_site=1.0)
    o = _make(lat, t1, alpha_max=10.0)
    e_dmd = o.force_calc(1)[0]
    sc = o.scalars()
    kT = sc["kbT"]
    per_site = (1.5 * kT * np.log(lat.x[:, 3]) + sc["c_0"][0]).sum()
    eam = O.Oracle(nt=1, variant="EAM")
    eam.set_fs(t1)
    eam.set_params(masses=t1.masses)
    x3 = lat.x[:, :3].copy()
    eam.set_system(lat.H, x3, np.zeros((lat.natms, 1)))
    eam.init(0.5)
    e_eam = eam.force_calc(1)[0]
    assert abs((e_dmd - per_site) - e_eam) < 1e-11 * abs(e_eam)
    assert abs(eam.energy_calc() - e_eam) < 1e-11 * abs(e_eam)
    fd = o.get("f")[:, :3]
    fe = eam.get("f")
    assert np.abs(fd - fe).max() < 1e-11 * np.abs(fe).max()


def test_eam_forces_are_minus_gradient(small_tables):
    t = small_tables
    t2 = W.FSTables(["Ni", "Cu"], t.nrho, t.drho, t.nr, t.dr, t.F, t.rho, t.rphi, t.masses)
    lat = W.fcc_alloy(4, nt=2, seed=11)
    types = (np.arange(lat.natms) % 2).astype(np.float64).reshape(-1, 1)
    eam = O.Oracle(nt=2, variant="EAM")
    eam.set_fs(t2)
    eam.set_params(masses=t2.masses)
    eam.set_system(lat.H, lat.x[:, :3].copy(), types)
    eam.init(0.5)
    eam.force_calc(1)
    f = eam.get("f").copy()
    x0 = eam.x_all()[: eam.natms].copy()
    for k, d in ((3, 0), (40, 1), (77, 2), (200, 0)):
        h = 1e-5
        xp = x0.copy(); xp[k, d] += h
        eam.set_x(xp); ep = eam.force_calc(0)[0]
        xm = x0.copy(); xm[k, d] -= h
        eam.set_x(xm); em = eam.force_calc(0)[0]
        assert abs(-(ep - em) / (2 * h) - f[k, d]) < 1e-6 * np.abs(f).max()
