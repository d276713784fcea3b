"""Pairs that cross the quadrature cut between evaluations without a list
rebuild.  The partner shares of list entries outside the cut are cleared when
the list is built and again only when an entry leaves the cut (the per-entry
cut state of pass 1), so this is where a stale share would show: sites are
moved back and forth by less than the skin, with energy_calc and force_calc
interleaved, and every evaluation is compared with the FP64 oracle at the same
positions.  The moves are checked on the host to carry pairs across the cut in
both directions."""
import numpy as np
import pytest

from tests.conftest import rel_inf

pytestmark = pytest.mark.gpu


def _inside(x, H, natms, cut):
    """half-list pairs (i < j, minimum image in the cubic box) inside the cut"""
    L = np.diag(H)
    d = x[:natms, None, :3] - x[None, :natms, :3]
    d -= L * np.round(d / L)
    r = np.sqrt((d * d).sum(-1))
    iu = np.triu_indices(natms, 1)
    return r[iu] < cut


@pytest.mark.parametrize("nt", [2, 3])
@pytest.mark.parametrize("precision,no_i,tol", [("fp64", 3, 1e-12), ("mixed", 5, 1e-5)])
def test_pairs_crossing_the_cut(tables, precision, no_i, tol, nt):
    """with three types a pair's share slot also holds two alpha-force columns
    beyond the first, cleared with it when the pair leaves the cut"""
    from mapp_amd import workload as W
    from mapp_amd.dmd import setup
    from oracle import oracle as O
    if nt != 2:
        tables = W.johnson_tables(elements=("Ni", "Cu", "Pd")[:nt], nr=2001, nrho=2001)
    lat = W.fcc_alloy(6, nt=nt, seed=21)
    ff = setup(lat, tables, no_i=no_i, variant="DMD_I", precision=precision)
    orc = O.make(lat, tables, no_i=no_i, variant="DMD_I")
    _, mod_rc, _ = ff.cutoffs()
    x0 = orc.x_all()[: lat.natms].copy()
    rng = np.random.default_rng(4)
    moves = [x0.copy(), x0.copy(), x0.copy(), x0.copy()]
    moves[1][:, :3] += rng.normal(0.0, 0.07, size=(lat.natms, 3))
    moves[3][:, :3] += rng.normal(0.0, 0.07, size=(lat.natms, 3))  # moves[2]: back to the start
    prev = _inside(x0, lat.H, lat.natms, mod_rc)
    ff.force_calc(1)
    orc.force_calc(1)
    left = entered = 0
    for step, x in enumerate(moves):
        now = _inside(x, lat.H, lat.natms, mod_rc)
        left += int(np.sum(prev & ~now))
        entered += int(np.sum(~prev & now))
        prev = now
        assert ff.set_x(x) == orc.set_x(x) == False  # within the skin: no rebuild
        if step % 2 == 1:  # an energy-only evaluation first, then the force
            e_g, e_o = ff.energy_calc(), orc.energy_calc()
            assert abs(e_g - e_o) <= tol * abs(e_o)
        en_g, en_o = ff.force_calc(1), orc.force_calc(1)
        ef = rel_inf(ff.get("f"), orc.get("f"))
        print(f"nt {nt} {precision} move {step}: pairs left the cut {left}, entered {entered}; "
              f"force rel {ef:.3e}", flush=True)
        assert ef <= tol
        assert abs(en_g[0] - en_o[0]) <= tol * abs(en_o[0])
        assert rel_inf(ff.get("E"), orc.get("E")) <= tol
    assert left > 0 and entered > 0
    ff.close()
