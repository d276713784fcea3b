"""GPU BDF clock (mapp_dmd_bdf_run) against the oracle's restatement of
Clock_BDF::init/run/solve/interpolate/hi_lo_err_est on the same seeded inputs.

Bars: the same number of steps, order and solve count per step; step sizes
within 1e-8 relative; final concentrations within 1e-8 (the device dot
products sum in another order than the oracle's loops, and the step-size
control feeds those sums back through pow(), so the trajectories agree to a
few ulps amplified over the run, not bit for bit)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("variant,fix", [("DMD_I", False), ("DMD_II", False), ("DMD_II", True)])
def test_bdf_matches_oracle(small_tables, variant, fix):
    from mapp_amd import workload as W
    from mapp_amd.dmd import setup
    from oracle import oracle as O
    lat = W.fcc_alloy(4, seed=4)
    ff = setup(lat, small_tables, no_i=3, variant=variant)
    orc = O.make(lat, small_tables, no_i=3, variant=variant)
    n = 16
    code_g, rep_g, ng_g = ff.bdf_run(n, 2, fix=fix)
    code_o, rep_o, ng_o = orc.bdf_run(n, 2, fix=fix)
    print(variant, "gpu", code_g, ng_g, "oracle", code_o, ng_o, flush=True)
    print(np.column_stack([rep_g[:, [0, 1, 4, 5]], rep_o[:, [0, 1, 4, 5]]]), flush=True)
    assert code_g == code_o == n
    np.testing.assert_array_equal(rep_g[:, 0], rep_o[:, 0])   # order
    np.testing.assert_array_equal(rep_g[:, 5], rep_o[:, 5])   # solves per step
    np.testing.assert_allclose(rep_g[:, 1], rep_o[:, 1], rtol=1e-8)
    c_g = ff.get("c")
    c_o = orc.get("c")
    assert np.max(np.abs(c_g - c_o)) <= 1e-8
    ff.close()
