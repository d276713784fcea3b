"""GPU BDF clock (mapp_dmd_bdf_run) against the oracle's restatement of
Clock_BDF::init/run/solve/interpolate/hi_lo_err_est on the same seeded inputs.

Bars: the same number of steps, order and solve count per step; step sizes
within 1e-8 relative; final concentrations within 1e-8 (the device dot
products sum in another order than the oracle's loops, and the step-size
control feeds those sums back through pow(), so the trajectories agree to a
few ulps amplified over the run, not bit for bit).  With the DMD_II mobility
and an initial step of 8 the controller holds the step size and raises the
order to 2 (src/clock_bdf.cpp:582-664), so those runs must contain order-2
steps: the hi/lo error estimates, the order-raise branch and the order-2
corrector (with and without the sign fix of :327) are all exercised.  With
the DMD_I mobility the solver's cost test keeps the step where the error ratio
is clamped to 2 every step, so the controller stays at order 1 (the oracle
agrees)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("variant,fix,dt0", [("DMD_I", False, -1.0), ("DMD_II", False, -1.0), ("DMD_II", False, 8.0),
                                            ("DMD_II", True, 8.0)])
def test_bdf_matches_oracle(small_tables, variant, fix, dt0):
    from mapp_amd import workload as W
    from mapp_amd.dmd import setup
    from oracle import oracle as O
    lat = W.fcc_alloy(4, seed=4)
    ff = setup(lat, small_tables, no_i=3, variant=variant)
    orc = O.make(lat, small_tables, no_i=3, variant=variant)
    n = 16
    code_g, rep_g, ng_g = ff.bdf_run(n, 2, initial_del_t=dt0, fix=fix)
    code_o, rep_o, ng_o = orc.bdf_run(n, 2, initial_del_t=dt0, fix=fix)
    print(variant, "gpu", code_g, ng_g, "oracle", code_o, ng_o, flush=True)
    print(np.column_stack([rep_g[:, [0, 1, 4, 5]], rep_o[:, [0, 1, 4, 5]]]), flush=True)
    assert code_g == code_o == n
    np.testing.assert_array_equal(rep_g[:, 0], rep_o[:, 0])   # order
    np.testing.assert_array_equal(rep_g[:, 5], rep_o[:, 5])   # solves per step
    np.testing.assert_allclose(rep_g[:, 1], rep_o[:, 1], rtol=1e-8)
    c_g = ff.get("c")
    c_o = orc.get("c")
    assert np.max(np.abs(c_g - c_o)) <= 1e-8
    if dt0 > 0:
        assert rep_g[:, 0].max() == 2
    ff.close()
