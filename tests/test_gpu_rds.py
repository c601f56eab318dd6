"""storm_fvt_drynwet and storm_fvt_fvsources (the RDS solver's node
routines) on the device: bit for bit against the reference routines' own
outputs (tests/golden/units.npz, from the per-routine harness), in both
cell orders, and against the oracle on a larger mesh."""
import os

import numpy as np
import pytest

from cases_small import small_cases

pytestmark = pytest.mark.gpu

UNITS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "units.npz")


@pytest.mark.parametrize("name", ["meander_crit", "dambreak_wetdry"])
@pytest.mark.parametrize("reorder", [1, 0])
def test_node_routines_match_the_reference(name, reorder):
    from storm_amd import fvt
    u = np.load(UNITS)
    case, _ = small_cases()[name]
    print(f"[gpu rds] {name} reorder={reorder}", flush=True)
    s = fvt.FVTSolver(case, dict(reorder=reorder))
    h = u[f"rds_{name}_h"]
    zo, pd = s.drynwet(case.zvtx, h)
    src = s.fvsources(case.zvtx, h)
    s.close()
    assert np.array_equal(zo, u[f"rds_{name}_zout"])
    assert np.array_equal(pd, u[f"rds_{name}_pd"])
    assert np.array_equal(src, u[f"rds_{name}_source"])


def test_node_routines_match_the_oracle_on_a_larger_mesh():
    import oracle as O
    from storm_amd import cases, fvt
    case = cases.meander(ns=200, nn=40, length=2000.0, width=600.0, amp=100.0,
                         wavelength=1000.0, chan_w=80.0, chan_d=4.0, n_bumps=20, seed=77)
    rng = np.random.default_rng(78)
    h = np.maximum(rng.uniform(-0.5, 2.0, case.n_pts), 0.0)
    print(f"[gpu rds] meander N={case.n_elems}", flush=True)
    s = fvt.FVTSolver(case)
    zo, pd = s.drynwet(case.zvtx, h)
    src = s.fvsources(case.zvtx, h)
    s.close()
    src_o, zo_o, pd_o = O.Oracle(case).rds(case.opts["h_dry"], case.zvtx, h)
    assert np.array_equal(zo, zo_o) and np.array_equal(pd, pd_o)
    assert np.array_equal(src, src_o)
    assert pd[:, 0].sum() > 100
