"""HIP path vs the CPU oracle on the same seeded cases (full time loop)."""
import os

import numpy as np
import pytest

from cases_small import golden_cases, small_cases

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

pytestmark = pytest.mark.gpu

# relative L-infinity bound on h, hu, hv after N steps (the parity bound, 1e-10)
RTOL = 1e-10


def _cmp(a, b):
    scale = max(np.abs(a).max(), 1e-300)
    return np.abs(a - b).max() / scale


@pytest.mark.parametrize("name", list(golden_cases().keys()))
def test_step_parity(name):
    import oracle as O
    from storm_amd import fvt
    case, n = golden_cases()[name]
    print(f"[gpu parity] {name} N={case.n_elems} steps={n}", flush=True)
    orc = O.Oracle(case)
    r_o = orc.run(n)
    orc.finalize()
    so = orc.state()
    s = fvt.FVTSolver(case)
    it, tt, rm = s.run(n)
    s.finalize()
    sg = s.state()
    assert it == r_o["iter"]
    assert tt == pytest.approx(r_o["ttime"], rel=1e-13)
    for k in ("h", "qh", "qu", "qv", "u", "v"):
        assert _cmp(so[k], sg[k]) <= RTOL, (k, _cmp(so[k], sg[k]))
    for k in ("hvtx", "zetavtx", "phi1", "phi2", "phi3"):
        assert _cmp(so[k], sg[k]) <= 1e-9, (k, _cmp(so[k], sg[k]))
    assert np.array_equal(so["dry"], sg["dry"])
    # delh of the last stage, against what the reference's error_est read
    g = np.load(os.path.join(GOLDEN, f"{name}.npz"))
    gx, gy = s.gradients()
    assert _cmp(g["delhx"], gx) <= 1e-9, _cmp(g["delhx"], gx)
    assert _cmp(g["delhy"], gy) <= 1e-9, _cmp(g["delhy"], gy)


@pytest.mark.parametrize("rk", [1, 3])
def test_steps_after_a_steady_stop_change_nothing(rk):
    """Steps launched after a steady run has converged run no stage; under
    an odd RK order each would swap the stage buffers an odd number of
    times.  State, final to_vertices and delh are those of the last step
    taken, whether 0, 1 or 2 steps were launched after it."""
    import oracle as O
    from storm_amd import cases, fvt
    case = cases.flume(nx=16, ny=6, length=100.0, width=30.0, seed=41, dt_op=0, cfl=0.3,
                       istop=1200, q_in=15.0, steady=True, stop_cr=1, rstop=1e-4, rkorder=rk)
    orc = O.Oracle(case)
    k = orc.run(1200)["iter"]
    orc.finalize()
    so = orc.state()
    assert k < 1200
    out = []
    for n in (k, k + 1, k + 2):
        print(f"[gpu steady stop] rk={rk} converged at {k}, launched {n}", flush=True)
        s = fvt.FVTSolver(case)
        it, _, _ = s.run(n)
        s.finalize()
        sg = s.state()
        gx, gy = s.gradients()
        s.close()
        assert it == k
        for key in ("h", "qh", "qu", "qv", "u", "v"):
            assert _cmp(so[key], sg[key]) <= RTOL, (n, key, _cmp(so[key], sg[key]))
        for key in ("hvtx", "zetavtx", "phi1", "phi2", "phi3"):
            assert _cmp(so[key], sg[key]) <= 1e-9, (n, key, _cmp(so[key], sg[key]))
        assert np.array_equal(so["dry"], sg["dry"])
        sg["gx"], sg["gy"] = gx, gy
        out.append(sg)
    for o in out[1:]:
        for key in out[0]:
            assert np.array_equal(out[0][key], o[key]), key


@pytest.mark.parametrize("cap", [384, 512])
def test_tile_capacity_leaves_results_unchanged(cap):
    """The tiling (and so the block size) changes only the order of the
    residual sums: the state is bit-identical."""
    from storm_amd import fvt
    case, n = small_cases()["meander_crit"]
    print(f"[gpu tiles] meander_crit cap={cap}", flush=True)
    out = []
    for block in (256, cap):
        s = fvt.FVTSolver(case, dict(block=block))
        s.run(n)
        s.finalize()
        out.append(s.state())
        s.close()
    for k in ("h", "qh", "qu", "qv", "u", "v", "hvtx", "zetavtx", "phi1", "phi2", "phi3", "dry"):
        assert np.array_equal(out[0][k], out[1][k]), k


@pytest.mark.parametrize("name,n", [("flume_rk2", 7), ("flume_rk3_slip", 30), ("flume_rk1_steady", 9)])
def test_graph_replay_leaves_results_unchanged(name, n):
    """Steps replayed from the captured two-step HIP graph (odd counts start
    with a direct step) give the results of directly launched steps."""
    from storm_amd import fvt
    case, _ = small_cases()[name]
    print(f"[gpu graph] {name} steps={n}", flush=True)
    out = []
    for graph in (0, 1):
        s = fvt.FVTSolver(case, dict(graph=graph))
        s.run(2)  # captures the graph at step 0 and replays it
        s.run(1)  # one direct step: out of the graph's phase
        s.run(n)  # realigns with a direct step, then replays
        s.finalize()
        out.append(s.state())
        s.close()
    for k in ("h", "qh", "qu", "qv", "u", "v", "hvtx", "zetavtx", "phi1", "phi2", "phi3", "dry"):
        assert np.array_equal(out[0][k], out[1][k]), k


@pytest.mark.parametrize("name", list(golden_cases().keys()))
def test_cell_orders_give_identical_results(name):
    """The device's cell order is a layout choice only: input order
    (reorder=0), the Hilbert curve (1) and reverse Cuthill-McKee (2) give
    the same bits, every flux being gathered in the reference's order.  The
    residual rmax alone is a sum over tiles, whose order follows the cell
    order: it agrees to rounding."""
    from storm_amd import fvt
    case, n = golden_cases()[name]
    print(f"[gpu cell order] {name}", flush=True)
    out = []
    for reorder in (0, 1, 2):
        s = fvt.FVTSolver(case, dict(reorder=reorder))
        it, tt, rm = s.run(n)
        s.finalize()
        st = s.state()
        st.update(s.output())
        gx, gy = s.gradients()
        st["gx"], st["gy"] = gx, gy
        st["run"] = np.array([it, tt])
        st_rmax = rm
        s.close()
        out.append((st, st_rmax))
    for st, rm in out[1:]:
        for k, v in out[0][0].items():
            assert np.array_equal(v, st[k], equal_nan=True), (name, k)
        assert rm == pytest.approx(out[0][1], rel=1e-12, abs=1e-300)
