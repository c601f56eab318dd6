"""HIP path vs the CPU oracle on the same seeded cases (full time loop)."""
import numpy as np
import pytest

from cases_small import small_cases

pytestmark = pytest.mark.gpu

# relative L-infinity bound on h, hu, hv after N steps (north-star bound 1e-10)
RTOL = 1e-10


def _cmp(a, b):
    scale = max(np.abs(a).max(), 1e-300)
    return np.abs(a - b).max() / scale


@pytest.mark.parametrize("name", list(small_cases().keys()))
def test_step_parity(name):
    import oracle as O
    from storm_amd import fvt
    case, n = small_cases()[name]
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
