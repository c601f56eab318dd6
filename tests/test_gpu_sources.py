"""The optional source branches on the device (wind, culverts, residual
smoothing) against the oracle beyond the golden cases: sweep counts and
factors the reference cannot read (read_options.f90:350 reads them from
inside the keyword), the xsf edge loop, batches, graph replay, tables
changed mid-run, the culverts' cells and discharges, and a larger meander
with every branch on."""
import numpy as np
import pytest

from cases_small import _meander_culverts, _with_wind, shuffled, source_cases

pytestmark = pytest.mark.gpu

RTOL = 1e-10


def _cmp(a, b):
    return np.abs(a - b).max() / max(np.abs(a).max(), 1e-300)


def _check(so, sg, tag):
    for k in ("h", "qh", "qu", "qv", "u", "v"):
        assert _cmp(so[k], sg[k]) <= RTOL, (tag, k, _cmp(so[k], sg[k]))
    for k in ("hvtx", "zetavtx", "phi1", "phi2", "phi3"):
        assert _cmp(so[k], sg[k]) <= 1e-9, (tag, k, _cmp(so[k], sg[k]))
    assert np.array_equal(so["dry"], sg["dry"]), tag


def _run_both(case, n, opts=None):
    import oracle as O
    from storm_amd import fvt
    o = O.Oracle(case, opts)
    r = o.run(n)
    o.finalize()
    s = fvt.FVTSolver(case, opts)
    it, tt, _ = s.run(n)
    s.finalize()
    assert it == r["iter"] and tt == pytest.approx(r["ttime"], rel=1e-13)
    return o, s


@pytest.mark.parametrize("nsw,eps", [(0, 0.5), (1, 0.5), (3, 0.25), (2, 1.0)])
@pytest.mark.parametrize("reorder", [1, 0])
def test_rsavg_sweeps_and_factors(nsw, eps, reorder):
    from cases_small import source_cases
    case, n = source_cases()["flume_rsavg_steady"]
    opts = dict(rsavg_n=nsw, rsavg_eps=eps, reorder=reorder)
    print(f"[gpu sources] rsavg n={nsw} eps={eps} reorder={reorder}", flush=True)
    o, s = _run_both(case, n, opts)
    _check(o.state(), s.state(), (nsw, eps))
    s.close()


def test_rsavg_with_the_xsf_edge_loop_and_culverts():
    from storm_amd import cases
    case = cases.meander(ns=40, nn=12, length=400.0, width=120.0, amp=20.0, wavelength=200.0,
                         chan_w=40.0, chan_d=2.0, n_bumps=5, q0=5.0, q1=20.0, t_ramp=10.0,
                         dt_op=0, cfl=0.02, istop=50, seed=71, rkorder=3, rsavg=True,
                         fluxes="xsf")
    case.culverts = _meander_culverts()
    print("[gpu sources] rsavg + xsf + culverts", flush=True)
    o, s = _run_both(case, 50)
    _check(o.state(), s.state(), "xsf")
    ci, co, src = s.culverts()
    oi, oo, osrc = o.culverts()
    assert np.array_equal(ci, oi) and np.array_equal(co, oo)
    assert _cmp(osrc, src) <= RTOL
    s.close()


def test_culvert_cells_and_discharges():
    case, n = source_cases()["meander_culverts"]
    print("[gpu sources] culvert cells", flush=True)
    for reorder in (1, 0):
        o, s = _run_both(case, n, dict(reorder=reorder))
        ci, co, src = s.culverts()
        oi, oo, osrc = o.culverts()
        assert np.array_equal(ci, oi) and np.array_equal(co, oo)
        assert _cmp(osrc, src) <= RTOL
        # a dry inlet, a curve's first and last discharges, interpolations
        assert src[1] == 0.0 and src[2] == 0.8 and src[3] == 0.3
        assert 0.4 < src[0] < 5.0 and 0.25 < src[4] < 0.7
        s.close()


def test_tables_changed_mid_run_and_graph_replay():
    """Wind cleared and culverts replaced between runs, with and without
    captured step graphs: the oracle gets the same changes."""
    import oracle as O
    from storm_amd import fvt
    case, _ = source_cases()["meander_sources_rk3"]
    for graph in (0, 1):
        print(f"[gpu sources] tables mid-run graph={graph}", flush=True)
        o = O.Oracle(case)
        s = fvt.FVTSolver(case, dict(graph=graph))
        o.run(10)
        s.run(10)
        o.set_wind(np.zeros(0), np.zeros(0), np.zeros(0))
        s.clear_wind()
        o.run(20)
        s.run(20)
        cv = _meander_culverts()[:2]
        o.set_culverts(cv)
        s.set_culverts(cv)
        o.run(30)
        s.run(30)
        o.finalize()
        s.finalize()
        _check(o.state(), s.state(), graph)
        s.close()


def test_batch_with_wind_and_rsavg():
    """Three steady flumes with wind at their vertices and residual
    smoothing, two of which converge (at steps 165 and 220) while the third
    runs on: each reach equals its own oracle."""
    import oracle as O
    from storm_amd import cases, fvt
    kw = dict(dt_op=0, cfl=0.3, istop=300, steady=True, stop_cr=1, rsavg=True, rstop=3e-4)
    reaches = [_with_wind(cases.flume(nx=16, ny=6, length=100.0, width=30.0, seed=81 + r,
                                      q_in=15.0 + 5 * r, **kw), True, 8100 + r)
               for r in range(3)]
    print("[gpu sources] batch wind + rsavg", flush=True)
    b = fvt.BatchSolver(reaches)
    b.run(300)
    b.finalize()
    parts = b.split(b.state())
    it, _, _, done = b.status()
    for r, case in enumerate(reaches):
        o = O.Oracle(case)
        res = o.run(300)
        o.finalize()
        assert it[r] == res["iter"]
        _check(o.state(), parts[r], r)
    assert list(it) == [165, 220, 300] and list(done) == [True, True, False]
    b.close()


def test_larger_meander_with_every_branch():
    """A 9,600-cell meander (ns=200, nn=24) at RK3 with wind at the cells,
    the culverts and residual smoothing: many levels per sweep."""
    from storm_amd import cases
    case = cases.meander(ns=200, nn=24, length=2000.0, width=240.0, amp=20.0,
                         wavelength=200.0, chan_w=40.0, chan_d=2.0, n_bumps=20, q0=10.0,
                         q1=40.0, t_ramp=20.0, dt_op=0, cfl=0.02, istop=40, seed=91,
                         rkorder=3, rsavg=True)
    case = _with_wind(case, False, 9191, coef=1.0)
    case.culverts = _meander_culverts()
    print(f"[gpu sources] meander N={case.n_elems}", flush=True)
    o, s = _run_both(case, 40)
    _check(o.state(), s.state(), "large")
    s.close()


def test_rsavg_on_a_shuffled_numbering():
    """Cells in random order: a few levels of thousands of cells each, wider
    than the shared-memory sweep's block, so the global-memory level sweep
    runs."""
    from storm_amd import cases
    case = shuffled(cases.meander(ns=100, nn=40, length=1000.0, width=400.0, amp=20.0,
                                  wavelength=200.0, chan_w=40.0, chan_d=2.0, n_bumps=10,
                                  q0=10.0, q1=40.0, t_ramp=20.0, dt_op=0, cfl=0.02, istop=30,
                                  seed=95, rkorder=2, rsavg=True), 5)
    print(f"[gpu sources] shuffled meander N={case.n_elems}", flush=True)
    o, s = _run_both(case, 30)
    _check(o.state(), s.state(), "shuffled")
    s.close()
