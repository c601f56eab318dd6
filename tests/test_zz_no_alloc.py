"""storm_fvt_get_gradients_noalloc through the C ABI: the same delh as
storm_fvt_get_gradients, bit for bit, on two calls in a row (the second on
a smaller context, which reuses the same device scratch array), with the
Green-Gauss option too; sizes beyond its scratch and a context with no run
are refused."""
import ctypes as C

import numpy as np
import pytest

from cases_small import option_cases, small_cases

pytestmark = pytest.mark.gpu


def _both(s):
    from storm_amd import fvt
    L = fvt.load()
    dp = C.POINTER(C.c_double)
    a = [np.zeros(s.N), np.zeros(s.N)]
    b = [np.full(s.N, np.nan), np.full(s.N, np.nan)]
    assert L.storm_fvt_get_gradients(s.h, a[0].ctypes.data_as(dp), a[1].ctypes.data_as(dp)) == 0
    rc = L.storm_fvt_get_gradients_noalloc(s.h, b[0].ctypes.data_as(dp), b[1].ctypes.data_as(dp))
    assert rc == 0, L.storm_fvt_last_error(s.h)
    return a, b


def test_noalloc_gradients_equal_the_allocating_entry_point():
    from storm_amd import fvt
    runs = [(small_cases()["meander_crit"], 40), (small_cases()["flume_rk2"], 9),
            (option_cases()["flume_gauss_rk3"], 12)]
    for (case, _), n in runs:
        print(f"[noalloc] {case.name} N={case.n_elems} steps={n}", flush=True)
        s = fvt.FVTSolver(case)
        s.run(n)
        a, b = _both(s)
        s.close()
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
        assert np.abs(a[0]).max() > 0


def test_noalloc_gradients_refusals():
    from storm_amd import cases, fvt
    L = fvt.load()
    dp = C.POINTER(C.c_double)
    case, _ = small_cases()["flume_rk2"]
    s = fvt.FVTSolver(case)
    out = np.zeros(s.N)
    # no run started: the state it would read is not on the device yet
    assert L.storm_fvt_get_gradients_noalloc(s.h, out.ctypes.data_as(dp), None) == -5
    s.run(1)
    assert L.storm_fvt_get_gradients_noalloc(s.h, out.ctypes.data_as(dp), None) == 0
    s.close()
    # a stage inflow: discharge strings are limited to 512 edges
    big = cases.flume(nx=1025, ny=1024, length=1025.0, width=1024.0, seed=3, in_kind=4,
                      q_in=2.5)
    assert big.n_elems > 2097152
    print(f"[noalloc] refusal N={big.n_elems}", flush=True)
    s = fvt.FVTSolver(big, dict(istop=1))
    s.run(1)
    out = np.zeros(s.N)
    rc = L.storm_fvt_get_gradients_noalloc(s.h, out.ctypes.data_as(dp), None)
    msg = L.storm_fvt_last_error(s.h).decode()
    s.close()
    assert rc == -1 and "scratch" in msg, (rc, msg)
