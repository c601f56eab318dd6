"""The C ABI without a GPU: libstorm_fvt.so and libstorm_mpnn.so load, export
every function include/storm_fvt.h and include/storm_mpnn.h declare, and the
Python mirrors bind exactly those.  The only calls made are host-only ones:
the default options, the least-squares molecule of a small mesh (the
set-up table of elemnt_db.f90, which needs no device) and an option check
that fails before any device work."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from storm_amd import cases, fvt, mpnn

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"\b(storm_(?:fvt|mpnn)_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header,lib,mirror", [
    ("storm_fvt.h", fvt.LIB_PATH, fvt.EXPORTS),
    ("storm_mpnn.h", mpnn.LIB_PATH, mpnn.EXPORTS),
])
def test_library_exports_every_declared_function(header, lib, mirror):
    names = declared(header)
    assert len(names) >= 10
    L = C.CDLL(lib)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(mirror) == names


def test_default_opts_are_the_reference_defaults():
    L = fvt.load()
    o = fvt.Opts()
    L.storm_fvt_default_opts(C.byref(o))
    # initialize_options.f90: rstop = REAL(1.0e-4) (a default-real literal),
    # h_dry 0.001; the library's own defaults: reorder on, graph off
    assert o.rstop == float(np.float32(1e-4))
    assert o.h_dry == 0.001 and o.reorder == 1 and o.graph == 0


def test_inspect_molecule_on_the_host():
    case = cases.flume(nx=6, ny=3, length=30.0, width=9.0, seed=5)
    keep = {}
    m = fvt.mesh_struct(case, keep)
    L = fvt.load()
    N = case.n_elems
    ptr = np.zeros(N + 1, dtype=np.int32)
    n = L.storm_fvt_inspect_molecule(C.byref(m), ptr.ctypes.data_as(C.POINTER(C.c_int32)), None)
    assert n > 0 and ptr[-1] == n
    cells = np.zeros(n, dtype=np.int32)
    L.storm_fvt_inspect_molecule(C.byref(m), ptr.ctypes.data_as(C.POINTER(C.c_int32)),
                                 cells.ctypes.data_as(C.POINTER(C.c_int32)))
    # every cell's molecule has at least its edge neighbours, none is itself
    counts = np.diff(ptr)
    nb = (case.edge_e[:, 1] >= 0)
    deg = np.bincount(case.edge_e[nb, 0], minlength=N) + np.bincount(case.edge_e[nb, 1], minlength=N)
    assert (counts >= deg).all()
    for i in range(N):
        assert i not in cells[ptr[i]:ptr[i + 1]]


def test_steady_batch_with_odd_rk_order_is_refused():
    """A steady reach that stops early keeps its state in place while the
    other reaches' steps swap the stage buffers; under an odd RK order that
    swap count is odd, so such batches are refused before any device work."""
    cs = [cases.flume(nx=6, ny=3, length=30.0, width=9.0, seed=5 + r) for r in range(2)]
    keep = [{}, {}]
    meshes = (fvt.Mesh * 2)(*[fvt.mesh_struct(c, k) for c, k in zip(cs, keep)])
    L = fvt.load()
    ctx = C.c_void_p()
    for rk in (1, 3):
        o = fvt.Opts()
        L.storm_fvt_default_opts(C.byref(o))
        o.rkorder, o.steady = rk, 1
        rc = L.storm_fvt_create_batch(2, meshes, C.byref(o), C.byref(ctx))
        assert rc == -1 and not ctx.value, rc
