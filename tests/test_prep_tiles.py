"""Host set-up of the stage kernel's tiles (no GPU): every tile fits its
capacity and every local id the kernel dereferences names the right cell,
for the small cases and the 2M-triangle meander reach."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from storm_amd import cases, fvt
from cases_small import shuffled, small_cases

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("probe") / "libprobe.so")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-o", out,
                           os.path.join(ROOT, "tests", "prep_probe.cpp"),
                           os.path.join(ROOT, "storm_amd", "csrc", "fvt_prep.cpp")])
    lib = C.CDLL(out)
    lib.probe_rsavg.restype = C.c_int
    lib.probe_rsavg.argtypes = [C.POINTER(fvt.Mesh), C.POINTER(fvt.Opts), C.c_int, C.c_uint,
                                C.POINTER(C.c_int), C.c_char_p, C.c_int]
    lib.probe_tiles.restype = C.c_int
    lib.probe_tiles.argtypes = [C.POINTER(fvt.Mesh), C.POINTER(fvt.Opts), C.c_int,
                                C.POINTER(C.c_int), C.c_char_p, C.c_int]
    return lib


def run_probe(lib, case, cap):
    keep = {}
    m = fvt.mesh_struct(case, keep)
    o = fvt.opts_struct(case.opts)
    stats = (C.c_int * 4)()
    msg = C.create_string_buffer(256)
    rc = lib.probe_tiles(C.byref(m), C.byref(o), cap, stats, msg, 256)
    assert rc == 0, msg.value.decode()
    return list(stats)


@pytest.mark.parametrize("cap", [256, 384, 512])
def test_small_case_tiles(probe, cap):
    n_stage = 0
    for name, (case, _) in small_cases().items():
        nt, maxw, nsv, halo = run_probe(probe, case, cap)
        assert 1 <= nt <= case.n_elems and maxw <= cap, name
        n_stage += nsv
    assert n_stage > 0  # some case has a stage string (its vertex list is checked)


def test_meander_tiles(probe):
    case = cases.meander()
    nt, maxw, nsv, halo = run_probe(probe, case, 512)
    # compact Hilbert patches: the halo stays a small share of the tile cells
    assert halo < 0.3 * case.n_elems
    assert maxw <= 512


def run_rsavg_probe(lib, case, nsweep=2):
    keep = {}
    m = fvt.mesh_struct(case, keep)
    o = dict(case.opts)
    o.update(rsavg=True, gradient="lsq")
    stats = (C.c_int * 4)()
    msg = C.create_string_buffer(256)
    rc = lib.probe_rsavg(C.byref(m), C.byref(fvt.opts_struct(o)), nsweep, 7, stats, msg, 256)
    assert rc == 0, msg.value.decode()
    return list(stats)


def test_rsavg_schedules_reproduce_the_serial_sweep(probe):
    """Residual smoothing (solverFVT.f90:244-255) by its level schedule and
    by the shared-memory sweep's tables, simulated on the host step by step,
    against the serial cell-order sweep: equal bit for bit, with no read of
    a stale slot or of a slot written in the same level.  The meander's
    row-by-row numbering gives thousands of narrow levels; a shuffled
    numbering gives a few wide ones (beyond the shared-memory sweep's block,
    so only the level schedule runs)."""
    for name, (case, _) in small_cases().items():
        fast, levels, nfar, width = run_rsavg_probe(probe, case)
        assert fast and levels > 10, name
    fast, levels, nfar, width = run_rsavg_probe(probe, cases.meander(ns=400, length=2000.0,
                                                                     n_bumps=20), 3)
    assert fast and levels > 1000 and width <= 512
    sh = shuffled(cases.meander(ns=100, nn=40, length=1000.0, width=400.0, n_bumps=10), 5)
    fast, levels, nfar, width = run_rsavg_probe(probe, sh)
    assert not fast and width > 512 and levels < 100


def test_rsavg_schedule_of_the_2m_meander(probe):
    fast, levels, nfar, width = run_rsavg_probe(probe, cases.meander())
    assert fast and levels > 8000 and nfar <= 1536
