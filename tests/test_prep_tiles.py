"""Host set-up of the stage kernel's tiles (no GPU): every tile fits its
capacity and every local id the kernel dereferences names the right cell,
for the small cases and the 2M-triangle meander reach."""
import copy
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
@pytest.mark.parametrize("reorder", [1, 2, 0])
def test_small_case_tiles(probe, cap, reorder):
    n_stage = 0
    for name, (case, _) in small_cases().items():
        case = copy.deepcopy(case)
        case.opts["reorder"] = reorder
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


ASAN_CHILD = r"""
import ctypes as C, sys
sys.path.insert(0, sys.argv[2]); sys.path.insert(0, sys.argv[2] + "/tests")
import numpy as np
from storm_amd import cases, fvt
from cases_small import small_cases, source_cases, _meander_culverts
import test_failure_injection as F
lib = C.CDLL(sys.argv[1])
lib.probe_tiles.argtypes = [C.POINTER(fvt.Mesh), C.POINTER(fvt.Opts), C.c_int,
                            C.POINTER(C.c_int), C.c_char_p, C.c_int]
lib.probe_rsavg.argtypes = [C.POINTER(fvt.Mesh), C.POINTER(fvt.Opts), C.c_int, C.c_uint,
                            C.POINTER(C.c_int), C.c_char_p, C.c_int]
dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int)
lib.probe_sources.argtypes = [C.POINTER(fvt.Mesh), C.POINTER(fvt.Opts), C.c_int64, dp, dp, dp,
                              C.c_int, dp, dp, ip, ip]
def p(a, t):
    return a.ctypes.data_as(C.POINTER(t))
n = 0
for name, (case, _) in list(small_cases().items()) + list(source_cases().items()):
    keep = {}
    m = fvt.mesh_struct(case, keep)
    o = dict(case.opts); o["rsavg"] = True
    st = (C.c_int * 4)(); msg = C.create_string_buffer(256)
    assert lib.probe_tiles(C.byref(m), C.byref(fvt.opts_struct(o)), 256, st, msg, 256) == 0, name
    if o.get("gradient") != "hdls":
        assert lib.probe_rsavg(C.byref(m), C.byref(fvt.opts_struct(o)), 2, 3, st, msg, 256) == 0, name
    w = [np.linspace(0.001, 0.002, case.n_pts), np.ones(case.n_pts), np.ones(case.n_pts)]
    cv = _meander_culverts() if name.startswith("meander") else []
    pin = np.array([c.inlet for c in cv] or [[0.0, 0, 0]]).reshape(-1)
    pout = np.array([c.outlet for c in cv] or [[0.0, 0]]).reshape(-1)
    rc = np.zeros(2, dtype=np.int32); cells = np.zeros(2 * max(len(cv), 1), dtype=np.int32)
    assert lib.probe_sources(C.byref(m), C.byref(fvt.opts_struct(o)), case.n_pts, p(w[0], C.c_double),
                             p(w[1], C.c_double), p(w[2], C.c_double), len(cv), p(pin, C.c_double),
                             p(pout, C.c_double), p(rc, C.c_int), p(cells, C.c_int)) == 0, name
    n += 1
# the failure paths of the set-up
bad = []
c = F.base_case(); c.tri = c.tri.copy(); c.tri[3, 1] = c.n_pts + 7; bad.append(c)
c = F.base_case(); c.tri = c.tri.copy(); c.tri[5, [1, 2]] = c.tri[5, [2, 1]]; bad.append(F.rebuilt(c))
c = F.base_case(); c.edge_e = c.edge_e.copy(); c.edge_e[2, 0] = c.n_elems + 3; bad.append(c)
c = F.base_case(); b = c.inflows[0]; b.nodes = np.concatenate([b.nodes[:1], b.nodes[2:]]); bad.append(c)
bad.append(cases.dambreak(nq=1, side=10.0, seed=1, n_bumps=1, x_dam=5.0, h_res=1.0))
for case in bad:
    keep = {}
    m = fvt.mesh_struct(case, keep)
    st = (C.c_int * 4)(); msg = C.create_string_buffer(256)
    assert lib.probe_tiles(C.byref(m), C.byref(fvt.opts_struct(case.opts)), 256, st, msg, 256) < 0
print("asan probes ok", n, len(bad))
"""


def test_prep_under_address_sanitizer(tmp_path):
    """fvt_prep.cpp (and the probes) built with AddressSanitizer and UBSan
    (g++, host code only), run over every small case (tiles, residual
    smoothing tables, wind and culvert set-up, dataprep copies) and the
    broken meshes of the failure-injection tests."""
    import sys
    gxx = "g++"
    asan = subprocess.run([gxx, "-print-file-name=libasan.so"], capture_output=True,
                          text=True).stdout.strip()
    if not os.path.isabs(asan) or not os.path.exists(asan):
        pytest.skip("libasan not available")
    so = str(tmp_path / "libprobe_asan.so")
    subprocess.check_call([gxx, "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined",
                           "-fno-omit-frame-pointer", "-fno-sanitize-recover=undefined",
                           "-shared", "-fPIC", "-o", so,
                           os.path.join(ROOT, "tests", "prep_probe.cpp"),
                           os.path.join(ROOT, "storm_amd", "csrc", "fvt_prep.cpp")])
    child = tmp_path / "child.py"
    child.write_text(ASAN_CHILD)
    env = dict(os.environ, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1")
    r = subprocess.run([sys.executable, str(child), so, ROOT], capture_output=True, text=True,
                       env=env, timeout=600)
    assert r.returncode == 0 and "asan probes ok" in r.stdout, (r.stdout[-2000:], r.stderr[-4000:])
    assert "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr
