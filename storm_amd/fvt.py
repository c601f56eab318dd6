"""Python host side of the MI355X FVT step: a ctypes binding of
libstorm_fvt.so (include/storm_fvt.h) shaped like the reference's call
surface.

The reference drives the FVT path with two module-state subroutines,
``CALL dataprepFVT`` and ``CALL solverFVT(iter, ttime, rmax, fn, title)``
(Source_xsy-64bit/solver.f90:50,57).  :class:`FVTSolver` keeps that shape:
``dataprepFVT`` builds the device context from a case (mesh, bed, initial
state, boundary strings, options) and ``solverFVT`` runs the time loop to
MAX ITERATION (or convergence) and ends with the reference's final
``to_vertices``.  ``run`` advances a given number of steps without that
final pass.  Errors raise :class:`StormError` (the role of byebye.f90).

There is no CPU fallback: if the library or a HIP device is missing, every
entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Dict, Optional

import numpy as np

from . import cases

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libstorm_fvt.so")

FLUX = {"ac": 0, "hll": 1, "rusanov": 2, "bgnvc": 3}
FLUXES = {"std": 0, "xsf": 1}  # invisc_fluxes.f90 or invisc_fluxes_xsf.f90
GRAD = {"lsq": 0, "gauss": 1, "hdls": 2}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class StormError(RuntimeError):
    pass


class Mesh(C.Structure):
    _fields_ = [("n_pts", C.c_int32), ("n_elems", C.c_int32), ("n_edges", C.c_int32),
                ("x", _dp), ("y", _dp), ("tri", _ip), ("edge_p", _ip), ("edge_e", _ip),
                ("z_at_vtx", C.c_int32), ("h_at_vtx", C.c_int32), ("u_at_vtx", C.c_int32),
                ("v_at_vtx", C.c_int32), ("cd_at_vtx", C.c_int32),
                ("z", _dp), ("h", _dp), ("u", _dp), ("v", _dp), ("cd", _dp),
                ("n_in", C.c_int32), ("in_type", _ip), ("in_ptr", _ip), ("in_nodes", _ip),
                ("qts_ptr", _ip), ("qts_t", _dp), ("qts_v", _dp),
                ("n_out", C.c_int32), ("out_type", _ip), ("out_ptr", _ip), ("out_nodes", _ip),
                ("hts_ptr", _ip), ("hts_t", _dp), ("hts_v", _dp)]


class Opts(C.Structure):
    _fields_ = [("rkorder", C.c_int32), ("dt_op", C.c_int32), ("steady", C.c_int32),
                ("stop_cr", C.c_int32), ("istop", C.c_int32), ("btype", C.c_int32),
                ("hvalve", C.c_int32), ("limiter", C.c_int32), ("flux", C.c_int32),
                ("gradient", C.c_int32), ("friction", C.c_int32), ("vclip", C.c_int32),
                ("cfl", C.c_double), ("rstop", C.c_double), ("h_dry", C.c_double),
                ("h_wet", C.c_double), ("vceiling", C.c_double),
                ("reorder", C.c_int32), ("block", C.c_int32), ("graph", C.c_int32),
                ("nonfinite_stop", C.c_int32), ("fluxes", C.c_int32),
                ("rsavg", C.c_int32), ("rsavg_n", C.c_int32), ("rsavg_eps", C.c_double)]


EXPORTS = ["storm_fvt_default_opts", "storm_fvt_create", "storm_fvt_create_batch",
           "storm_fvt_reach_status", "storm_fvt_set_state",
           "storm_fvt_set_bc_values", "storm_fvt_run", "storm_fvt_finalize",
           "storm_fvt_get_state", "storm_fvt_get_conserved", "storm_fvt_get_vertex",
           "storm_fvt_get_prep_fields", "storm_fvt_get_output", "storm_fvt_get_c2v_weights",
           "storm_fvt_get_gradients", "storm_fvt_get_gradients_noalloc",
           "storm_fvt_get_bc_values", "storm_fvt_drynwet", "storm_fvt_fvsources",
           "storm_fvt_set_wind", "storm_fvt_set_culverts", "storm_fvt_get_culverts",
           "storm_fvt_device_info",
           "storm_fvt_last_run_seconds", "storm_fvt_last_error", "storm_fvt_destroy",
           "storm_fvt_inspect_molecule"]

STORM_E_STATE = -5
VERTEX_KEYS = ("hvtx", "uvtx", "vvtx", "zetavtx", "cdvtx", "zvtx")

_lib = None


def load(path: str = LIB_PATH):
    """Load libstorm_fvt.so; raises StormError when it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise StormError(f"{path} is not built (run __graft_entry__.build())")
    L = C.CDLL(path)
    L.storm_fvt_default_opts.argtypes = [C.POINTER(Opts)]
    L.storm_fvt_create.restype = C.c_int
    L.storm_fvt_create.argtypes = [C.POINTER(Mesh), C.POINTER(Opts), C.POINTER(C.c_void_p)]
    L.storm_fvt_create_batch.restype = C.c_int
    L.storm_fvt_create_batch.argtypes = [C.c_int32, C.POINTER(Mesh), C.POINTER(Opts),
                                         C.POINTER(C.c_void_p)]
    L.storm_fvt_reach_status.argtypes = [C.c_void_p, _ip, _dp, _dp, C.POINTER(C.c_uint8)]
    L.storm_fvt_set_state.argtypes = [C.c_void_p, _dp, _dp, _dp]
    L.storm_fvt_set_bc_values.argtypes = [C.c_void_p, _dp, _dp]
    L.storm_fvt_run.argtypes = [C.c_void_p, C.c_int32, _ip, _dp, _dp]
    L.storm_fvt_finalize.argtypes = [C.c_void_p]
    L.storm_fvt_get_state.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, C.POINTER(C.c_uint8)]
    L.storm_fvt_get_conserved.argtypes = [C.c_void_p, _dp, _dp, _dp]
    L.storm_fvt_get_vertex.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
    L.storm_fvt_get_prep_fields.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp]
    L.storm_fvt_get_output.argtypes = [C.c_void_p, _dp, _dp]
    L.storm_fvt_get_c2v_weights.restype = C.c_int64
    L.storm_fvt_get_c2v_weights.argtypes = [C.c_void_p, _ip, _ip, _dp, _dp]
    L.storm_fvt_get_gradients.argtypes = [C.c_void_p, _dp, _dp]
    L.storm_fvt_get_gradients_noalloc.argtypes = [C.c_void_p, _dp, _dp]
    L.storm_fvt_drynwet.argtypes = [C.c_void_p, _dp, _dp, _dp, _ip]
    L.storm_fvt_fvsources.argtypes = [C.c_void_p, _dp, _dp, _dp]
    L.storm_fvt_get_bc_values.argtypes = [C.c_void_p, _dp, _dp]
    L.storm_fvt_set_wind.argtypes = [C.c_void_p, C.c_int64, _dp, _dp, _dp]
    L.storm_fvt_set_culverts.argtypes = [C.c_void_p, C.c_int32, _dp, _dp, _ip, _dp, _dp]
    L.storm_fvt_get_culverts.argtypes = [C.c_void_p, _ip, _ip, _dp]
    L.storm_fvt_device_info.argtypes = [C.c_void_p, _ip, _ip, C.POINTER(C.c_void_p)]
    L.storm_fvt_last_run_seconds.restype = C.c_double
    L.storm_fvt_last_run_seconds.argtypes = [C.c_void_p]
    L.storm_fvt_last_error.restype = C.c_char_p
    L.storm_fvt_last_error.argtypes = [C.c_void_p]
    L.storm_fvt_destroy.argtypes = [C.c_void_p]
    L.storm_fvt_inspect_molecule.restype = C.c_int64
    L.storm_fvt_inspect_molecule.argtypes = [C.POINTER(Mesh), _ip, _ip]
    _lib = L
    return L


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def _csr(bs, field):
    ptr = np.zeros(len(bs) + 1, dtype=np.int32)
    for i, b in enumerate(bs):
        ptr[i + 1] = ptr[i] + len(getattr(b, field))
    return ptr


def mesh_struct(case, keep: dict) -> Mesh:
    """C-ABI mesh of a :class:`storm_amd.cases.Case` (arrays kept alive in
    ``keep``)."""
    k = keep
    k["x"] = np.ascontiguousarray(case.x, dtype=np.float64)
    k["y"] = np.ascontiguousarray(case.y, dtype=np.float64)
    k["tri"] = np.ascontiguousarray(case.tri, dtype=np.int32)
    k["ep"] = np.ascontiguousarray(case.edge_p, dtype=np.int32)
    k["ee"] = np.ascontiguousarray(case.edge_e, dtype=np.int32)
    data = cases.inputs(case)
    for nm in ("z", "h", "u", "v", "cd"):
        k[nm] = np.ascontiguousarray(data[nm][0], dtype=np.float64)
    ins, outs = case.inflows, case.outflows
    z1 = np.zeros(1)
    k["in_type"] = np.array([b.kind for b in ins] or [0], dtype=np.int32)
    k["in_ptr"] = _csr(ins, "nodes")
    k["in_nodes"] = np.concatenate([b.nodes for b in ins] or [z1]).astype(np.int32)
    k["qts_ptr"] = _csr(ins, "t")
    k["qts_t"] = np.concatenate([b.t for b in ins] or [z1]).astype(np.float64)
    k["qts_v"] = np.concatenate([b.v for b in ins] or [z1]).astype(np.float64)
    k["out_type"] = np.array([b.kind for b in outs] or [0], dtype=np.int32)
    k["out_ptr"] = _csr(outs, "nodes")
    k["out_nodes"] = np.concatenate([b.nodes for b in outs] or [z1]).astype(np.int32)
    k["hts_ptr"] = _csr(outs, "t")
    k["hts_t"] = np.concatenate([b.t for b in outs] or [z1]).astype(np.float64)
    k["hts_v"] = np.concatenate([b.v for b in outs] or [z1]).astype(np.float64)
    d, i = C.c_double, C.c_int32
    return Mesh(n_pts=case.n_pts, n_elems=case.n_elems, n_edges=case.n_edges,
                x=_p(k["x"], d), y=_p(k["y"], d), tri=_p(k["tri"], i),
                edge_p=_p(k["ep"], i), edge_e=_p(k["ee"], i),
                z_at_vtx=int(data["z"][1]), h_at_vtx=int(data["h"][1]),
                u_at_vtx=int(data["u"][1]), v_at_vtx=int(data["v"][1]),
                cd_at_vtx=int(data["cd"][1]),
                z=_p(k["z"], d), h=_p(k["h"], d), u=_p(k["u"], d), v=_p(k["v"], d),
                cd=_p(k["cd"], d), n_in=len(ins), in_type=_p(k["in_type"], i),
                in_ptr=_p(k["in_ptr"], i), in_nodes=_p(k["in_nodes"], i),
                qts_ptr=_p(k["qts_ptr"], i), qts_t=_p(k["qts_t"], d), qts_v=_p(k["qts_v"], d),
                n_out=len(outs), out_type=_p(k["out_type"], i), out_ptr=_p(k["out_ptr"], i),
                out_nodes=_p(k["out_nodes"], i), hts_ptr=_p(k["hts_ptr"], i),
                hts_t=_p(k["hts_t"], d), hts_v=_p(k["hts_v"], d))


def opts_struct(o: Dict[str, object]) -> Opts:
    """Options in the reference's meaning (options.f90); h_wet follows
    read_options.f90:441 (C THRESHOLD times H THRESHOLD)."""
    s = Opts()
    load().storm_fvt_default_opts(C.byref(s))
    h_dry = float(o.get("h_dry", 0.001))
    s.rkorder = int(o.get("rkorder", 2))
    s.dt_op = int(o.get("dt_op", 0))
    s.steady = int(bool(o.get("steady", True)))
    s.stop_cr = int(o.get("stop_cr", 0))
    s.istop = int(o.get("istop", 0))
    s.btype = int(o.get("btype", 0))
    s.hvalve = int(bool(o.get("hvalve", True)))
    s.limiter = int(o.get("limiter", 2))
    s.flux = FLUX[str(o.get("flux", "ac"))]
    s.gradient = GRAD[str(o.get("gradient", "lsq"))]
    s.friction = int(o.get("friction", 0))
    s.vclip = int(bool(o.get("vclip", False)))
    s.cfl = float(o.get("cfl", 1.0))
    s.rstop = float(o.get("rstop", 1e-4))
    s.h_dry = h_dry
    s.h_wet = max(1.0, float(o.get("c_threshold", 1.0))) * h_dry
    s.vceiling = float(o.get("vceiling", 35.0))
    s.reorder = int(o.get("reorder", 1))
    s.block = int(o.get("block", 0))
    s.graph = int(o.get("graph", 0))
    s.nonfinite_stop = int(bool(o.get("nonfinite_stop", False)))
    s.fluxes = FLUXES[str(o.get("fluxes", "std"))]
    s.rsavg = int(bool(o.get("rsavg", False)))
    s.rsavg_n = int(o.get("rsavg_n", 2))
    s.rsavg_eps = float(o.get("rsavg_eps", 0.5))
    return s


class FVTSolver:
    """One FVT context on the current HIP device."""

    def __init__(self, case, opts: Optional[Dict[str, object]] = None):
        self.dataprepFVT(case, opts)

    # -- reference call surface -----------------------------------------
    def dataprepFVT(self, case, opts: Optional[Dict[str, object]] = None):
        """dataprepFVT.f90: set up mesh tables, coefficients and state."""
        L = load()
        o = dict(case.opts)
        if opts:
            o.update(opts)
        self.opts = o
        self._keep = {}
        m = mesh_struct(case, self._keep)
        self._o = opts_struct(o)
        h = C.c_void_p()
        rc = L.storm_fvt_create(C.byref(m), C.byref(self._o), C.byref(h))
        if rc != 0 or not h.value:
            raise StormError(f"storm_fvt_create failed ({rc})")
        self.h = h
        self.N, self.P = case.n_elems, case.n_pts
        self.n_in, self.n_out = len(case.inflows), len(case.outflows)
        self.ncv = 0
        # windforcing and ioculvert's data, read before dataprepFVT's end
        if getattr(case, "wind", None) is not None:
            self.set_wind(*case.wind)
        if getattr(case, "culverts", None):
            self.set_culverts(case.culverts)

    def solverFVT(self):
        """solverFVT.f90: run to MAX ITERATION (or steady convergence), then
        the final to_vertices.  Returns (iter, ttime, rmax)."""
        r = self.run(int(self.opts.get("istop", 0)))
        self.finalize()
        return r

    # -- finer control ------------------------------------------------------
    def _chk(self, rc):
        if rc != 0:
            msg = load().storm_fvt_last_error(self.h)
            raise StormError(f"storm_fvt error {rc}: {msg.decode() if msg else ''}")

    def run(self, nsteps: int):
        it = C.c_int32(0)
        tt = C.c_double(0)
        rm = C.c_double(0)
        self._chk(load().storm_fvt_run(self.h, int(nsteps), C.byref(it), C.byref(tt), C.byref(rm)))
        return it.value, tt.value, rm.value

    def finalize(self):
        self._chk(load().storm_fvt_finalize(self.h))

    def set_state(self, h, u, v):
        h = np.ascontiguousarray(h, dtype=np.float64)
        u = np.ascontiguousarray(u, dtype=np.float64)
        v = np.ascontiguousarray(v, dtype=np.float64)
        self._chk(load().storm_fvt_set_state(self.h, _p(h, C.c_double), _p(u, C.c_double),
                                             _p(v, C.c_double)))

    def set_bc_values(self, qin=None, hbc=None):
        q = None if qin is None else np.ascontiguousarray(qin, dtype=np.float64)
        hb = None if hbc is None else np.ascontiguousarray(hbc, dtype=np.float64)
        self._chk(load().storm_fvt_set_bc_values(
            self.h, None if q is None else _p(q, C.c_double), None if hb is None else _p(hb, C.c_double)))

    def state(self) -> Dict[str, np.ndarray]:
        N = self.N
        h, u, v = np.zeros(N), np.zeros(N), np.zeros(N)
        phi = np.zeros(3 * N)
        dry = np.zeros(N, dtype=np.uint8)
        self._chk(load().storm_fvt_get_state(self.h, _p(h, C.c_double), _p(u, C.c_double),
                                             _p(v, C.c_double), _p(phi, C.c_double),
                                             _p(dry, C.c_uint8)))
        qh, qu, qv = np.zeros(N), np.zeros(N), np.zeros(N)
        self._chk(load().storm_fvt_get_conserved(self.h, _p(qh, C.c_double), _p(qu, C.c_double),
                                                 _p(qv, C.c_double)))
        out = dict(h=h, u=u, v=v, qh=qh, qu=qu, qv=qv, phi1=phi[:N], phi2=phi[N:2 * N],
                   phi3=phi[2 * N:], dry=dry.astype(np.int32))
        out.update(self.vertex(missing_ok=True))
        return out

    def vertex(self, missing_ok: bool = False) -> Dict[str, np.ndarray]:
        """hvtx, uvtx, vvtx, zetavtx of the last to_vertices (dataprepFVT's
        before any step, the final ones after finalize).  After steps with no
        finalize only hvtx is current: with missing_ok the others are left
        out (else StormError)."""
        P = self.P
        a = {k: np.zeros(P) for k in ("hvtx", "uvtx", "vvtx", "zetavtx")}
        L = load()
        rc = L.storm_fvt_get_vertex(self.h, *[_p(a[k], C.c_double) for k in a])
        if rc == 0:
            return a
        if rc == STORM_E_STATE and missing_ok:
            rc = L.storm_fvt_get_vertex(self.h, _p(a["hvtx"], C.c_double), None, None, None)
            return {"hvtx": a["hvtx"]} if rc == 0 else {}
        self._chk(rc)

    def prep_fields(self) -> Dict[str, np.ndarray]:
        """The module copies dataprepFVT leaves: z, zeta, cd (cells), zvtx,
        cdvtx (vertices); the vertex depth, velocities and surface come from
        vertex() before the first step."""
        N, P = self.N, self.P
        a = dict(z=np.zeros(N), zvtx=np.zeros(P), zeta=np.zeros(N), cd=np.zeros(N),
                 cdvtx=np.zeros(P))
        self._chk(load().storm_fvt_get_prep_fields(self.h, *[_p(a[k], C.c_double) for k in
                                                             ("z", "zvtx", "zeta", "cd", "cdvtx")]))
        return a

    def output(self) -> Dict[str, np.ndarray]:
        """After finalize: cdvtx = ctr2vtx(cd) and e_est (error_est(2))."""
        cd, ee = np.zeros(self.P), np.zeros(self.N)
        self._chk(load().storm_fvt_get_output(self.h, _p(cd, C.c_double), _p(ee, C.c_double)))
        return dict(cdvtx=cd, e_est=ee)

    def c2v_weights(self):
        """(ptr, cells, w, wid): the centre-to-vertex weights as CSR."""
        L = load()
        n = L.storm_fvt_get_c2v_weights(self.h, None, None, None, None)
        if n < 0:
            self._chk(int(n))
        ptr = np.zeros(self.P + 1, dtype=np.int32)
        cells = np.zeros(max(n, 1), dtype=np.int32)
        w, wid = np.zeros(max(n, 1)), np.zeros(max(n, 1))
        L.storm_fvt_get_c2v_weights(self.h, _p(ptr, C.c_int32), _p(cells, C.c_int32),
                                    _p(w, C.c_double), _p(wid, C.c_double))
        return ptr, cells[:n], w[:n], wid[:n]

    def gradients(self):
        """delh of the last stage (x, y), what error_est reads after the loop."""
        gx, gy = np.zeros(self.N), np.zeros(self.N)
        self._chk(load().storm_fvt_get_gradients(self.h, _p(gx, C.c_double), _p(gy, C.c_double)))
        return gx, gy

    def drynwet(self, z, h):
        """drynwet + part_dry per triangle for vertex bed z and depth h:
        (zout (N, 3), pd (N, 4))."""
        z = np.ascontiguousarray(z, dtype=np.float64)
        h = np.ascontiguousarray(h, dtype=np.float64)
        zo = np.zeros(3 * self.N)
        pd = np.zeros(4 * self.N, dtype=np.int32)
        self._chk(load().storm_fvt_drynwet(self.h, _p(z, C.c_double), _p(h, C.c_double),
                                           _p(zo, C.c_double), _p(pd, C.c_int32)))
        return zo.reshape(self.N, 3), pd.reshape(self.N, 4)

    def fvsources(self, z, h):
        """fvsources' source (3, P) for vertex bed z and depth h."""
        z = np.ascontiguousarray(z, dtype=np.float64)
        h = np.ascontiguousarray(h, dtype=np.float64)
        src = np.zeros(3 * self.P)
        self._chk(load().storm_fvt_fvsources(self.h, _p(z, C.c_double), _p(h, C.c_double),
                                             _p(src, C.c_double)))
        return src.reshape(3, self.P)

    def set_wind(self, c, wx, wy):
        """Wind data (windforcing.f90): one value per cell or per vertex."""
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (c, wx, wy)]
        if not (len(a[0]) == len(a[1]) == len(a[2])):
            raise StormError("wind arrays differ in length")
        self._chk(load().storm_fvt_set_wind(self.h, len(a[0]), *[_p(x, C.c_double) for x in a]))

    def clear_wind(self):
        self._chk(load().storm_fvt_set_wind(self.h, 0, None, None, None))

    def set_culverts(self, culverts):
        """Culverts (cases.Culvert records, ioculvert.f90's data)."""
        n = len(culverts)
        if n == 0:
            self._chk(load().storm_fvt_set_culverts(self.h, 0, None, None, None, None, None))
            self.ncv = 0
            return
        pin = np.array([cv.inlet for cv in culverts], dtype=np.float64).reshape(-1)
        pout = np.array([cv.outlet for cv in culverts], dtype=np.float64).reshape(-1)
        ntab = np.array([len(cv.head) for cv in culverts], dtype=np.int32)
        head = np.concatenate([np.asarray(cv.head, dtype=np.float64) for cv in culverts])
        q = np.concatenate([np.asarray(cv.q, dtype=np.float64) for cv in culverts])
        self._chk(load().storm_fvt_set_culverts(self.h, n, _p(pin, C.c_double),
                                                _p(pout, C.c_double), _p(ntab, C.c_int32),
                                                _p(head, C.c_double), _p(q, C.c_double)))
        self.ncv = n

    def culverts(self):
        """cvtrigin, cvtrigout (0-based input cells) and the latest stage's
        cvsrc."""
        n = self.ncv
        cin = np.zeros(max(n, 1), dtype=np.int32)
        cout = np.zeros(max(n, 1), dtype=np.int32)
        src = np.zeros(max(n, 1))
        self._chk(load().storm_fvt_get_culverts(self.h, _p(cin, C.c_int32), _p(cout, C.c_int32),
                                                _p(src, C.c_double)))
        return cin[:n], cout[:n], src[:n]

    def bc_values(self):
        """qin, hbc of the latest step (get_time_bcs' module values)."""
        q, hb = np.zeros(max(self.n_in, 1)), np.zeros(max(self.n_out, 1))
        self._chk(load().storm_fvt_get_bc_values(self.h, _p(q, C.c_double), _p(hb, C.c_double)))
        return q[:self.n_in], hb[:self.n_out]

    def last_run_seconds(self) -> float:
        return float(load().storm_fvt_last_run_seconds(self.h))

    def perm(self) -> np.ndarray:
        p = np.zeros(self.N, dtype=np.int32)
        self._chk(load().storm_fvt_device_info(self.h, None, _p(p, C.c_int32), None))
        return p

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            load().storm_fvt_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchSolver(FVTSolver):
    """An ensemble of independent reaches advanced together in one context
    (storm_fvt_create_batch).  Every reach gets the same options; state()
    returns the reaches' arrays laid end to end, split() cuts them apart."""

    def __init__(self, cases_, opts: Optional[Dict[str, object]] = None):
        L = load()
        o = dict(cases_[0].opts)
        if opts:
            o.update(opts)
        self.opts = o
        self._keep = []
        meshes = (Mesh * len(cases_))()
        for r, c in enumerate(cases_):
            k = {}
            meshes[r] = mesh_struct(c, k)
            self._keep.append(k)
        self._meshes = meshes
        self._o = opts_struct(o)
        h = C.c_void_p()
        rc = L.storm_fvt_create_batch(len(cases_), meshes, C.byref(self._o), C.byref(h))
        if rc != 0 or not h.value:
            raise StormError(f"storm_fvt_create_batch failed ({rc})")
        self.h = h
        self.nr = len(cases_)
        self.cell_ptr = np.cumsum([0] + [c.n_elems for c in cases_])
        self.pts_ptr = np.cumsum([0] + [c.n_pts for c in cases_])
        self.N, self.P = int(self.cell_ptr[-1]), int(self.pts_ptr[-1])
        self.n_in = sum(len(c.inflows) for c in cases_)
        self.n_out = sum(len(c.outflows) for c in cases_)
        self.ncv = 0
        winds = [getattr(c, "wind", None) for c in cases_]
        if any(w is not None for w in winds):
            # the reaches' wind data end to end, all at cells or all at vertices
            at_c = [w is not None and len(w[0]) == c.n_elems for w, c in zip(winds, cases_)]
            at_v = [w is not None and len(w[0]) == c.n_pts for w, c in zip(winds, cases_)]
            if all(at_c):
                self.set_wind(*[np.concatenate([w[k] for w in winds]) for k in range(3)])
            elif all(at_v):
                self.set_wind(*[np.concatenate([w[k] for w in winds]) for k in range(3)])
            else:
                raise StormError("a batch takes wind data for every reach, all at cells or all "
                                 "at vertices")
        if any(getattr(c, "culverts", None) for c in cases_):
            raise StormError("culverts need a context of one reach")

    def status(self):
        """Per reach: (iter, ttime, rmax, done) arrays."""
        it = np.zeros(self.nr, dtype=np.int32)
        tt, rm = np.zeros(self.nr), np.zeros(self.nr)
        dn = np.zeros(self.nr, dtype=np.uint8)
        self._chk(load().storm_fvt_reach_status(self.h, _p(it, C.c_int32), _p(tt, C.c_double),
                                                _p(rm, C.c_double), _p(dn, C.c_uint8)))
        return it, tt, rm, dn.astype(bool)

    def split(self, st: Dict[str, np.ndarray]):
        """Cut a state() dictionary into one dictionary per reach."""
        out = []
        for r in range(self.nr):
            a, b = self.cell_ptr[r], self.cell_ptr[r + 1]
            pa, pb = self.pts_ptr[r], self.pts_ptr[r + 1]
            out.append({k: (v[pa:pb] if k in VERTEX_KEYS else v[a:b]) for k, v in st.items()})
        return out


def inspect_molecule(case):
    """The least-squares molecule (elemnt_db.f90:42-128) the library builds,
    without a device."""
    keep = {}
    m = mesh_struct(case, keep)
    L = load()
    n = L.storm_fvt_inspect_molecule(C.byref(m), None, None)
    if n < 0:
        raise StormError(f"inspect_molecule failed ({n})")
    ptr = np.zeros(case.n_elems + 1, dtype=np.int32)
    cells = np.zeros(max(n, 1), dtype=np.int32)
    L.storm_fvt_inspect_molecule(C.byref(m), _p(ptr, C.c_int32), _p(cells, C.c_int32))
    return ptr, cells[:n]
