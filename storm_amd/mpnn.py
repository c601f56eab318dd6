"""Python host side of the warm-start network (libstorm_mpnn.so,
include/storm_mpnn.h): an edge message-passing network over the cells of a
reach, trained and run by hand-written BF16 MFMA kernels.  There is no CPU
fallback: a missing library raises."""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libstorm_mpnn.so")
EXPORTS = ["storm_mpnn_create", "storm_mpnn_destroy", "storm_mpnn_num_params",
           "storm_mpnn_get_params", "storm_mpnn_set_params", "storm_mpnn_set_graph",
           "storm_mpnn_forward", "storm_mpnn_backward", "storm_mpnn_get_grads",
           "storm_mpnn_adamw_step", "storm_mpnn_set_lr", "storm_mpnn_last_error",
           "storm_mpnn_gvf_profile"]

_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int32)


class Cfg(C.Structure):
    _fields_ = [("layers", C.c_int32), ("hidden", C.c_int32), ("f_node", C.c_int32),
                ("f_edge", C.c_int32), ("f_out", C.c_int32), ("lr", C.c_float),
                ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float),
                ("weight_decay", C.c_float), ("seed", C.c_uint64)]


class MpnnError(RuntimeError):
    pass


_lib = None


def load(path: str = LIB_PATH):
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise MpnnError(f"{path} is not built (run __graft_entry__.build())")
    L = C.CDLL(path)
    L.storm_mpnn_create.argtypes = [C.POINTER(Cfg), C.POINTER(C.c_void_p)]
    L.storm_mpnn_destroy.argtypes = [C.c_void_p]
    L.storm_mpnn_num_params.restype = C.c_int64
    L.storm_mpnn_num_params.argtypes = [C.c_void_p]
    L.storm_mpnn_get_params.argtypes = [C.c_void_p, _fp]
    L.storm_mpnn_set_params.argtypes = [C.c_void_p, _fp]
    L.storm_mpnn_set_graph.argtypes = [C.c_void_p, C.c_int32, C.c_int32, _ip, _ip, _fp, _fp]
    L.storm_mpnn_forward.argtypes = [C.c_void_p, _fp]
    L.storm_mpnn_backward.argtypes = [C.c_void_p, _fp, _fp, C.POINTER(C.c_double)]
    L.storm_mpnn_get_grads.argtypes = [C.c_void_p, _fp]
    L.storm_mpnn_adamw_step.argtypes = [C.c_void_p]
    L.storm_mpnn_set_lr.argtypes = [C.c_void_p, C.c_float]
    L.storm_mpnn_last_error.restype = C.c_char_p
    L.storm_mpnn_last_error.argtypes = [C.c_void_p]
    L.storm_mpnn_gvf_profile.argtypes = [C.c_double] * 6 + [C.c_int32, C.POINTER(C.c_double)]
    _lib = L
    return L


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def gvf_profile(q: float, S: float, manning: float, h0: float, h_floor: float, ds: float,
                n_steps: int) -> np.ndarray:
    """storm_mpnn_gvf_profile: the RK4 depths of the steady 1-D gradually
    varied flow at n_steps + 1 points ds apart, upstream from h0 (host
    code)."""
    h = np.empty(n_steps + 1, dtype=np.float64)
    if load().storm_mpnn_gvf_profile(q, S, manning, h0, h_floor, ds, n_steps,
                                     _p(h, C.c_double)) != 0:
        raise MpnnError("storm_mpnn_gvf_profile: bad arguments")
    return h


class Net:
    """The network and its AdamW state on the current HIP device."""

    def __init__(self, layers=6, f_node=16, f_edge=16, f_out=16, lr=1e-3, beta1=0.9,
                 beta2=0.999, eps=1e-8, weight_decay=1e-4, seed=0):
        L = load()
        self.cfg = Cfg(layers, 64, f_node, f_edge, f_out, lr, beta1, beta2, eps, weight_decay, seed)
        h = C.c_void_p()
        if L.storm_mpnn_create(C.byref(self.cfg), C.byref(h)) != 0 or not h.value:
            raise MpnnError("storm_mpnn_create failed")
        self.h = h
        self.n_params = int(L.storm_mpnn_num_params(h))
        self.n = 0

    def _chk(self, rc):
        if rc != 0:
            msg = load().storm_mpnn_last_error(self.h)
            raise MpnnError(msg.decode() if msg else "storm_mpnn error")

    def set_graph(self, src, dst, node_feat, edge_feat):
        self._g = (np.ascontiguousarray(src, dtype=np.int32), np.ascontiguousarray(dst, dtype=np.int32),
                   np.ascontiguousarray(node_feat, dtype=np.float32),
                   np.ascontiguousarray(edge_feat, dtype=np.float32))
        s, d, nf, ef = self._g
        self.n = nf.shape[0]
        self._chk(load().storm_mpnn_set_graph(self.h, self.n, s.shape[0], _p(s, C.c_int32),
                                              _p(d, C.c_int32), _p(nf, C.c_float),
                                              _p(ef, C.c_float)))

    def forward(self, copy: bool = True) -> Optional[np.ndarray]:
        """Forward pass; returns y [n][f_out] (None with copy=False, which
        leaves the outputs on the device for backward)."""
        if not copy:
            self._chk(load().storm_mpnn_forward(self.h, None))
            return None
        y = np.zeros((self.n, self.cfg.f_out), dtype=np.float32)
        self._chk(load().storm_mpnn_forward(self.h, _p(y, C.c_float)))
        return y

    def backward(self, target, w=None) -> float:
        t = np.ascontiguousarray(target, dtype=np.float32)
        ww = None if w is None else np.ascontiguousarray(w, dtype=np.float32)
        loss = C.c_double(0)
        self._chk(load().storm_mpnn_backward(self.h, _p(t, C.c_float),
                                             None if ww is None else _p(ww, C.c_float),
                                             C.byref(loss)))
        return loss.value

    def step(self):
        self._chk(load().storm_mpnn_adamw_step(self.h))

    def set_lr(self, lr: float):
        self._chk(load().storm_mpnn_set_lr(self.h, float(lr)))

    def params(self) -> np.ndarray:
        p = np.zeros(self.n_params, dtype=np.float32)
        self._chk(load().storm_mpnn_get_params(self.h, _p(p, C.c_float)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.shape[0] == self.n_params
        self._chk(load().storm_mpnn_set_params(self.h, _p(p, C.c_float)))

    def grads(self) -> np.ndarray:
        g = np.zeros(self.n_params, dtype=np.float32)
        self._chk(load().storm_mpnn_get_grads(self.h, _p(g, C.c_float)))
        return g

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            load().storm_mpnn_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
