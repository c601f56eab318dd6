"""Learned warm start of the steady FVT solver.

An edge message-passing network (storm_amd/mpnn.py over libstorm_mpnn.so:
BF16 MFMA GEMMs with FP32 accumulation, hand-written forward, backward and
AdamW kernels) predicts an initial cell state (h, hu, hv) of a reach from its
bed elevation, Manning n, and boundary discharge and stage.  The FP64 solver
then starts from the prediction and runs to the same residual as a cold
start, so it both backs up and checks the prediction: a poor
prediction costs steps, never accuracy.  No counterpart exists in the
reference, which starts from the user's initial conditions
(Source_xsy-64bit/dataprepFVT.f90:90-130).

Graph: one node per cell, one directed edge each way across every interior
side.

Reach frame.  The inputs are put in a frame of the reach: the axis ``e``
runs from the midpoint of the inflow string to the midpoint of the outflow
string, ``dist`` is a cell's distance to the outlet along it, ``H`` is the
outflow stage, ``q`` the inflow discharge per metre of inflow string, and
``zfit = a + S*dist`` the least-squares line of the cell bed elevations
against ``dist`` (the fitted bed slope S).  The frame assumes a reach whose
axis is straight (the flumes of the ensemble generator); a curved reach
would need a curvilinear coordinate in its place.

Node features (16), set 2: bed relative to the outlet stage,
bed relative to the fitted line, logs of Manning n, q, the outlet depth h_out
(outlet stage over the fitted bed) and the fitted slope, distance to the
outlet and to the inlet, inflow / outflow / wall flags, cell size, the
dimensionless distance S*dist/h_out, the log of the outlet Froude number
squared q^2/(g h_out^3), and the depth of the outlet stage over the fitted
bed at the cell.  Set 1 (checkpoints trained before set 2) has n, q, h_out
and S on linear scales, the discharge and stage of the flagged boundary cells
in place of the two dimensionless groups, and a zero last.  The logs and the
dimensionless groups let the layers form the power laws a steady profile
depends on as sums.  Edge features (16): the centroid offset in the reach
frame and its length, the bed step and the step of the bed relative to the
fitted line, a constant, ten zeros.

Set 3 (the default) adds a non-learned base: the steady 1-D gradually
varied flow over the fitted bed from the outlet stage (``gvf_depth``), with
the Manning coefficient scaled by the steady mode's friction factor
(``friction_multiplier``: in steady runs every cell's implicit friction acts
with the local time step of the reach's last cell).  Its surface over the
plane is the zero output and node feature 15, and the log of each cell's
friction factor dt_last/dt_i at the edge's two cells are edge features 6
and 7.

Outputs, each scaled to order one:
  y0 = (zeta - H - S*dist - base) / 0.05 m
                                     water surface relative to the base: the
                                     plane through the outlet stage parallel
                                     to the fitted bed (sets 1, 2), or the
                                     backwater surface (set 3)
  y1 = (hu_e / q - 1) / 0.04         discharge along the axis
  y2 = (hv_e / q) / 0.04             discharge across the axis
The scales weight a relative error of the depth (about 4 m) twice as much as
the same relative error of the discharge, whose steady values carry the
inflow fixed point's tolerance of 0.5% (inflowQbyH.f90:120-121).
The other 13 outputs are trained to zero.  ``decode`` turns them back into
h = max(zeta - zc, 0), hu, hv and the u, v that storm_fvt_set_state takes.
With all outputs zero the decode is the base with uniform discharge: the
plane (``baseline_state``) or the backwater start (``gvf_state``), the
non-learned guesses the network improves on.
"""
from __future__ import annotations

import dataclasses
import json
import os
import time
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import cases as _cases

F_NODE = 16
F_EDGE = 16
F_OUT = 16
D_SCALE = 5.0       # m, depth/elevation scale of the features
Y0_SCALE = 0.05     # m, water-surface output
Q_SCALE = 0.04      # relative discharge outputs


@dataclasses.dataclass
class Frame:
    """The reach frame and the graph of one reach."""
    src: np.ndarray
    dst: np.ndarray
    node: np.ndarray        # (N, 16) float32
    edge: np.ndarray        # (E, 16) float32
    zc: np.ndarray          # cell bed elevation
    dist: np.ndarray        # distance to the outlet along the axis
    e: np.ndarray           # unit axis (2,)
    H: float
    q: float
    S: float
    a: float
    order: Optional[np.ndarray] = None   # cells along the axis (set by window)
    rank: Optional[np.ndarray] = None
    eperm: Optional[np.ndarray] = None   # edges by their source's rank (set by window)
    ers: Optional[np.ndarray] = None     # the source ranks in that order
    # water surface of the zero output relative to the plane (feature set 3:
    # the 1-D backwater profile's; 0 otherwise)
    base: Optional[np.ndarray] = None


def _string_geometry(case, b):
    xs, ys = case.x[b.nodes], case.y[b.nodes]
    length = float(np.hypot(np.diff(xs), np.diff(ys)).sum())
    mid = np.array([xs.mean(), ys.mean()])
    return length, mid


def _string_cells(case, nodes) -> np.ndarray:
    """Cells with a side on the string (consecutive nodes of the chain)."""
    bk = np.nonzero(case.edge_e[:, 1] < 0)[0]  # string sides are boundary sides
    nodes = np.asarray(nodes, dtype=np.int64)
    if bk.size == 0 or nodes.size < 2:
        return np.zeros(0, dtype=np.int64)
    ep = case.edge_p[bk].astype(np.int64)
    P = int(max(ep.max(), nodes.max())) + 1

    def key(a, b):  # an undirected side as one integer
        return np.minimum(a, b) * P + np.maximum(a, b)

    kb = key(ep[:, 0], ep[:, 1])
    order = np.argsort(kb, kind="stable")
    kb = kb[order]
    ks = key(nodes[:-1], nodes[1:])
    pos = np.minimum(np.searchsorted(kb, ks), kb.size - 1)
    hit = kb[pos] == ks
    return np.unique(case.edge_e[bk[order[pos[hit]]], 0].astype(np.int64))


FEATURES = 3  # node feature set of new networks


def frame(case, features: int = FEATURES) -> Frame:
    """Graph, features and reach frame of a case (one inflow string with a
    discharge and one outflow string with a stage)."""
    if len(case.inflows) != 1 or len(case.outflows) != 1:
        raise ValueError("the warm start needs exactly one inflow and one outflow string")
    bi, bo = case.inflows[0], case.outflows[0]
    if bi.kind not in (1, 5) or bo.kind != 1 or len(bo.v) == 0:
        raise ValueError("the warm start needs a discharge inflow (vtype 1/5) and a stage outflow")
    w_in, m_in = _string_geometry(case, bi)
    _, m_out = _string_geometry(case, bo)
    ax = m_out - m_in
    e = ax / np.hypot(*ax)
    tri = case.tri
    xc = case.x[tri].sum(axis=1) / 3.0
    yc = case.y[tri].sum(axis=1) / 3.0
    zc = case.zvtx[tri].sum(axis=1) / 3.0
    dist = (m_out[0] - xc) * e[0] + (m_out[1] - yc) * e[1]
    span = float(np.hypot(*ax))
    H = float(bo.v[0])
    q = float(bi.v[0]) / w_in
    A = np.vstack([np.ones_like(dist), dist]).T
    a, S = np.linalg.lstsq(A, zc, rcond=None)[0]
    zfit = a + S * dist
    # cell areas (t_areas.f90:40-41 form)
    x0, x1, x2 = (case.x[tri[:, j]] for j in range(3))
    y0, y1, y2 = (case.y[tri[:, j]] for j in range(3))
    area = 0.5 * np.abs((x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0))
    N = case.n_elems
    is_in = np.zeros(N)
    is_in[_string_cells(case, bi.nodes)] = 1.0
    is_out = np.zeros(N)
    is_out[_string_cells(case, bo.nodes)] = 1.0
    ee = case.edge_e
    bnd = ee[:, 1] < 0
    is_wall = np.zeros(N)
    is_wall[ee[bnd, 0]] = 1.0
    is_wall = np.clip(is_wall - is_in - is_out, 0.0, 1.0)
    X = np.zeros((N, F_NODE), dtype=np.float32)
    X[:, 0] = (zc - H) / D_SCALE
    X[:, 1] = (zc - zfit) / Y0_SCALE
    X[:, 6] = dist / 1000.0
    X[:, 7] = (span - dist) / 1000.0
    X[:, 8] = is_in
    X[:, 9] = is_out
    X[:, 10] = is_wall
    X[:, 11] = np.sqrt(area) / D_SCALE
    if features == 1:
        X[:, 2] = case.cd / 0.04 - 1.0
        X[:, 3] = q / D_SCALE
        X[:, 4] = (H - zfit) / D_SCALE
        X[:, 5] = S * 1000.0
        X[:, 12] = is_in * q / D_SCALE
        X[:, 13] = is_out * (H - zc) / D_SCALE
    elif features in (2, 3):
        h_out = H - a  # outlet stage over the fitted bed at the outlet (dist 0)
        if not (h_out > 0 and S > 0):
            raise ValueError("the warm start needs a positive outlet depth and bed slope")
        X[:, 2] = np.log(case.cd / 0.04)
        X[:, 3] = np.log(q / D_SCALE)
        X[:, 4] = np.log(h_out / D_SCALE)
        X[:, 5] = np.log(S * 1000.0)
        X[:, 12] = S * dist / h_out
        X[:, 13] = np.log(q * q / (9.81 * h_out ** 3))
        X[:, 14] = (H - zfit) / D_SCALE
    else:
        raise ValueError("unknown feature set %r" % features)
    inter = ee[:, 1] >= 0
    i, j = ee[inter, 0], ee[inter, 1]
    src = np.concatenate([i, j]).astype(np.int32)
    dst = np.concatenate([j, i]).astype(np.int32)
    dx = xc[dst] - xc[src]
    dy = yc[dst] - yc[src]
    E = src.shape[0]
    EF = np.zeros((E, F_EDGE), dtype=np.float32)
    EF[:, 0] = (dx * e[0] + dy * e[1]) / D_SCALE
    EF[:, 1] = (-dx * e[1] + dy * e[0]) / D_SCALE
    EF[:, 2] = np.hypot(dx, dy) / D_SCALE
    za = zc - zfit
    EF[:, 3] = (zc[dst] - zc[src]) / Y0_SCALE
    EF[:, 4] = (za[dst] - za[src]) / Y0_SCALE
    EF[:, 5] = 1.0
    fr = Frame(src, dst, X, EF, zc, dist, e, H, q, float(S), float(a))
    fr.base = np.zeros(N)
    if features == 3:
        # log of each cell's steady friction factor dt_last / dt_i
        # (friction_multiplier) at the edge's two cells
        per = sum(np.hypot(case.x[tri[:, j]] - case.x[tri[:, (j + 1) % 3]],
                           case.y[tri[:, j]] - case.y[tri[:, (j + 1) % 3]]) for j in range(3))
        ap = area / per
        lr = np.log(ap[-1] / ap)
        EF[:, 6] = lr[src]
        EF[:, 7] = lr[dst]
        # the gradually varied flow's surface over the plane, as the zero
        # output and as a node feature
        n_eff = float(np.mean(case.cd)) * np.sqrt(friction_multiplier(case))
        fr.base = (a + gvf_depth(fr, n_eff)) - H
        X[:, 15] = fr.base / Y0_SCALE
    return fr


def targets(fr: Frame, h, qu, qv) -> np.ndarray:
    """Network outputs that decode to the cell state (h, hu, hv)."""
    N = h.shape[0]
    T = np.zeros((N, F_OUT), dtype=np.float64)
    zeta = h + fr.zc
    T[:, 0] = (zeta - fr.H - fr.S * fr.dist - fr.base) / Y0_SCALE
    T[:, 1] = ((qu * fr.e[0] + qv * fr.e[1]) / fr.q - 1.0) / Q_SCALE
    T[:, 2] = ((qv * fr.e[0] - qu * fr.e[1]) / fr.q) / Q_SCALE
    return T.astype(np.float32)


def decode(fr: Frame, y: np.ndarray, h_dry: float = 1e-3, discharge: str = "learned"):
    """(h, u, v) of network outputs y (N, >=3).  discharge="uniform" keeps
    the learned water surface but takes the plane's uniform discharge along
    the axis in place of the learned one (outputs 1 and 2 read as zero)."""
    y = np.asarray(y, dtype=np.float64)
    zeta = fr.H + fr.S * fr.dist + fr.base + Y0_SCALE * y[:, 0]
    h = np.maximum(zeta - fr.zc, 0.0)
    if discharge not in ("learned", "uniform"):
        raise ValueError("discharge must be 'learned' or 'uniform'")
    lq = discharge == "learned"
    qa = fr.q * (1.0 + Q_SCALE * y[:, 1]) if lq else np.full_like(h, fr.q)
    qc = fr.q * Q_SCALE * y[:, 2] if lq else np.zeros_like(h)
    hu = qa * fr.e[0] - qc * fr.e[1]
    hv = qa * fr.e[1] + qc * fr.e[0]
    wet = h > h_dry
    u = np.where(wet, hu / np.where(wet, h, 1.0), 0.0)
    v = np.where(wet, hv / np.where(wet, h, 1.0), 0.0)
    return h, u, v


def cold_state(case):
    """The uninformed start: still water at one stage, the larger of the
    outlet stage and the highest cell bed plus the outlet depth, so that
    every cell (the inflow string's too) starts under at least the outlet's
    depth.  Flat water at the outlet stage alone leaves the upper part of a
    reach whose bed drops more than its outlet depth shallow or dry, and 30
    of the 256 ensemble reaches then blow up within ten steps (on the oracle
    as on the device); from this start every one converges."""
    zc = case.zvtx[case.tri].sum(axis=1) / 3.0
    bo = case.outflows[0]
    H = float(bo.v[0])
    d_out = H - float(case.zvtx[bo.nodes].mean())
    stage = max(H, float(zc.max()) + d_out)
    h = np.maximum(stage - zc, 0.0)
    return h, np.zeros(case.n_elems), np.zeros(case.n_elems)


def _pool_map(fn, items):
    """fn over items on a thread pool (the array work releases the GIL),
    results in order."""
    from concurrent.futures import ThreadPoolExecutor
    items = list(items)
    if len(items) < 2:
        return [fn(x) for x in items]
    with ThreadPoolExecutor(min(16, os.cpu_count() or 1, len(items))) as ex:
        return list(ex.map(fn, items))


def frames(cases_: Sequence, features: int = FEATURES) -> List[Frame]:
    """frame() of every case, built in parallel."""
    return _pool_map(lambda c: frame(c, features), cases_)


def starts(cases_: Sequence, kind: str) -> List:
    """The non-learned starts of every case, built in parallel: "plane"
    (baseline_state), "gvf" (gvf_state) or "cold" (cold_state)."""
    fn = {"plane": baseline_state, "gvf": gvf_state, "cold": cold_state}[kind]
    return _pool_map(fn, cases_)


def baseline_state(case, h_dry: float = 1e-3):
    """The decode of zero outputs of feature sets 1 and 2: the plane
    through the outlet stage parallel to the fitted bed, with uniform
    discharge along the axis."""
    fr = frame(case, 2)
    return decode(fr, np.zeros((case.n_elems, 3)), h_dry)


G = 9.81


def friction_multiplier(case) -> float:
    """The steady mode's friction scale over the reach.  In steady runs the
    implicit friction of every cell is applied with the local time step of
    the reach's last cell (solverFVT.f90:260-265,336) while its fluxes take
    the cell's own, so at the fixed point cell i's Manning term carries the
    factor r_i = dt_last / dt_i.  With the flow's speed about the same in
    every cell, r_i is the ratio of area-over-perimeter of the two cells.
    The reach responds to its harmonic mean, 1 / mean(1 / r_i): the last
    cell's area-over-perimeter over the reach's mean (on converged training
    reaches the multiplier that best fits the 2-D surface was within 1.2 %
    of it, where the plain mean of r_i overshoots by up to 3.5 %).  The
    reference's last cell is the last one in input order."""
    tri = case.tri
    x, y = case.x[tri], case.y[tri]
    area = 0.5 * np.abs((x[:, 1] - x[:, 0]) * (y[:, 2] - y[:, 0])
                        - (x[:, 2] - x[:, 0]) * (y[:, 1] - y[:, 0]))
    per = sum(np.hypot(x[:, j] - x[:, (j + 1) % 3], y[:, j] - y[:, (j + 1) % 3])
              for j in range(3))
    ap = area / per
    return float(ap[-1] / np.mean(ap))


def gvf_depth(fr: Frame, manning: float, n_steps: int = 4000) -> np.ndarray:
    """Depth over the fitted bed along the reach axis of the steady 1-D
    gradually varied flow with the frame's unit discharge q:
    (1 - Fr^2) dh/dx = S - Sf, Fr^2 = q^2 / (g h^3), Sf = n^2 q^2 / h^(10/3)
    (the depth-based Manning slope of the 2-D model's friction term), x
    downstream.  Integrated upstream from the outlet depth H - a by the
    classical fourth-order Runge-Kutta method on n_steps equal steps of the
    axis; returns h at each cell's distance from the outlet.  Where the
    profile would reach 1.1 times the critical depth it is held there (the
    frame's subcritical assumption)."""
    from . import mpnn
    q, S = fr.q, fr.S
    span = float(fr.dist.max()) if fr.dist.size else 0.0
    hc = (q * q / G) ** (1.0 / 3.0)
    h_floor = 1.1 * hc
    s = np.linspace(0.0, max(span, 1e-9), n_steps + 1)
    ds = float(s[1] - s[0])
    h0 = max(float(fr.H - fr.a), h_floor)
    # the integration itself is host C++ in the network's library
    # (storm_mpnn_gvf_profile): the same operations as in Python, without
    # holding the interpreter while the starts of many reaches are built
    h = mpnn.gvf_profile(float(q), float(S), float(manning), h0, h_floor, ds, n_steps)
    return np.interp(fr.dist, s, h)


def gvf_state(case, h_dry: float = 1e-3):
    """A non-learned start from 1-D hydraulics: the water surface of the
    gradually varied flow (gvf_depth with the reach's mean Manning n scaled
    by the steady mode's friction factor, friction_multiplier) over the
    fitted bed, with uniform discharge along the axis.  Where the plane
    guess assumes normal flow throughout, this follows the backwater curve
    the outlet stage sets."""
    fr = frame(case, 3)  # zero outputs of set 3 decode to this start
    return decode(fr, np.zeros((case.n_elems, 3)), h_dry)


# --------------------------------------------------------------------------
# Training data: steady states from the batched FP64 solver
# --------------------------------------------------------------------------

STEADY_OPTS = dict(flux="hll", limiter=0, rkorder=2, cfl=0.9, steady=True, stop_cr=1,
                   rstop=1e-6, dt_op=0, graph=1)


def steady_states(reaches: Sequence, opts: Optional[Dict] = None, max_steps: int = 60000,
                  chunk: int = 1000, init: Optional[List] = None, log=None):
    """Run reaches together (one batched context) until each converges.
    Returns per reach: (state dict or None if it did not converge, steps),
    and the wall time.  init: optional per-reach (h, u, v) starting states."""
    from . import fvt
    o = dict(STEADY_OPTS)
    if opts:
        o.update(opts)
    o["istop"] = max_steps
    t0 = time.perf_counter()
    b = fvt.BatchSolver(list(reaches), o)
    if init is not None:
        b.set_state(np.concatenate([s[0] for s in init]), np.concatenate([s[1] for s in init]),
                    np.concatenate([s[2] for s in init]))
    while True:
        b.run(chunk)
        it, tt, rm, dn = b.status()
        if log:
            log("steady: %d/%d converged, step %d" % (dn.sum(), len(dn), it.max()))
        if dn.all() or it.max() >= max_steps:
            break
    wall = time.perf_counter() - t0
    parts = b.split(b.state())
    b.close()
    out = []
    for r, p in enumerate(parts):
        ok = bool(dn[r]) and np.isfinite(rm[r])
        out.append((p if ok else None, int(it[r])))
    return out, wall


# --------------------------------------------------------------------------
# The network
# --------------------------------------------------------------------------

def window(fr: Frame, y: np.ndarray, start: int, size: int, hops: int):
    """A training sample cut from a reach: the `size` cells ranked start..
    start+size-1 along the axis (downstream order), the edges between them,
    their features and targets, and loss weights that are zero within `hops`
    edges of the cut (where the network's receptive field of `hops` layers
    is truncated) and one elsewhere."""
    if fr.rank is None:
        order = np.argsort(fr.dist, kind="stable")[::-1]  # upstream first
        fr.order = order.astype(np.int64)
        fr.rank = np.empty_like(fr.order)
        fr.rank[fr.order] = np.arange(fr.order.shape[0])
        rs = fr.rank[fr.src]
        fr.eperm = np.argsort(rs, kind="stable")
        fr.ers = rs[fr.eperm]
    a, b = start, start + size
    # the edges whose source lies in the window, then those whose
    # destination does too, in edge order
    j0, j1 = np.searchsorted(fr.ers, [a, b])
    cand = fr.eperm[j0:j1]
    rd = fr.rank[fr.dst[cand]]
    ind = (rd >= a) & (rd < b)
    keep = np.sort(cand[ind])
    src = (fr.rank[fr.src[keep]] - a).astype(np.int32)
    dst = (fr.rank[fr.dst[keep]] - a).astype(np.int32)
    near = np.zeros(size, dtype=bool)
    near[fr.ers[j0:j1][~ind] - a] = True  # cells with a neighbour outside the window
    for _ in range(hops - 1):
        near[dst[near[src]]] = True
    cells = fr.order[a:b]
    return (src, dst, fr.node[cells], fr.edge[keep], y[cells],
            np.where(near, 0.0, 1.0).astype(np.float32))


def _union(frames: Sequence[Frame], ys: Optional[Sequence[np.ndarray]] = None):
    """Disjoint union of reach graphs (one batch)."""
    off = 0
    src, dst, X, EF, T = [], [], [], [], []
    for k, fr in enumerate(frames):
        src.append(fr.src + off)
        dst.append(fr.dst + off)
        X.append(fr.node)
        EF.append(fr.edge)
        if ys is not None:
            T.append(ys[k])
        off += fr.node.shape[0]
    cat = np.concatenate
    return (cat(src).astype(np.int32), cat(dst).astype(np.int32), cat(X), cat(EF),
            cat(T) if ys is not None else None)


class WarmStart:
    """The trained network and what it needs to predict a reach's state."""

    def __init__(self, layers: int = 6, lr: float = 1e-3, weight_decay: float = 1e-4,
                 seed: int = 0, params: Optional[np.ndarray] = None, features: int = FEATURES):
        from . import mpnn
        self.layers = layers
        self.features = features
        self.net = mpnn.Net(layers=layers, f_node=F_NODE, f_edge=F_EDGE, f_out=F_OUT, lr=lr,
                            weight_decay=weight_decay, seed=seed)
        if params is not None:
            self.net.set_params(params)
        self.history: List[Dict] = []

    # -- training -------------------------------------------------------
    def train(self, frames: Sequence[Frame], ys: Sequence[np.ndarray], epochs: int,
              batch: int = 4, seed: int = 0, log=None, time_limit: Optional[float] = None,
              lr: float = 1e-3, lr_final: float = 1e-5):
        """AdamW over mini-batches of whole reaches (disjoint unions), the
        learning rate decaying from lr to lr_final on a cosine over the
        epochs."""
        rng = np.random.default_rng(seed)
        n = len(frames)
        t0 = time.perf_counter()
        for ep in range(epochs):
            c = 0.5 * (1.0 + np.cos(np.pi * ep / max(epochs - 1, 1)))
            self.net.set_lr(lr_final + (lr - lr_final) * c)
            order = rng.permutation(n)
            tot, cnt = 0.0, 0
            for s in range(0, n, batch):
                idx = order[s:s + batch]
                src, dst, X, EF, T = _union([frames[i] for i in idx], [ys[i] for i in idx])
                self.net.set_graph(src, dst, X, EF)
                self.net.forward(copy=False)
                loss = self.net.backward(T)
                self.net.step()
                tot += loss * len(idx)
                cnt += len(idx)
            rec = dict(epoch=ep, loss=tot / cnt, seconds=time.perf_counter() - t0)
            self.history.append(rec)
            if log:
                log("epoch %d loss %.5f (%.1f s)" % (ep, rec["loss"], rec["seconds"]))
            if time_limit and time.perf_counter() - t0 > time_limit:
                break
        return self.history

    def train_windows(self, frames: Sequence[Frame], ys: Sequence[np.ndarray], steps: int,
                      reaches_per_step: int = 32, size: int = 4096, seed: int = 0, log=None,
                      lr: float = 1e-3, lr_final: float = 1e-5, log_every: int = 100,
                      time_limit: Optional[float] = None):
        """AdamW over batches that each join windows of `size` cells cut from
        `reaches_per_step` different reaches (sampled without replacement),
        so every update sees many reaches' parameters at once.  The next
        batch is assembled on a host thread while the device runs the
        current one.  The learning rate decays from lr to lr_final on a
        cosine over the steps.  time_limit (s): after the first 50 updates
        the step count is lowered, if needed, to what fits in that time at
        the measured rate, and the cosine runs over the lowered count."""
        from concurrent.futures import ThreadPoolExecutor
        rng = np.random.default_rng(seed)
        n = len(frames)
        r = min(reaches_per_step, n)

        def batch():
            parts, off = [], 0
            for i in rng.choice(n, r, replace=False):
                fr, y = frames[i], ys[i]
                k = min(size, fr.node.shape[0])
                a = int(rng.integers(0, fr.node.shape[0] - k + 1))
                src, dst, X, EF, T, w = window(fr, y, a, k, self.layers)
                parts.append((src + off, dst + off, X, EF, T, w))
                off += k
            return tuple(np.concatenate([p[j] for p in parts]) for j in range(6))

        t0 = time.perf_counter()
        tot, cnt = 0.0, 0
        with ThreadPoolExecutor(1) as ex:
            nxt = ex.submit(batch)
            step = 0
            while step < steps:
                if time_limit and step == 50:
                    fit = int(50 * time_limit / max(time.perf_counter() - t0, 1e-9))
                    if fit < steps:
                        steps = max(fit, 51)
                        if log:
                            log("time limit: %d updates" % steps)
                src, dst, X, EF, T, w = nxt.result()
                if step + 1 < steps:
                    nxt = ex.submit(batch)
                c = 0.5 * (1.0 + np.cos(np.pi * step / max(steps - 1, 1)))
                self.net.set_lr(lr_final + (lr - lr_final) * c)
                self.net.set_graph(src, dst, X, EF)
                self.net.forward(copy=False)
                tot += self.net.backward(T, w)
                cnt += 1
                self.net.step()
                if (step + 1) % log_every == 0 or step + 1 == steps:
                    rec = dict(step=step + 1, loss=tot / cnt, seconds=time.perf_counter() - t0)
                    self.history.append(rec)
                    if log:
                        log("step %d loss %.5f (%.1f s)" % (rec["step"], rec["loss"], rec["seconds"]))
                    tot, cnt = 0.0, 0
                step += 1
        return self.history

    def loss(self, frames: Sequence[Frame], ys: Sequence[np.ndarray]) -> float:
        """Mean over reaches of each reach's mean loss, without updating."""
        tot = 0.0
        for fr, y in zip(frames, ys):
            self.net.set_graph(fr.src, fr.dst, fr.node, fr.edge)
            out = self.net.forward()
            tot += float(((out - y) ** 2).mean())
        return tot / len(frames)

    # -- inference ------------------------------------------------------
    def predict_outputs(self, frames: Sequence[Frame], batch: int = 16) -> List[np.ndarray]:
        """Network outputs of every frame, `batch` reaches per forward pass
        (the reaches' graphs are disjoint, so a reach's outputs do not
        depend on the others in its batch)."""
        out = []
        for s in range(0, len(frames), batch):
            fs = frames[s:s + batch]
            src, dst, X, EF, _ = _union(fs)
            self.net.set_graph(src, dst, X, EF)
            y = self.net.forward()
            o = 0
            for fr in fs:
                n = fr.node.shape[0]
                out.append(y[o:o + n])
                o += n
        return out

    def predict(self, cases_: Sequence, batch: int = 16, h_dry: float = 1e-3,
                discharge: str = "uniform"):
        """Per case: (h, u, v) for storm_fvt_set_state.  By default the
        network's water surface with uniform discharge along the axis: a
        steady flow carries the same discharge through every section, and
        the network's cell-scale discharge errors would start transients
        that the surface alone does not (64 held-out 200k reaches reached
        the steady test in 18.4 s this way, 21.5 s with the network's
        discharge outputs).  discharge="learned" decodes those too."""
        frs = frames(cases_, self.features)
        ys = self.predict_outputs(frs, batch)
        return _pool_map(lambda a: decode(a[0], a[1], h_dry, discharge), list(zip(frs, ys)))

    # -- checkpoint ---------------------------------------------------------
    def save(self, path: str, meta: Optional[Dict] = None):
        m = dict(layers=self.layers, f_node=F_NODE, f_edge=F_EDGE, f_out=F_OUT,
                 features=self.features, y0_scale=Y0_SCALE, q_scale=Q_SCALE,
                 history=self.history)
        if meta:
            m.update(meta)
        np.savez_compressed(path, params=self.net.params(), meta=json.dumps(m))

    @classmethod
    def load(cls, path: str) -> "WarmStart":
        d = np.load(path, allow_pickle=False)
        meta = json.loads(str(d["meta"]))
        if meta.get("y0_scale", Y0_SCALE) != Y0_SCALE or meta.get("q_scale", Q_SCALE) != Q_SCALE:
            raise ValueError("checkpoint %s was trained with other output scales" % path)
        ws = cls(layers=int(meta["layers"]), params=d["params"],
                 features=int(meta.get("features", 1)))
        ws.meta = meta
        ws.history = list(meta.get("history", []))
        return ws

    def close(self):
        self.net.close()


DEFAULT_CHECKPOINT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data",
                                  "warmstart_mpnn.npz")


def training_reach(k: int, **kw):
    """Training reach k: the ensemble generator at 20k-100k triangles,
    seeds 5005+k."""
    return _cases.ensemble_reach(k, n_min=20000, n_max=100000, base_seed=5005, steady=True, **kw)


def heldout_reach(k: int, n: int = 200000, **kw):
    """Held-out reach k: the ensemble generator at n triangles, seeds
    9009+k (disjoint from the training seeds)."""
    return _cases.ensemble_reach(k, n_min=n, n_max=n, base_seed=9009, steady=True, **kw)
