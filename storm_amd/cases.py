"""Synthetic SToRM cases: meshes, beds, initial states and boundary strings.

Every generator returns a :class:`Case` holding the same information the
reference reads from a ".run" case (Source_xsy-64bit/read_mfa.f90:67-173):

* nodes ``x, y`` (P), triangles ``tri`` (N x 3, 0-based, counter-clockwise),
* the edge table ``edge_p`` / ``edge_e`` (E x 2, 0-based, -1 = no cell) in the
  order the reference's own edge search produces it
  (Source_xsy-64bit/find_edges.f90:47-54,165-187: first appearance while
  sweeping cells and their local sides, p(1)->p(2) oriented as in the first
  cell, e(1) = first cell),
* bed elevation at the vertices ``zvtx`` and cell-centred Manning n ``cd``,
  initial depth and velocities at the cells,
* inflow strings (vtype, node chain, Q(t) table) and outflow strings (htype,
  node chain, H(t) table) in the reference's ioflow format
  (Source_xsy-64bit/ioflow.f90:179-313).

:func:`write_run` writes a case as the reference's ".run" input (Tecplot
FETRIANGLE block file with edge sections, boundary file, options file), which
is how the oracle drives the reference.  All randomness comes from numpy's
PCG64 seeded generator, drawn in a fixed order documented per generator.
"""
from __future__ import annotations

import dataclasses
import math
import os
from typing import Dict, List, Optional, Tuple

import numpy as np

GRAVITY = 9.81


@dataclasses.dataclass
class Boundary:
    """One inflow or outflow node string and its time table."""
    kind: int                 # vtype (inflow) or htype (outflow)
    nodes: np.ndarray         # 0-based node ids, consecutive ones share an edge
    t: np.ndarray             # time table (may be empty for htype 0/2)
    v: np.ndarray             # Q (inflow) or stage H (outflow)


@dataclasses.dataclass
class Culvert:
    """One culvert as ioculvert.f90:84-131 reads it: inlet point and invert
    elevation, outlet point, and the performance curve (head above the
    invert -> discharge)."""
    inlet: Tuple[float, float, float]   # x, y, invert z
    outlet: Tuple[float, float]         # x, y
    head: np.ndarray                    # ascending heads
    q: np.ndarray                       # discharge at each head


@dataclasses.dataclass
class Case:
    name: str
    x: np.ndarray
    y: np.ndarray
    tri: np.ndarray           # (N, 3) int32, CCW
    edge_p: np.ndarray        # (E, 2) int32
    edge_e: np.ndarray        # (E, 2) int32, -1 = boundary
    zvtx: np.ndarray          # (P,)
    cd: np.ndarray            # (N,) Manning n
    h: np.ndarray             # (N,)
    u: np.ndarray             # (N,)
    v: np.ndarray             # (N,)
    inflows: List[Boundary]
    outflows: List[Boundary]
    opts: Dict[str, object]
    meta: Dict[str, object] = dataclasses.field(default_factory=dict)
    # Inputs given at the other location than the fields above, as a grid
    # file may hold them (dataprepFVT.f90:73-140 takes either): "z" the bed
    # at the cells (N); "h", "u", "v", "cd" values at the vertices (P).
    alt: Dict[str, np.ndarray] = dataclasses.field(default_factory=dict)
    # Wind forcing data as windforcing.f90 reads it: coefficient, x and y
    # wind components, each per cell (N values) or per vertex (P values).
    wind: Optional[Tuple[np.ndarray, np.ndarray, np.ndarray]] = None
    # Culverts (ioculvert.f90); a single reach only.
    culverts: List[Culvert] = dataclasses.field(default_factory=list)

    @property
    def n_elems(self) -> int:
        return int(self.tri.shape[0])

    @property
    def n_pts(self) -> int:
        return int(self.x.shape[0])

    @property
    def n_edges(self) -> int:
        return int(self.edge_p.shape[0])


def inputs(case: Case) -> Dict[str, Tuple[np.ndarray, bool]]:
    """The initial data as the reference reads it: name -> (values, given at
    the vertices) for z, h, u, v and cd."""
    out = {"z": (case.alt["z"], False) if "z" in case.alt else (case.zvtx, True)}
    for k in ("h", "u", "v", "cd"):
        out[k] = (case.alt[k], True) if k in case.alt else (getattr(case, k), False)
    return out


# --------------------------------------------------------------------------
# Topology
# --------------------------------------------------------------------------

def build_edges(tri: np.ndarray, n_pts: int) -> Tuple[np.ndarray, np.ndarray]:
    """Edge table in the reference's first-appearance order.

    Follows Source_xsy-64bit/find_edges.f90:47-54 (edges added while sweeping
    cell i=1..N, local side j=1..3 from vertex(j) to vertex(j+1)) and :174-187
    (e(1) is the first cell met, e(2) the second).
    """
    n = tri.shape[0]
    a = tri.reshape(-1).astype(np.int64)
    b = tri[:, [1, 2, 0]].reshape(-1).astype(np.int64)
    lo = np.minimum(a, b)
    hi = np.maximum(a, b)
    key = lo * np.int64(n_pts) + hi
    order = np.argsort(key, kind="stable")
    ks = key[order]
    start = np.ones(ks.shape[0], dtype=bool)
    start[1:] = ks[1:] != ks[:-1]
    grp = np.cumsum(start) - 1
    first_pos = order[start]                      # flat (cell*3+side) of first hit
    n_e = first_pos.shape[0]
    counts = np.bincount(grp, minlength=n_e)
    if counts.max() > 2:
        raise ValueError("mesh connectivity error: an edge is shared by >2 cells")
    second = np.full(n_e, -1, dtype=np.int64)
    two = np.where(counts == 2)[0]
    # positions of the second occurrence (sorted stable => order[start+1])
    sidx = np.where(start)[0]
    second[two] = order[sidx[two] + 1]
    # edge numbering = order of first appearance
    perm = np.argsort(first_pos, kind="stable")
    first_pos = first_pos[perm]
    second = second[perm]
    edge_p = np.stack([a[first_pos], b[first_pos]], axis=1).astype(np.int32)
    e1 = (first_pos // 3).astype(np.int32)
    e2 = np.where(second >= 0, second // 3, -1).astype(np.int32)
    edge_e = np.stack([e1, e2], axis=1)
    return edge_p, edge_e


def _grid_triangles(nx: int, ny: int, diag: np.ndarray) -> np.ndarray:
    """Split an (nx, ny) quad grid into CCW triangles, node id = i*(ny+1)+j.

    Quad (i, j) has corners a=(i,j), b=(i+1,j), c=(i+1,j+1), d=(i,j+1).
    ``diag`` = 0 splits along a-c, 1 along b-d.  Cells of quad (i, j) are
    2*(i*ny+j) and 2*(i*ny+j)+1.
    """
    ii, jj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    a = ii * (ny + 1) + jj
    b = (ii + 1) * (ny + 1) + jj
    c = (ii + 1) * (ny + 1) + jj + 1
    d = ii * (ny + 1) + jj + 1
    dg = diag.astype(bool)
    t0 = np.where(dg[..., None], np.stack([a, b, d], -1), np.stack([a, b, c], -1))
    t1 = np.where(dg[..., None], np.stack([b, c, d], -1), np.stack([a, c, d], -1))
    tri = np.stack([t0, t1], axis=2).reshape(-1, 3)
    return tri.astype(np.int32)


def _jitter(rng, nx, ny, dx, dy, amp=0.2):
    """Jitter of grid nodes: interior in both directions, boundary nodes only
    along their own boundary line, corners fixed.  Draws x then y jitter."""
    jx = rng.uniform(-amp, amp, size=(nx + 1, ny + 1)) * dx
    jy = rng.uniform(-amp, amp, size=(nx + 1, ny + 1)) * dy
    jx[0, :] = 0.0
    jx[nx, :] = 0.0
    jy[:, 0] = 0.0
    jy[:, ny] = 0.0
    return jx, jy


def _cell_centroids(x, y, tri):
    return (x[tri].sum(axis=1) / 3.0, y[tri].sum(axis=1) / 3.0)


def default_opts(**kw) -> Dict[str, object]:
    """Options with the reference defaults (initialize_options.f90:23-141)
    except where a case overrides them."""
    o: Dict[str, object] = dict(
        rkorder=2, cfl=1.0, dt_op=0, steady=False, stop_cr=0, rstop=1.0e-4,
        istop=0, h_dry=0.001, c_threshold=1.0, btype=0, hvalve=True,
        limiter=2, flux="ac", gradient="lsq", friction=0, vclip=False,
        vceiling=35.0, fluxes="std", rsavg=False, rsavg_n=2, rsavg_eps=0.5)
    o.update(kw)
    return o


# --------------------------------------------------------------------------
# Config 1: straight flume, and the ensemble reaches built like it
# --------------------------------------------------------------------------

def flume(nx: int = 200, ny: int = 50, length: float = 1000.0,
          width: float = 100.0, seed: int = 1001, slope: float = 0.001,
          rough_sd: float = 0.02, manning: float = 0.03, q_in: float = 200.0,
          h_out: float = 2.0, h0: Optional[float] = 2.0, name: str = "flume",
          in_kind: int = 1, out_kind: int = 1, **opt_kw) -> Case:
    """Straight flume [0,L] x [0,W].

    Random draws, in order: node x-jitter, node y-jitter (each U(-0.2,0.2)
    times the spacing), quad diagonal (Bernoulli 1/2), bed roughness
    N(0, rough_sd^2) per node.  Bed: slope*(L-x) + roughness.  Inflow string
    (vtype 1, Q constant) on x=0, outflow string (htype 1, stage constant) on
    x=L, walls on y=0 and y=W.  Initial state: h = h0 (or stage h_out when
    h0 is None), u = v = 0.  in_kind 4 (stage inflow, q_in is the stage) or 5
    (discharge into wet cells only) and out_kind 0 (free outflow) or 2
    (critical depth) select the other boundary routines.
    """
    rng = np.random.default_rng(seed)
    dx, dy = length / nx, width / ny
    jx, jy = _jitter(rng, nx, ny, dx, dy)
    diag = rng.integers(0, 2, size=(nx, ny))
    ii, jj = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), indexing="ij")
    x = (ii * dx + jx).reshape(-1)
    y = (jj * dy + jy).reshape(-1)
    tri = _grid_triangles(nx, ny, diag)
    edge_p, edge_e = build_edges(tri, x.shape[0])
    zvtx = slope * (length - x) + rng.normal(0.0, rough_sd, size=x.shape[0])
    n = tri.shape[0]
    cd = np.full(n, manning)
    if h0 is None:
        zc = zvtx[tri].sum(axis=1) / 3.0
        h = np.maximum(h_out - zc, 0.0)
    else:
        h = np.full(n, float(h0))
    u = np.zeros(n)
    v = np.zeros(n)
    inflow_nodes = np.arange(0, ny + 1, dtype=np.int32)            # i = 0
    outflow_nodes = (nx * (ny + 1) + np.arange(0, ny + 1)).astype(np.int32)
    inflows = [Boundary(in_kind, inflow_nodes, np.array([0.0]), np.array([q_in]))]
    if out_kind == 1:
        outflows = [Boundary(1, outflow_nodes, np.array([0.0]), np.array([h_out]))]
    else:
        outflows = [Boundary(out_kind, outflow_nodes, np.zeros(0), np.zeros(0))]
    opts = default_opts(rkorder=2, cfl=0.9, dt_op=1, steady=False, istop=2000)
    opts.update(opt_kw)
    meta = dict(generator="flume", nx=nx, ny=ny, length=length, width=width,
                seed=seed, slope=slope, rough_sd=rough_sd, manning=manning,
                q_in=q_in, h_out=h_out, in_kind=in_kind, out_kind=out_kind)
    return Case(name, x, y, tri, edge_p, edge_e, zvtx, cd, h, u, v, inflows,
                outflows, opts, meta)


def normal_depth_rect(q: float, width: float, n: float, slope: float) -> float:
    """Manning normal depth of a rectangular channel, Q = A R^(2/3) S^(1/2) / n,
    solved by bisection."""
    lo, hi = 1e-6, 100.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        a = width * mid
        r = a / (width + 2 * mid)
        qq = a * r ** (2.0 / 3.0) * math.sqrt(slope) / n
        if qq > q:
            hi = mid
        else:
            lo = mid
    return 0.5 * (lo + hi)


def ensemble_reach(k: int, n_min: int = 5000, n_max: int = 200000,
                   base_seed: int = 4004, **opt_kw) -> Case:
    """Reach k of the batched ensemble: a flume like :func:`flume` whose
    triangle count is log-uniform in [n_min, n_max] with aspect 10:1, and
    per-reach slope U[2e-4,2e-3], Manning n U[0.02,0.06], Q U[50,1500].
    Outflow stage = bed at the outlet + Manning normal depth of a 100 m
    rectangle.  Parameter draws (own generator, seed base_seed+k, in order):
    log N, slope, n, Q; the mesh uses seed base_seed+k+1_000_000."""
    rng = np.random.default_rng(base_seed + k)
    logn = rng.uniform(math.log(n_min), math.log(n_max))
    slope = rng.uniform(2e-4, 2e-3)
    manning = rng.uniform(0.02, 0.06)
    q = rng.uniform(50.0, 1500.0)
    ntri = math.exp(logn)
    ny = max(4, int(round(math.sqrt(ntri / 20.0))))
    nx = max(4, int(round(ntri / (2 * ny))))
    hn = normal_depth_rect(q, 100.0, manning, slope)
    c = flume(nx=nx, ny=ny, length=1000.0, width=100.0,
              seed=base_seed + k + 1_000_000, slope=slope, manning=manning,
              q_in=q, h_out=hn, h0=None, name=f"reach{k}", **opt_kw)
    c.meta.update(dict(generator="ensemble_reach", k=k, normal_depth=hn))
    return c


# --------------------------------------------------------------------------
# Config 2: meandering channel with a dry floodplain
# --------------------------------------------------------------------------

def meander(ns: int = 4000, nn: int = 250, length: float = 20000.0,
            width: float = 1000.0, amp: float = 300.0, wavelength: float = 2000.0,
            chan_w: float = 80.0, chan_d: float = 4.0, valley_slope: float = 5e-4,
            n_bumps: int = 200, seed: int = 2002, n_chan: float = 0.035,
            n_fp: float = 0.06, q0: float = 500.0, q1: float = 3000.0,
            t_ramp: float = 3600.0, name: str = "meander", wet: Optional[float] = None,
            **opt_kw) -> Case:
    """Sinusoidal meandering channel.

    Grid in valley coordinates (s, n), s in [0, L], n in [-W/2, W/2], mapped
    by a vertical shear onto the centreline y = amp*sin(2 pi s / wavelength):
    (x, y) = (s, yc(s) + n).  Parabolic main channel |n| < chan_w/2 of depth
    chan_d below the valley floor z_v(s) = valley_slope*(L - s); the
    floodplain carries n_bumps Gaussian bumps (height U[0.5,2] m, sigma
    U[20,80] m, centres uniform over the floodplain).  Initial water at bank
    full in the channel (stage z_v), floodplain dry; with ``wet`` the stage is
    z_v + wet over the whole valley instead (wet = 4 puts every cell under
    water).
    Inflow string on s=0
    across the channel only (vtype 1, Q ramp q0 -> q1 over t_ramp), critical
    outflow (htype 2) across the whole section at s=L, walls elsewhere.
    Random draws, in order: s-jitter, n-jitter, diagonals, bump centres s,
    bump centres n (as |n| in the floodplain with a random side), heights,
    sigmas.
    """
    rng = np.random.default_rng(seed)
    ds, dn = length / ns, width / nn
    js, jn = _jitter(rng, ns, nn, ds, dn)
    diag = rng.integers(0, 2, size=(ns, nn))
    ii, jj = np.meshgrid(np.arange(ns + 1), np.arange(nn + 1), indexing="ij")
    s = (ii * ds + js).reshape(-1)
    nrm = (jj * dn + jn).reshape(-1) - 0.5 * width
    half_c = 0.5 * chan_w
    # keep channel-bank node rows unjittered in n so that the inflow string
    # spans exactly the channel
    x = s
    y = amp * np.sin(2.0 * np.pi * s / wavelength) + nrm
    tri = _grid_triangles(ns, nn, diag)
    edge_p, edge_e = build_edges(tri, x.shape[0])
    zv = valley_slope * (length - s)
    inch = np.abs(nrm) < half_c
    zvtx = zv.copy()
    zvtx[inch] -= chan_d * (1.0 - (nrm[inch] / half_c) ** 2)
    bs = rng.uniform(0.0, length, size=n_bumps)
    bn_abs = rng.uniform(half_c + 20.0, 0.5 * width, size=n_bumps)
    side = np.where(rng.integers(0, 2, size=n_bumps) == 0, -1.0, 1.0)
    bn = side * bn_abs
    bh = rng.uniform(0.5, 2.0, size=n_bumps)
    bsig = rng.uniform(20.0, 80.0, size=n_bumps)
    fp = ~inch
    sf, nf = s[fp], nrm[fp]
    add = np.zeros(sf.shape[0])
    for k in range(n_bumps):
        r2 = (sf - bs[k]) ** 2 + (nf - bn[k]) ** 2
        near = r2 < (6.0 * bsig[k]) ** 2
        add[near] += bh[k] * np.exp(-r2[near] / (2.0 * bsig[k] ** 2))
    zvtx[fp] += add
    n_cells = tri.shape[0]
    sc = s[tri].sum(axis=1) / 3.0
    nc = nrm[tri].sum(axis=1) / 3.0
    zc = zvtx[tri].sum(axis=1) / 3.0
    stage = valley_slope * (length - sc)
    if wet is None:
        h = np.where(np.abs(nc) < half_c, np.maximum(stage - zc, 0.0), 0.0)
    else:  # the whole valley under water, wet metres above the floor
        h = np.maximum(stage + wet - zc, 0.0)
    cd = np.where(np.abs(nc) < half_c, n_chan, n_fp)
    u = np.zeros(n_cells)
    v = np.zeros(n_cells)
    col0 = np.arange(0, nn + 1)
    nrow = col0 * dn - 0.5 * width
    in_sel = col0[np.abs(nrow) <= half_c + 1e-9]
    inflow_nodes = in_sel.astype(np.int32)
    outflow_nodes = (ns * (nn + 1) + col0).astype(np.int32)
    inflows = [Boundary(1, inflow_nodes, np.array([0.0, t_ramp]),
                        np.array([q0, q1]))]
    outflows = [Boundary(2, outflow_nodes, np.zeros(0), np.zeros(0))]
    opts = default_opts(rkorder=3, cfl=0.9, dt_op=1, steady=False, istop=10000)
    opts.update(opt_kw)
    meta = dict(generator="meander", ns=ns, nn=nn, seed=seed, length=length,
                width=width, amp=amp, wavelength=wavelength,
                valley_slope=valley_slope)
    return Case(name, x, y, tri, edge_p, edge_e, zvtx, cd, h, u, v, inflows,
                outflows, opts, meta)


# --------------------------------------------------------------------------
# Config 3: dam break over a dry, bumpy bed
# --------------------------------------------------------------------------

def dambreak(nq: int = 707, side: float = 2000.0, seed: int = 3003,
             n_bumps: int = 50, x_dam: float = 800.0, h_res: float = 10.0,
             manning: float = 0.025, name: str = "dambreak", **opt_kw) -> Case:
    """Square dam break: reservoir depth h_res for x < x_dam (stage
    h_res above the lowest reservoir bed is not enforced: depth is), dry bed
    downstream over a bed of n_bumps Gaussians (amplitude U[0,1.5] m, sigma
    U[50,200] m).  Closed walls.  Random draws, in order: x-jitter, y-jitter,
    diagonals, bump x, bump y, amplitude, sigma."""
    rng = np.random.default_rng(seed)
    d = side / nq
    jx, jy = _jitter(rng, nq, nq, d, d)
    diag = rng.integers(0, 2, size=(nq, nq))
    ii, jj = np.meshgrid(np.arange(nq + 1), np.arange(nq + 1), indexing="ij")
    x = (ii * d + jx).reshape(-1)
    y = (jj * d + jy).reshape(-1)
    tri = _grid_triangles(nq, nq, diag)
    edge_p, edge_e = build_edges(tri, x.shape[0])
    bx = rng.uniform(0.0, side, size=n_bumps)
    by = rng.uniform(0.0, side, size=n_bumps)
    ba = rng.uniform(0.0, 1.5, size=n_bumps)
    bsig = rng.uniform(50.0, 200.0, size=n_bumps)
    zvtx = np.zeros(x.shape[0])
    for k in range(n_bumps):
        zvtx += ba[k] * np.exp(-((x - bx[k]) ** 2 + (y - by[k]) ** 2)
                               / (2 * bsig[k] ** 2))
    xc, _ = _cell_centroids(x, y, tri)
    n_cells = tri.shape[0]
    h = np.where(xc < x_dam, h_res, 0.0)
    cd = np.full(n_cells, manning)
    opts = default_opts(rkorder=2, cfl=0.5, dt_op=1, steady=False,
                        h_dry=1e-4, flux="rusanov", istop=100000)
    opts.update(opt_kw)
    meta = dict(generator="dambreak", nq=nq, seed=seed, side=side)
    return Case(name, x, y, tri, edge_p, edge_e, zvtx, cd, h,
                np.zeros(n_cells), np.zeros(n_cells), [], [], opts, meta)


# --------------------------------------------------------------------------
# Reference ".run" writer
# --------------------------------------------------------------------------

def _fmt(a):
    return "\n".join("%.17e" % float(t) for t in a)


def write_run(case: Case, dirpath: str, out_prefix: str = "case") -> str:
    """Write ``case`` as a reference ".run" input under ``dirpath``.

    Formats: run file = five file names (read_mfa.f90:67-75); grid file =
    Tecplot FETRIANGLE block data with the cell-centred variables 4-7
    (h, u, v, n) flagged in VARLOCATION (tecplot.f90:42-65,
    tecplot_read.f90:49-99) followed by the edge and edge-neighbour sections
    (tecplot_read.f90:117-145); boundary file (ioflow.f90:179-313); options
    file (read_options.f90).  Returns the path of the ".run" file."""
    os.makedirs(dirpath, exist_ok=True)
    P, N, E = case.n_pts, case.n_elems, case.n_edges
    o = case.opts
    grid = os.path.join(dirpath, "grid.dat")
    with open(grid, "w") as f:
        f.write('TITLE = "%s"\n' % case.name)
        f.write('VARIABLES = "X" "Y" "Z" "H" "U" "V" "N" "V8" "V9" "V10" "V11"\n')
        data = inputs(case)
        cells = ["%02d" % (3 + j) for j, k in enumerate(("z", "h", "u", "v", "cd"))
                 if not data[k][1]]
        loc = (", VARLOCATION=([%s]=CELLCENTERED)" % ",".join(cells)) if cells else ""
        f.write("ZONE N=%d, E=%d, DATAPACKING=BLOCK, ZONETYPE=FETRIANGLE%s\n" % (P, N, loc))
        for arr in (case.x, case.y) + tuple(data[k][0] for k in ("z", "h", "u", "v", "cd")):
            f.write(_fmt(arr))
            f.write("\n")
        zeros = "\n".join(["0.0"] * P) + "\n"
        for _ in range(4):
            f.write(zeros)
        t1 = case.tri + 1
        f.write("\n".join("%d %d %d" % (a, b, c) for a, b, c in t1))
        f.write("\n%d\n" % E)
        ep = case.edge_p + 1
        f.write("\n".join("%d %d %d" % (k + 1, a, b) for k, (a, b) in enumerate(ep)))
        f.write("\n%d\n" % E)
        ee = np.where(case.edge_e >= 0, case.edge_e + 1, -1)
        f.write("\n".join("%d %d %d" % (k + 1, a, b) for k, (a, b) in enumerate(ee)))
        f.write("\n")
    bcs = os.path.join(dirpath, "bcs.dat")
    with open(bcs, "w") as f:
        for b in case.inflows:
            f.write("1 %d %d\n" % (b.kind, len(b.t)))
            for t, q in zip(b.t, b.v):
                f.write("%.10e %.10e\n" % (t, q))
            f.write("%d\n" % len(b.nodes))
            f.write("\n".join(str(int(p) + 1) for p in b.nodes) + "\n")
        for b in case.outflows:
            nts = len(b.t) if b.kind == 1 else 0
            f.write("2 %d %d\n" % (b.kind, nts))
            if b.kind == 1:
                for t, hh in zip(b.t, b.v):
                    f.write("%.10e %.10e\n" % (t, hh))
            f.write("%d\n" % len(b.nodes))
            f.write("\n".join(str(int(p) + 1) for p in b.nodes) + "\n")
    opsf = os.path.join(dirpath, "options.dat")
    with open(opsf, "w") as f:
        f.write("SOLVER FVT\n")
        # HEADER 0 leaves DATE_AND_TIME unset before header.f90 indexes month() with it
        f.write("MASTERFILE 0\nGEOMETRY 0\nHEADER 1\nINITIAL CONDS 0\nOPTIONS 0\n")
        f.write("TIME STEPPING %d\n" % (0 if o["steady"] else 1))
        f.write("DELT TYPE %d\n" % int(o["dt_op"]))
        f.write("CFL %.17e\n" % float(o["cfl"]))
        f.write("RK ORDER %d\n" % int(o["rkorder"]))
        f.write("LIMITER %s\n" % {0: "ZERO", 1: "NONE", 2: "BJ"}[int(o["limiter"])])
        f.write("H THRESHOLD %.17e\n" % float(o["h_dry"]))
        # always written: without it h_wet keeps the default 0.001 whatever
        # H THRESHOLD says (read_options.f90:441)
        f.write("C THRESHOLD %.17e\n" % max(1.0, float(o["c_threshold"])))
        f.write("WALLS %s\n" % ("NOSLIP" if int(o["btype"]) == 0 else "FREE"))
        f.write("HVALVE %d\n" % (1 if o["hvalve"] else 0))
        f.write("MAX ITERATION %d\n" % int(o["istop"]))
        f.write("STOPPING CRIT %d\n" % int(o["stop_cr"]))
        f.write("MAX RESIDUAL %.17e\n" % float(o["rstop"]))
        f.write("ROUGHNESS %d\n" % int(o["friction"]))
        if o["vclip"]:
            f.write("MAX VELOCITY %.17e\n" % float(o["vceiling"]))
        if str(o.get("gradient", "lsq")) == "hdls":
            f.write("ACTIVATE HDLS\n")
        if o.get("rsavg"):
            # read_options.f90:347-351 reads values after the keyword from
            # column 13 (inside the keyword), so only the bare form runs: the
            # defaults epsRSAVG = 0.5, nRSAVG = 2
            if int(o.get("rsavg_n", 2)) != 2 or float(o.get("rsavg_eps", 0.5)) != 0.5:
                raise ValueError("the reference reads only the bare ACTIVATE RSAVG (0.5, 2)")
            f.write("ACTIVATE RSAVG\n")
        # file names are upper case: readline.f90 upper-cases every line
        if case.wind is not None:
            f.write("WIND WIND.DAT\n")
        if case.culverts:
            f.write("CULV CULV.DAT\n")
    if case.wind is not None:
        write_wind(case, os.path.join(dirpath, "WIND.DAT"))
    if case.culverts:
        write_culverts(case, os.path.join(dirpath, "CULV.DAT"))
    run = os.path.join(dirpath, out_prefix + ".run")
    with open(run, "w") as f:
        f.write("grid.dat\nbcs.dat\n%s.out\n%s.sol\noptions.dat\n"
                % (out_prefix, out_prefix))
    return run


def _short(v: float) -> str:
    s = repr(float(v))
    if len(s) > 12:
        raise ValueError("value %r does not fit the reference's 40-column lines" % v)
    return s


def write_wind(case: Case, path: str) -> None:
    """windforcing.f90:43-63: the count, then one "c wx wy" line per cell or
    vertex.  readline.f90 keeps 40 columns, so every value must print in at
    most 12 characters (wind_data() rounds them so)."""
    c, wx, wy = case.wind
    with open(path, "w") as f:
        f.write("%d\n" % len(c))
        for a, b, d in zip(c, wx, wy):
            f.write("%s %s %s\n" % (_short(a), _short(b), _short(d)))


def write_culverts(case: Case, path: str) -> None:
    """ioculvert.f90:84-131: the count; per culvert the inlet "x y z" line,
    the outlet "x y" line, the number of curve points and "head q" lines."""
    with open(path, "w") as f:
        f.write("%d\n" % len(case.culverts))
        for cv in case.culverts:
            f.write("%s %s %s\n" % tuple(_short(v) for v in cv.inlet))
            f.write("%s %s\n" % tuple(_short(v) for v in cv.outlet))
            f.write("%d\n" % len(cv.head))
            for hh, qq in zip(cv.head, cv.q):
                f.write("%s %s\n" % (_short(hh), _short(qq)))


def wind_data(rng: np.random.Generator, n: int, c: float = 1.5e-3, speed: float = 12.0,
              direction: float = 0.6) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Random wind data for n cells or vertices: coefficient c(1 + U(-0.2,
    0.2)), components speed(cos, sin)(direction + N(0, 0.3^2)) times U(0.5,
    1.5); each value rounded to 4 significant digits, so that it prints in
    the reference's 40-column lines and reads back to the same double."""
    def r4(a):
        return np.array([float("%.4g" % v) for v in a])
    cc = r4(c * (1.0 + rng.uniform(-0.2, 0.2, n)))
    ang = direction + rng.normal(0.0, 0.3, n)
    mag = speed * rng.uniform(0.5, 1.5, n)
    return cc, r4(mag * np.cos(ang)), r4(mag * np.sin(ang))


def read_ref_output(path: str) -> Dict[str, np.ndarray]:
    """Read the binary dump written by the reference driver."""
    with open(path, "rb") as f:
        hdr = np.frombuffer(f.read(16), dtype=np.int32)
        n, p, it, rko = (int(t) for t in hdr)
        tt = np.frombuffer(f.read(16), dtype=np.float64)
        out: Dict[str, np.ndarray] = dict(iter=np.int64(it), rkorder=np.int64(rko),
                                          ttime=tt[0], rmax=tt[1])
        for k in ("h", "u", "v", "qh", "qu", "qv", "phi1", "phi2", "phi3"):
            out[k] = np.frombuffer(f.read(8 * n), dtype=np.float64).copy()
        out["dry"] = np.frombuffer(f.read(4 * n), dtype=np.int32).copy()
        out["hvtx"] = np.frombuffer(f.read(8 * p), dtype=np.float64).copy()
        out["zetavtx"] = np.frombuffer(f.read(8 * p), dtype=np.float64).copy()
        # the post-loop fields (uvtx, vvtx, cdvtx, zvtx; delh, e_est) follow
        for k in ("uvtx", "vvtx", "cdvtx", "zvtx"):
            out[k] = np.frombuffer(f.read(8 * p), dtype=np.float64).copy()
        for k in ("delhx", "delhy", "e_est"):
            out[k] = np.frombuffer(f.read(8 * n), dtype=np.float64).copy()
    return out


PREP_KEYS = (("z", "c"), ("zvtx", "p"), ("zeta", "c"), ("cd", "c"), ("cdvtx", "p"), ("h", "c"),
             ("u", "c"), ("v", "c"), ("hvtx", "p"), ("uvtx", "p"), ("vvtx", "p"),
             ("zetavtx", "p"))


def read_prep_output(path: str, n: int, p: int) -> Dict[str, np.ndarray]:
    """Read the module copies the reference driver dumps after dataprepFVT."""
    out: Dict[str, np.ndarray] = {}
    with open(path, "rb") as f:
        for k, where in PREP_KEYS:
            m = n if where == "c" else p
            out[k] = np.frombuffer(f.read(8 * m), dtype=np.float64).copy()
    return out
