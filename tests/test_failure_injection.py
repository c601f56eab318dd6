"""Failure injection through the C ABI, on the host (no GPU needed): broken
meshes, boundary strings and options reach storm_fvt_create, which must
refuse them with the matching STORM_E_* code and a message naming the
fault (printed on stderr, as the reference's byebye path prints its own)
before any device work; null handles are refused by every entry point."""
import ctypes as C
import copy

import numpy as np
import pytest

from storm_amd import cases, fvt

E_ARG, E_MESH, E_DEVICE, E_STATE = -1, -2, -3, -5


def base_case():
    return cases.flume(nx=6, ny=3, length=30.0, width=9.0, seed=5, dt_op=0, cfl=0.05, istop=5)


def create(case, capfd, **opts):
    keep = {}
    m = fvt.mesh_struct(case, keep)
    o = dict(case.opts)
    o.update(opts)
    h = C.c_void_p()
    rc = fvt.load().storm_fvt_create(C.byref(m), C.byref(fvt.opts_struct(o)), C.byref(h))
    err = capfd.readouterr().err
    if h.value:
        fvt.load().storm_fvt_destroy(h)
    return rc, err


def rebuilt(c):
    c.edge_p, c.edge_e = cases.build_edges(c.tri, c.n_pts)
    return c


def test_too_few_triangles(capfd):
    c = cases.dambreak(nq=1, side=10.0, seed=1, n_bumps=1, x_dam=5.0, h_res=1.0)  # 2 triangles
    rc, err = create(c, capfd)
    assert rc == E_MESH and "too few triangles" in err


def test_vertex_id_out_of_range(capfd):
    c = copy.deepcopy(base_case())
    c.tri = c.tri.copy()
    c.tri[3, 1] = c.n_pts + 7
    rc, err = create(c, capfd)
    assert rc == E_MESH and "vertex id out of range" in err


def test_clockwise_triangle(capfd):
    c = copy.deepcopy(base_case())
    c.tri = c.tri.copy()
    c.tri[5, [1, 2]] = c.tri[5, [2, 1]]
    rc, err = create(rebuilt(c), capfd)
    assert rc == E_MESH and "oriented" in err


def test_disconnected_triangle(capfd):
    c = copy.deepcopy(base_case())
    P = c.n_pts
    c.x = np.concatenate([c.x, [100.0, 101.0, 100.0]])
    c.y = np.concatenate([c.y, [100.0, 100.0, 101.0]])
    c.zvtx = np.concatenate([c.zvtx, [0.0, 0.0, 0.0]])
    c.tri = np.vstack([c.tri, [[P, P + 1, P + 2]]]).astype(np.int32)
    for k in ("cd", "h", "u", "v"):
        setattr(c, k, np.concatenate([getattr(c, k), [getattr(c, k)[0]]]))
    rc, err = create(rebuilt(c), capfd)
    assert rc == E_MESH and ("connectivity" in err or "wall" in err)


def test_edge_table_entry_out_of_range(capfd):
    c = copy.deepcopy(base_case())
    c.edge_e = c.edge_e.copy()
    c.edge_e[2, 0] = c.n_elems + 3
    rc, err = create(c, capfd)
    assert rc == E_MESH and "edge table entry out of range" in err


def test_non_contiguous_boundary_string(capfd):
    c = copy.deepcopy(base_case())
    b = c.inflows[0]
    b.nodes = np.concatenate([b.nodes[:1], b.nodes[2:]])  # skips a node: no shared edge
    rc, err = create(c, capfd)
    assert rc == E_MESH and "boundary string nodes" in err


def test_string_through_the_interior(capfd):
    c = copy.deepcopy(base_case())
    ny = 3
    # the column of nodes at x index 2 (node id i*(ny+1)+j) runs through the interior
    c.inflows[0].nodes = np.array([2 * (ny + 1) + j for j in range(ny + 1)], dtype=np.int32)
    rc, err = create(c, capfd)
    assert rc == E_MESH and "interior" in err


@pytest.mark.parametrize("opts,code,text", [
    (dict(rkorder=4), E_ARG, "Runge-Kutta order"),
    (dict(fluxes="xsf", flux="hll"), E_ARG, "flux_ac"),
    (dict(rsavg=True, rsavg_n=-1), E_ARG, "invalid option value"),
    (dict(h_dry=0.0), E_ARG, "invalid threshold"),
    (dict(limiter=3), E_ARG, "invalid option value"),
])
def test_invalid_options(capfd, opts, code, text):
    rc, err = create(base_case(), capfd, **opts)
    assert rc == code and text in err


@pytest.mark.parametrize("vtype,text", [(2, "VTYPE = 2"), (3, "VTYPE = 3"), (7, "vtype")])
def test_unsupported_inflow_types(capfd, vtype, text):
    c = copy.deepcopy(base_case())
    c.inflows[0].kind = vtype
    rc, err = create(c, capfd)
    assert rc == E_ARG and text in err


def test_valid_mesh_without_a_device(capfd):
    """On a host without a GPU a valid mesh passes every check and stops at
    the device."""
    rc, err = create(base_case(), capfd)
    if rc == 0:
        pytest.skip("a HIP device is present")
    assert rc == E_DEVICE and "no HIP device" in err


def test_null_handles_are_refused():
    L = fvt.load()
    it, tt, rm = C.c_int32(0), C.c_double(0), C.c_double(0)
    assert L.storm_fvt_run(None, 1, C.byref(it), C.byref(tt), C.byref(rm)) == E_ARG
    assert L.storm_fvt_finalize(None) == E_ARG
    assert L.storm_fvt_set_state(None, None, None, None) == E_ARG
    assert L.storm_fvt_set_wind(None, 0, None, None, None) == E_ARG
    assert L.storm_fvt_set_culverts(None, 0, None, None, None, None, None) == E_ARG
    assert L.storm_fvt_get_culverts(None, None, None, None) == E_ARG
    assert L.storm_fvt_get_state(None, None, None, None, None, None) == E_ARG
    assert L.storm_fvt_get_gradients_noalloc(None, None, None) == E_ARG
