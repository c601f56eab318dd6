"""Oracle functions against single routines of the reference, run by the
per-routine harness oracle/ref_units.f90 (fixture tests/golden/units.npz,
written by tests/golden/make_units.py), and the closed forms that pin what
the reference cannot: the HLL star region with a dry side, and the
Green-Gauss ghost cell at walls."""
import math
import os

import numpy as np
import pytest

from cases_small import small_cases
from storm_amd import cases

UNITS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "units.npz")
G = 9.81


@pytest.fixture(scope="module")
def units():
    return np.load(UNITS)


def test_rusanov_is_the_reference_routine(units):
    """flux_rusanov.f90:1-82, bit for bit, dry sides included."""
    import oracle as O
    for x, f in zip(units["rusanov_in"], units["rusanov_out"]):
        nx, ny, hl, ul, vl, cl, hr, ur, vr, cr = x
        got = O.flux("rusanov", nx, ny, hl, ul, vl, hr, ur, vr, cl, cr)
        assert np.array_equal(got, f), (x, got, f)


def _ref_hll_speeds(x):
    # flux_HLL.f90:42-68 as written (the both-wet right speed subtracts c_R)
    nx, ny, hl, ul, vl, hr, ur, vr, h0 = x
    cl, cr = math.sqrt(G * hl), math.sqrt(G * hr)
    us = 0.5 * ((ul + ur) * nx + (vl + vr) * ny) + cl - cr
    cs = 0.5 * (cl + cr) + ((ul - ur) * nx + (vl - vr) * ny) / 4.0
    unl, unr = ul * nx + vl * ny, ur * nx + vr * ny
    if hl < h0:
        return unr - 2.0 * cr, unr + cr
    if hr < h0:
        return unl - cl, unl + 2.0 * cl
    return min(unl - cl, us - cs), max(unr - cr, us + cs)


def test_hll_dry_bed_speeds_and_upwind_fluxes_are_the_reference_routine(units):
    """flux_HLL.f90 on states whose waves both run one way.  With a dry side
    (hl or hr below h0) the dry-bed speeds of :50-69 decide the branch, and
    the oracle takes the reference's branch and flux bit for bit.  Both
    sides wet, the left-upwind branch (s_L > 0) is the reference's too.  The
    reference's both-wet s_R is u_R.n - c_R (:64) where Yoon & Kang eq. 8
    has u_R.n + c_R; the oracle keeps eq. 8, so right-upwind wet states are
    compared with the closed forms below instead."""
    import oracle as O
    n_dry = n_left = 0
    for x, f in zip(units["hll_in"], units["hll_out"]):
        nx, ny, hl, ul, vl, hr, ur, vr, h0 = x
        sl, sr = _ref_hll_speeds(x)
        dry = hl < h0 or hr < h0
        if not dry and not sl > 0:
            continue
        got = O.flux("hll", nx, ny, hl, ul, vl, hr, ur, vr, 0.0, 0.0, h0)
        assert np.array_equal(got, f), (x, got, f)
        n_dry += dry
        n_left += not dry
    assert n_dry >= 150 and n_left >= 40


def _phys(h, u, v, nx, ny):
    return np.array([h * u * nx + h * v * ny,
                     (h * u * u + 0.5 * G * h * h) * nx + h * u * v * ny,
                     h * u * v * nx + (h * v * v + 0.5 * G * h * h) * ny])


@pytest.mark.parametrize("side", ["left", "right"])
def test_hll_star_region_with_a_dry_side(side):
    """Yoon & Kang (2004) eqs. 6-8 with the dry-bed speeds: left dry,
    s_L = u_R.n - 2 c_R, s_R = u_R.n + c_R; right dry, s_L = u_L.n - c_L,
    s_R = u_L.n + 2 c_L; F = (s_R F_L - s_L F_R + s_L s_R (q_R - q_L)) /
    (s_R - s_L).  (A speed with c for 2c, or the other side's velocity,
    moves the flux far beyond the tolerance.)"""
    import oracle as O
    rng = np.random.default_rng(11)
    n = 0
    for _ in range(200):
        a = rng.uniform(0, 2 * math.pi)
        nx, ny = math.cos(a), math.sin(a)
        h = rng.uniform(0.1, 3.0)
        u, v = rng.normal(0, 0.7, 2)
        c = math.sqrt(G * h)
        un = u * nx + v * ny
        if side == "left":
            hl, ul, vl, hr, ur, vr = 0.0, 0.0, 0.0, h, u, v
            sl, sr = un - 2 * c, un + c
        else:
            hl, ul, vl, hr, ur, vr = h, u, v, 0.0, 0.0, 0.0
            sl, sr = un - c, un + 2 * c
        if not (sl < -1e-3 and sr > 1e-3):
            continue
        n += 1
        fl, fr = _phys(hl, ul, vl, nx, ny), _phys(hr, ur, vr, nx, ny)
        q = np.array([hr - hl, hr * ur - hl * ul, hr * vr - hl * vl])
        ref = (sr * fl - sl * fr + sl * sr * q) / (sr - sl)
        got = O.flux("hll", nx, ny, hl, ul, vl, hr, ur, vr, 0.0, 0.0, 1e-3)
        np.testing.assert_allclose(got, ref, rtol=1e-13, atol=1e-13 * (1 + np.abs(ref).max()))
    assert n > 100


def test_explicit_friction_is_the_reference_routine(units):
    """fric_terms.f90:22-98 (Manning), bit for bit; the oracle's step runs
    the same function."""
    import oracle as O
    c = units["fric_in"]
    p2, p3 = O.fric_explicit(units["fric_dry"], c[:, 0], c[:, 1], c[:, 2], c[:, 3], c[:, 4],
                             c[:, 5], c[:, 6], c[:, 7])
    assert np.array_equal(p2, units["fric_out"][0])
    assert np.array_equal(p3, units["fric_out"][1])


@pytest.mark.parametrize("name", ["meander_crit", "dambreak_wetdry"])
def test_rds_node_routines_are_the_reference(units, name):
    """fvsources (node least squares over n2n, the median-dual integral with
    its literal c1), part_dry and drynwet (the lone-dry-vertex-3 case
    included) on vertex depths with dry nodes, bit for bit."""
    import oracle as O
    case, _ = small_cases()[name]
    src, zo, pd = O.Oracle(case).rds(case.opts["h_dry"], case.zvtx, units[f"rds_{name}_h"])
    assert np.array_equal(src, units[f"rds_{name}_source"])
    assert np.array_equal(zo, units[f"rds_{name}_zout"])
    assert np.array_equal(pd, units[f"rds_{name}_pd"])
    assert ((pd[:, 1] == 1) & (pd[:, 2] == 3)).sum() > 10 and (pd[:, 1] == 2).sum() > 10


def test_gauss_gradient_ghost_cells_mirror_the_centroid():
    """At a wall side the Green-Gauss molecule's ghost point is the cell
    centroid mirrored across the side (ghost_cell.f90:34-41), carrying the
    cell's own value.  A linear field constant across the straight walls
    y = 0 and y = W (phi = a + b x) takes the same value at the centroid and
    its mirror, so the gradient is exact at wall cells too -- unless the
    ghost point is elsewhere (e.g. the side's midpoint)."""
    import oracle as O
    c = cases.flume(nx=14, ny=6, length=70.0, width=30.0, seed=23, dt_op=0, cfl=0.01, istop=1,
                    gradient="gauss", limiter=1, rough_sd=0.0, slope=0.0)
    x = c.x[c.tri].mean(axis=1)
    zc = c.zvtx[c.tri].mean(axis=1)
    c.h = (3.0 + 0.01 * x) - zc
    c.u = 0.5 + 0.003 * x
    c.v = -0.2 + 0.002 * x
    o = O.Oracle(c)
    g = o.cell_grads()
    _, _, t3 = o.topology()
    # cells (and their neighbours) clear of the inflow and outflow ends
    far = (x > 12.0) & (x < 58.0)
    wall = far & np.any(t3 < 0, axis=1)
    assert wall.sum() > 10
    np.testing.assert_allclose(g[far, 0], 0.01, atol=1e-12)
    np.testing.assert_allclose(g[far, 1], 0.0, atol=1e-12)
    np.testing.assert_allclose(g[far, 2], 0.003, atol=1e-12)
    np.testing.assert_allclose(g[far, 4], 0.002, atol=1e-12)
