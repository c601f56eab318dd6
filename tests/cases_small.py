"""Small seeded cases shared by the parity tests (the oracle runs them in
well under a second each)."""
import numpy as np

from storm_amd import cases


def shuffled(case, seed):
    """The same mesh with its triangles in a seeded random order (edges
    rebuilt in the reference's order for the new numbering)."""
    import copy
    rng = np.random.default_rng(seed)
    perm = rng.permutation(case.n_elems)
    c = copy.deepcopy(case)
    c.tri = np.ascontiguousarray(case.tri[perm])
    c.edge_p, c.edge_e = cases.build_edges(c.tri, c.n_pts)
    for k in ("cd", "h", "u", "v"):
        setattr(c, k, np.ascontiguousarray(getattr(case, k)[perm]))
    c.meta["shuffled"] = seed
    return c


def vtx_data_case():
    """A flume whose initial data are given where the reference's other
    branch of dataprepFVT.f90:73-140 reads them: the bed at the cells, depth,
    velocities and Manning n at the vertices; free outflow (htype 0)."""
    c = cases.flume(nx=16, ny=6, length=100.0, width=30.0, seed=51, dt_op=0, cfl=0.05,
                    istop=30, q_in=12.0, out_kind=0)
    rng = np.random.default_rng(5151)
    zc = c.zvtx[c.tri].mean(axis=1) + rng.normal(0.0, 0.01, c.n_elems)
    c.alt = {"z": zc,
             "h": np.maximum(1.2 - c.zvtx, 0.0) + rng.uniform(0.0, 0.05, c.n_pts),
             "u": 0.2 + rng.normal(0.0, 0.02, c.n_pts),
             "v": rng.normal(0.0, 0.02, c.n_pts),
             "cd": rng.uniform(0.025, 0.04, c.n_pts)}
    c.meta["alt"] = "z at cells (flume zvtx averaged + N(0,0.01^2)); h, u, v, n at vertices; seed 5151"
    return c


def small_cases():
    return {
        "flume_rk2": (cases.flume(nx=20, ny=5, length=100.0, width=25.0, seed=7, dt_op=0,
                                  cfl=0.05, istop=20, q_in=20.0), 20),
        "dambreak_wetdry": (cases.dambreak(nq=14, side=200.0, seed=5, x_dam=80.0, h_res=2.0,
                                           flux="ac", dt_op=0, cfl=0.02, istop=60, h_dry=1e-3), 60),
        "meander_crit": (cases.meander(ns=40, nn=12, length=400.0, width=120.0, amp=20.0,
                                       wavelength=200.0, chan_w=40.0, chan_d=2.0, n_bumps=5,
                                       q0=5.0, q1=20.0, t_ramp=10.0, dt_op=0, cfl=0.02,
                                       istop=80), 80),
        "flume_rk3_slip": (cases.flume(nx=20, ny=6, length=100.0, width=30.0, seed=11, dt_op=0,
                                       cfl=0.05, istop=30, q_in=20.0, rkorder=3, btype=1), 30),
        "flume_rk1_steady": (cases.flume(nx=16, ny=6, length=100.0, width=30.0, seed=12,
                                         dt_op=0, cfl=0.3, istop=40, q_in=15.0, rkorder=1,
                                         steady=True, stop_cr=1, rstop=1e-12, hvalve=False), 40),
        "flume_unlimited_vclip": (cases.flume(nx=16, ny=6, length=100.0, width=30.0, seed=13,
                                              dt_op=0, cfl=0.05, istop=25, q_in=15.0, limiter=1,
                                              vclip=True, vceiling=0.3), 25),
        "flume_zero_grad_hyst": (cases.flume(nx=16, ny=6, length=100.0, width=30.0, seed=14,
                                             dt_op=0, cfl=0.05, istop=25, q_in=15.0, limiter=0,
                                             c_threshold=3.0, h0=None, h_out=0.3), 25),
        "flume_inH_free": (cases.flume(nx=16, ny=6, length=100.0, width=30.0, seed=21,
                                       dt_op=0, cfl=0.05, istop=30, in_kind=4, out_kind=0,
                                       q_in=2.3), 30),
        "flume_q5_crit": (cases.flume(nx=16, ny=6, length=100.0, width=30.0, seed=21,
                                      dt_op=0, cfl=0.05, istop=30, in_kind=5, out_kind=2,
                                      q_in=12.0), 30),
        # higher-order least squares (ACTIVATE HDLS): the vertex-neighbour
        # molecule, unit weights; a wet/dry front and a meander with banks
        "dambreak_hdls": (cases.dambreak(nq=14, side=200.0, seed=6, x_dam=80.0, h_res=2.0,
                                         flux="ac", dt_op=0, cfl=0.02, istop=60, h_dry=1e-3,
                                         gradient="hdls"), 60),
        "meander_hdls_rk3": (cases.meander(ns=40, nn=12, length=400.0, width=120.0, amp=20.0,
                                           wavelength=200.0, chan_w=40.0, chan_d=2.0,
                                           n_bumps=5, q0=5.0, q1=20.0, t_ramp=10.0, dt_op=0,
                                           cfl=0.02, istop=60, rkorder=3, gradient="hdls"), 60),
        # a one-quad-wide strip: cell 1 has only two vertex neighbours, so
        # t2tHD takes the augmentation of elemnt_db.f90:168-197 (whose row
        # differs from one built from the molecule's own members); unlimited
        # gradients, so that every cell's gradient enters the fluxes
        "flume_strip_hdls": (cases.flume(nx=8, ny=1, length=40.0, width=5.0, seed=3, dt_op=0,
                                         cfl=0.05, istop=30, q_in=6.0, h0=0.5, slope=0.01,
                                         limiter=1, gradient="hdls"), 30),
        # steady local time stepping that converges (RMS dh < 1e-4) at step 390
        "flume_steady_converged": (cases.flume(nx=16, ny=6, length=100.0, width=30.0, seed=41,
                                               dt_op=0, cfl=0.3, istop=600, q_in=15.0,
                                               steady=True, stop_cr=1, rstop=1e-4), 600),
        # initial data at the other locations dataprepFVT reads (vtx2ctr,
        # ctr2vtx of the bed), free outflow
        "flume_vtx_data": (vtx_data_case(), 30),
        # a velocity inflow string (vtype 0: no flux routine, enforce_bcs sets
        # the vertex velocity), slip walls, free outflow
        "flume_vtype0_slip_free": (cases.flume(nx=16, ny=6, length=100.0, width=30.0, seed=52,
                                               dt_op=0, cfl=0.05, istop=30, in_kind=0,
                                               out_kind=0, q_in=0.4, btype=1), 30),
    }


def option_cases():
    """Cases for options whose reference routine is defective or not on the
    reference's active path (so only the oracle, pinned by closed forms,
    checks them): the corrected HLL and BGNVC fluxes, Rusanov, Green-Gauss
    gradients, explicit Manning friction, the corrected CFL time step, and
    stopping test 2."""
    return {
        "flume_hll_cfl": (cases.flume(nx=20, ny=6, length=100.0, width=30.0, seed=31,
                                      dt_op=1, cfl=0.9, istop=40, q_in=15.0, flux="hll"), 40),
        "dambreak_rusanov_cfl": (cases.dambreak(nq=14, side=200.0, seed=32, x_dam=80.0,
                                                h_res=2.0, flux="rusanov", dt_op=1, cfl=0.5,
                                                istop=60, h_dry=1e-4), 60),
        "flume_bgnvc": (cases.flume(nx=20, ny=6, length=100.0, width=30.0, seed=33, dt_op=0,
                                    cfl=0.05, istop=30, q_in=15.0, flux="bgnvc"), 30),
        "flume_gauss_rk3": (cases.flume(nx=20, ny=6, length=100.0, width=30.0, seed=34,
                                        dt_op=1, cfl=0.5, istop=30, q_in=15.0, rkorder=3,
                                        gradient="gauss"), 30),
        "flume_explicit_friction": (cases.flume(nx=20, ny=6, length=100.0, width=30.0,
                                                seed=35, dt_op=1, cfl=0.5, istop=30, q_in=15.0,
                                                friction=1), 30),
        # stopping test 2 (discharge and depth), converging near step 6,500
        "flume_stage_steady_stop2": (cases.flume(nx=16, ny=6, length=100.0, width=30.0,
                                                 seed=21, dt_op=0, cfl=0.3, istop=20000,
                                                 in_kind=4, out_kind=1, q_in=2.3, h_out=2.0,
                                                 h0=2.0, steady=True, stop_cr=2, rstop=1e-5,
                                                 limiter=0, flux="hll"), 20000),
    }


def xsf_cases():
    """Cases of the invisc_fluxes_xsf.f90 interior-edge loop (option fluxes
    "xsf"), pinned by the reference linked with that file in place of
    invisc_fluxes.f90: wet/dry fronts (dry sides keep their depth, ALGORITHM
    I1 walls), a fully wet flume (the left state's y-gradients taken from the
    right cell) and a meander with banks at RK3."""
    return {
        "dambreak_xsf": (cases.dambreak(nq=14, side=200.0, seed=8, x_dam=80.0, h_res=2.0,
                                        flux="ac", dt_op=0, cfl=0.02, istop=60, h_dry=1e-3,
                                        fluxes="xsf"), 60),
        "flume_xsf": (cases.flume(nx=20, ny=5, length=100.0, width=25.0, seed=9, dt_op=0,
                                  cfl=0.05, istop=30, q_in=20.0, fluxes="xsf"), 30),
        "meander_xsf_rk3": (cases.meander(ns=40, nn=12, length=400.0, width=120.0, amp=20.0,
                                          wavelength=200.0, chan_w=40.0, chan_d=2.0, n_bumps=5,
                                          q0=5.0, q1=20.0, t_ramp=10.0, dt_op=0, cfl=0.02,
                                          istop=60, rkorder=3, seed=10, fluxes="xsf"), 60),
    }


def _with_wind(c, at_vtx, seed, coef=2.0):
    rng = np.random.default_rng(seed)
    c.wind = cases.wind_data(rng, c.n_pts if at_vtx else c.n_elems, c=coef)
    c.meta["wind"] = "wind_data(seed %d, %s, c=%g)" % (seed, "vertices" if at_vtx else "cells", coef)
    return c


def _meander_culverts():
    """Culverts on the small meander (centreline y = 20 sin(2 pi x / 200),
    channel |n| < 20, bank-full stage): a channel-to-floodplain culvert
    whose head moves through its curve, one whose inlet sits on the dry
    floodplain (never runs), one whose head is above its curve (its last
    discharge), one below its first head (its first discharge), and two
    sharing an outlet cell with opposite flows."""
    C = cases.Culvert
    return [
        C((100.0, 0.0, -1.0), (150.0, 20.0), np.array([0.2, 0.6, 1.2, 2.5]),
          np.array([0.4, 1.5, 3.0, 5.0])),
        C((60.0, 60.0, 3.0), (250.0, 70.0), np.array([0.1, 1.0]), np.array([1.0, 2.0])),
        C((300.0, 0.0, -1.0), (330.0, -35.0), np.array([0.05, 0.1]), np.array([0.5, 0.8])),
        C((200.5, 1.5, -1.0), (251.0, 40.0), np.array([5.0, 6.0]), np.array([0.3, 0.9])),
        C((120.0, -12.0, -1.2), (251.0, 40.0), np.array([0.5, 1.3, 2.0]),
          np.array([0.2, 0.25, 0.7])),
    ]


def source_cases():
    """The optional source branches of solverFVT.f90:218-256 against the
    reference: wind (given at the vertices on a closed wet/dry basin, and at
    the cells), culverts on a meander with banks (wet and dry inlets, heads
    below, inside and above the curves, a shared outlet), residual smoothing
    (ACTIVATE RSAVG: the bare option, so epsRSAVG 0.5 and nRSAVG 2) in a
    steady flume, with HDLS gradients (the smoothing keeps elemnt_db's
    t2t), and all three together at RK3."""
    mkw = dict(ns=40, nn=12, length=400.0, width=120.0, amp=20.0, wavelength=200.0,
               chan_w=40.0, chan_d=2.0, n_bumps=5, q0=5.0, q1=20.0, t_ramp=10.0, dt_op=0,
               cfl=0.02)
    basin = _with_wind(cases.dambreak(nq=14, side=200.0, seed=61, x_dam=80.0, h_res=2.0,
                                      flux="ac", dt_op=0, cfl=0.02, istop=60, h_dry=1e-3),
                       True, 6161)
    flume_w = _with_wind(cases.flume(nx=20, ny=5, length=100.0, width=25.0, seed=62, dt_op=0,
                                     cfl=0.05, istop=30, q_in=20.0), False, 6262)
    mc = cases.meander(istop=80, seed=63, **mkw)
    mc.culverts = _meander_culverts()
    all3 = _with_wind(cases.meander(istop=60, seed=64, rkorder=3, rsavg=True, **mkw), False,
                      6464, coef=1.0)
    all3.culverts = _meander_culverts()
    return {
        "dambreak_wind_vtx": (basin, 60),
        "flume_wind_cells": (flume_w, 30),
        "meander_culverts": (mc, 80),
        "flume_rsavg_steady": (cases.flume(nx=16, ny=6, length=100.0, width=30.0, seed=65,
                                           dt_op=0, cfl=0.3, istop=60, q_in=15.0, steady=True,
                                           stop_cr=1, rstop=1e-12, rsavg=True), 60),
        "meander_hdls_rsavg": (cases.meander(istop=50, seed=66, gradient="hdls", rsavg=True,
                                             **mkw), 50),
        "meander_sources_rk3": (all3, 60),
    }


def large_cases():
    """Full-length runs against the reference (fixtures of the final state
    only): the config-1 flume (20,000 triangles, 2,000 SSP-RK2 steps) with
    the reference's own Roe flux (flux_ac) at a fixed dt of 0.025 s (CFL
    about 0.7-0.9; the reference's CFL time step reads out of bounds), and a
    200,000-triangle cut of the config-2 meander (2 km, 20 bumps) whose
    inflow ramps 500 -> 3,000 m3/s in 60 s so that the bank-full channel
    spills and wet/dry fronts move over 1,500 SSP-RK3 steps."""
    return {
        "flume20k_ac": (cases.flume(flux="ac", dt_op=0, cfl=0.025, istop=2000), 2000),
        "meander200k": (cases.meander(ns=400, length=2000.0, n_bumps=20, t_ramp=60.0, dt_op=0,
                                      cfl=0.025, istop=1500), 1500),
    }


def golden_cases():
    """Every case with a golden fixture from the reference."""
    out = dict(small_cases())
    out.update(xsf_cases())
    out.update(source_cases())
    return out
