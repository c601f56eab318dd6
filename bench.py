#!/usr/bin/env python3
"""Benchmark of the MI355X FVT time step on the single-reach workload.

Workload (one GPU): sinusoidal meandering channel, 20 km long, 2,000,000
jittered triangles (4000 x 250 quads, each split in two), parabolic main
channel 80 m x 4 m, 1 km floodplain initially dry with 200 Gaussian bumps,
Manning n 0.035 / 0.06, inflow ramp 500 -> 3000 m^3/s over 1 h, critical-depth
outflow, SSP-RK3, CFL 0.9 with the time step from the CFL reduction, moving
wet/dry fronts.  A step is one full SSP-RK3 time step of the whole reach
(three stages, implicit friction, CFL reduction, residual); the state is
resident in HBM when timing starts.

Prints one JSON line.  `value` is cell-updates/s (triangles x steps / s).
`wet_valley` times the same reach, the same way, with the whole valley under
water (every cell wet: each forms its gradients, limiter, faces and fluxes,
where a dry cell skips them), the stage kernel's harder case.

`--workload ensemble` instead runs the batched ensemble (storm_amd/cases.py
`ensemble_reach`: reaches of log-uniform size in [5k, 200k] triangles, own
slope, roughness, discharge and stage) from a cold start to steady state
(RMS dh <= 1e-6 in every reach; a reach that diverges stops and is counted
apart) in one context, with the autotuned launch configuration, and reports
the wall time to steady and the cell-updates/s.

`--workload warmstart` times held-out reaches (200k triangles each) to steady
state from a cold start, from the plane guess, from the 1-D backwater start
and from the learned warm start (storm_amd/warmstart.py), and reports the
cold / learned wall-time ratio.
`--workload warmstart-train` times AdamW updates of the warm-start network
(BF16 MFMA, FP32 accumulation) on batches of 32 windows of 4,096 cells cut
from 64 training reaches, against synthetic targets.
`roofline` prices the step with the algorithmic bytes of one cell-update of
SSP-RK3 (1,248 B: three stages of 384 B average plus 96 B of friction, dt and
residual) against 8 TB/s; one "launch" is one time step.  `cpu_baseline`
times the reference Fortran itself (oracle/_ref/ref_driver, built from the
reference's sources with flang -O2, serial, as the reference runs) on a
bounded sample of the same workload: the meander generator cut to 4 km
(400,000 triangles, the same cross-section, cell size and bump density) for
20 steps, timing its solverFVT call alone.  Where that build is absent it
times the CPU oracle (a serial C restatement of the reference step) instead;
the oracle's figure on 10 steps of the full 2M reach is reported beside it.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

B_STEP_RK3 = 1248.0          # algorithmic bytes per cell-update, SSP-RK3
B_STAGE_RK3 = 384.0          # algorithmic bytes per cell-stage, SSP-RK3 average
PEAK_HBM = 8.0e12            # B/s, MI355X
METRIC = "FP64 cell-updates/s + HBM % roofline; time-to-steady warm vs cold"


REF_SAMPLE = dict(ns=800, length=4000.0, n_bumps=40)  # meander cut to 4 km, 400k triangles
REF_STEPS = 20


def cpu_reference(steps=REF_STEPS):
    """The flang-built reference on the bounded meander sample: cell-updates
    per second of its solverFVT call, and that call's seconds."""
    import subprocess
    import tempfile
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # the baseline only
    from storm_amd import cases
    if not os.path.exists(O.REF_DRIVER):
        print("bench.py: %s is not built; timing the oracle only" % O.REF_DRIVER,
              file=sys.stderr)
        return None
    case = cases.meander(istop=steps, **REF_SAMPLE)
    with tempfile.TemporaryDirectory() as d:
        run = cases.write_run(case, d)
        r = subprocess.run([O.REF_DRIVER, os.path.basename(run), "out.bin"], cwd=d,
                           capture_output=True, text=True, timeout=600)
        t = [float(l.split()[-1]) for l in r.stdout.splitlines()
             if l.startswith("solverFVT seconds")]
        if r.returncode != 0 or not t:
            print("bench.py: the reference run failed (exit %d): %s" %
                  (r.returncode, (r.stdout + r.stderr)[-2000:]), file=sys.stderr)
            return None
    return case.n_elems * steps / t[0], t[0]


def cpu_baseline(case, steps):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # the checker/baseline only
    O.build()
    o = O.Oracle(case, dict(istop=10 ** 9))
    o.run(1)  # warm-up step (first-touch of the arrays)
    t0 = time.perf_counter()
    o.run(steps)
    t = time.perf_counter() - t0
    return case.n_elems * steps / t, t


KERNEL_SOURCES = ("storm_amd/csrc/fvt_device.hip", "storm_amd/csrc/fvt_math.h",
                  "storm_amd/csrc/fvt_prep.cpp", "storm_amd/csrc/fvt_prep.h")


def kernel_source_hash():
    """SHA-1 (first 16 hex digits) of the sources the step's kernels are
    built from: a traffic figure measured on other sources is not reported."""
    import hashlib
    h = hashlib.sha1()
    for rel in KERNEL_SOURCES:
        with open(os.path.join(ROOT, rel), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def traffic_from_profiles(workload):
    """HBM bytes of one step as rocprofv3's FETCH_SIZE/WRITE_SIZE counted
    them (profiles/traffic_meander2m.json, one entry per workload), when that
    measurement was taken on the current kernel sources; else None."""
    path = os.path.join(ROOT, "profiles", "traffic_meander2m.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            d = json.load(f)
        e = d.get("workloads", {}).get(workload)
        if e and e.get("kernel_sources") == kernel_source_hash():
            return e.get("bytes_per_step")
    except Exception:
        return None
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-wet", action="store_true",
                    help="meander: skip the wet-valley line (the same reach with every cell wet)")
    ap.add_argument("--graph", type=int, default=0,
                    help="1: replay whole steps from a captured HIP graph")
    ap.add_argument("--block", type=int, default=0,
                    help="stage-kernel tile capacity (256, 384 or 512; 0 = library default)")
    ap.add_argument("--workload", choices=["meander", "ensemble", "warmstart", "warmstart-train",
                                           "dambreak", "flume"],
                    default="meander")
    ap.add_argument("--t-end", type=float, default=300.0, help="dambreak: simulated seconds")
    ap.add_argument("--reaches", type=int, default=256,
                    help="ensemble: number of reaches; warmstart: held-out reaches (default 64)")
    ap.add_argument("--max-steps", type=int, default=60000, help="ensemble/warmstart: step limit")
    ap.add_argument("--size", type=int, default=200000, help="warmstart: triangles per reach")
    ap.add_argument("--rstop", type=float, default=1e-6, help="warmstart: RMS dh tolerance")
    ap.add_argument("--checkpoint", default="", help="warmstart: trained network (.npz)")
    ap.add_argument("--start", choices=["cold", "plane", "gvf", "learned"], default="cold",
                    help="ensemble: initial state (plane, gvf (the 1-D backwater profile) and "
                         "learned: storm_amd/warmstart.py, their set-up inside the timed region)")
    ap.add_argument("--accuracy", action="store_true",
                    help="warmstart: time to a fixed accuracy against tightly converged "
                         "states (on --reaches reaches, default 8) instead of the RMS dh test")
    ap.add_argument("--modes", default="cold,plane,gvf,learned",
                    help="warmstart: starts to time (the ratio needs cold and learned)")
    ap.add_argument("--wet", type=float, default=None,
                    help="meander: start with the whole valley under water, this many metres "
                         "above the valley floor (4 puts every cell under water)")
    ap.add_argument("--limiter", type=int, default=2,
                    help="ensemble: LIMITER (2 Barth-Jespersen, the stated scheme; 0 first order)")
    ap.add_argument("--reorder", type=int, default=1,
                    help="meander: device cell order (1 Hilbert, 2 reverse Cuthill-McKee, 0 input)")
    ap.add_argument("--rsavg", type=int, default=0,
                    help="meander: residual smoothing (ACTIVATE RSAVG) with this many sweeps")
    ap.add_argument("--calib-copy", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (counter "
                         "calibration for rocprofv3 FETCH_SIZE/WRITE_SIZE runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: the FVT step runs on one GPU only; --gpus %d is not supported "
              "(a reach would shard into contiguous ranges of tiles with a two-cell halo)."
              % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    from storm_amd import cases, fvt
    if args.workload == "ensemble":
        ensemble(args)
        return
    if args.workload == "warmstart":
        if args.accuracy:
            warmstart_accuracy(args)
        else:
            warmstart(args)
        return
    if args.workload == "warmstart-train":
        warmstart_train(args)
        return
    if args.workload == "dambreak":
        dambreak(args)
        return
    if args.workload == "flume":
        flume(args)
        return
    if args.calib_copy > 0:
        import torch
        nbytes = args.calib_copy << 20
        src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
    case = cases.meander(wet=args.wet)
    workload = "meander-2M-rk3" if args.wet is None else "meander-2M-rk3-wet"
    extra = {}
    if args.rsavg > 0:
        workload += "-rsavg%d" % args.rsavg
        extra = dict(rsavg=True, rsavg_n=args.rsavg)
    t0 = time.perf_counter()
    if args.reorder != 1:
        workload += "-order%d" % args.reorder
        extra["reorder"] = args.reorder
    solver = fvt.FVTSolver(case, dict(istop=10 ** 9, graph=args.graph, block=args.block, **extra))
    setup_s = time.perf_counter() - t0
    solver.run(args.warmup)
    t0 = time.perf_counter()
    it, ttime, rmax = solver.run(args.steps)
    wall = time.perf_counter() - t0
    gpu_s = solver.last_run_seconds()
    wet_frac = float((solver.state()["h"] > case.opts["h_dry"]).mean())
    n = case.n_elems
    value = n * args.steps / wall
    step_s = gpu_s / args.steps
    achieved = B_STEP_RK3 * n / step_s
    traffic = traffic_from_profiles(workload)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded meander generator, storm_amd/cases.py)",
        "config": {"workload": workload, "triangles": n, "vertices": case.n_pts,
                   "scheme": "Roe flux_ac, LSQ+Barth-Jespersen, SSP-RK3, CFL 0.9",
                   "boundaries": "Q ramp 500->3000 m3/s inflow, critical outflow, walls",
                   "initial": ("channel at bank-full, floodplain dry" if args.wet is None else
                               "whole valley under water, stage %g m above the floor"
                               % args.wet),
                   "wet_fraction": round(wet_frac, 4),
                   "setup_s": round(setup_s, 3), "sim_time_s": ttime},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": PEAK_HBM / 1e9,
                     "unit": "GB/s", "frac": achieved / PEAK_HBM,
                     "traffic": traffic,
                     "kernel": "one SSP-RK3 time step (all its launches)",
                     "algorithmic_bytes_per_launch": B_STEP_RK3 * n,
                     "avg_launch_ms": 1e3 * step_s,
                     # the three stages' 384 B per cell-stage over the whole
                     # step's time (friction, dt and residual counted as
                     # overhead of the stages)
                     "frac_stage": B_STAGE_RK3 * 3 * n / step_s / PEAK_HBM,
                     # measured HBM bytes over the step time
                     "frac_traffic": (traffic / step_s / PEAK_HBM) if traffic else None,
                     "kernel_sources": kernel_source_hash()},
    }
    if args.wet is None and not args.no_wet and not extra:
        # the same reach with every cell wet (whole valley under water, 4 m
        # above the floor), timed the same way: the stage kernel's harder case
        solver.close()
        wcase = cases.meander(wet=4.0)
        ws = fvt.FVTSolver(wcase, dict(istop=10 ** 9, graph=args.graph, block=args.block))
        ws.run(args.warmup)
        t0 = time.perf_counter()
        ws.run(args.steps)
        wwall = time.perf_counter() - t0
        wstep = ws.last_run_seconds() / args.steps
        wtraffic = traffic_from_profiles("meander-2M-rk3-wet")
        out["wet_valley"] = {
            "workload": "meander-2M-rk3-wet", "ms_per_step": 1e3 * wwall / args.steps,
            "value": n * args.steps / wwall, "unit": "cell-updates/s",
            "wet_fraction": round(float((ws.state()["h"] > case.opts["h_dry"]).mean()), 4),
            "roofline_frac": B_STEP_RK3 * n / wstep / PEAK_HBM,
            "roofline_frac_stage": B_STAGE_RK3 * 3 * n / wstep / PEAK_HBM,
            "traffic": wtraffic, "avg_launch_ms": 1e3 * wstep}
        ws.close()
    if not args.no_cpu:
        cpu_rate, cpu_t = cpu_baseline(case, args.cpu_steps)
        port = {"value": cpu_rate, "unit": "cell-updates/s", "cores": 1, "kind": "port",
                "sample": "%d SSP-RK3 steps of the same 2M-triangle reach (%.1f s), serial C "
                          "oracle" % (args.cpu_steps, cpu_t)}
        ref = cpu_reference()
        if ref is None:
            out["cpu_baseline"] = port
        else:
            out["cpu_baseline"] = {
                "value": ref[0], "unit": "cell-updates/s", "cores": 1, "kind": "reference",
                "sample": "reference Fortran (flang -O2, serial), solverFVT call alone: %d "
                          "SSP-RK3 steps of the meander generator cut to 4 km (%d triangles, "
                          "same cross-section, cell size and bump density), %.1f s"
                          % (REF_STEPS, 400000, ref[1]),
                "port": port}
    print(json.dumps(out))


def ensemble(args):
    """Time to steady state of a batched ensemble from the cold start
    (warmstart.cold_state: still water covering every cell), or (--start)
    from the plane guess, the backwater profile or the learned warm start."""
    from storm_amd import cases, fvt, tune
    t0 = time.perf_counter()
    reaches = [cases.ensemble_reach(k, steady=True) for k in range(args.reaches)]
    gen_s = time.perf_counter() - t0
    # the stated scheme is least squares + Barth-Jespersen (--limiter 2);
    # --limiter 0 runs first-order HLL
    opts = dict(flux="hll", limiter=args.limiter, rkorder=2, cfl=0.9, steady=True, stop_cr=1,
                rstop=1e-6, dt_op=0, istop=args.max_steps, nonfinite_stop=True)
    if args.rsavg > 0:
        opts.update(rsavg=True, rsavg_n=args.rsavg)
    best, times = tune.tune(reaches, opts, steps=4)
    opts.update(best)
    t0 = time.perf_counter()
    b = fvt.BatchSolver(reaches, opts)
    setup_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    if args.start != "cold":
        from storm_amd import warmstart as W
        if args.start == "plane":
            init = W.starts(reaches, "plane")
        elif args.start == "gvf":
            init = W.starts(reaches, "gvf")
        else:
            ws = W.WarmStart.load(args.checkpoint or W.DEFAULT_CHECKPOINT)
            init = ws.predict(reaches)
            ws.close()
        b.set_state(np.concatenate([x[0] for x in init]), np.concatenate([x[1] for x in init]),
                    np.concatenate([x[2] for x in init]))
    else:
        from storm_amd import warmstart as W
        init = W.starts(reaches, "cold")
        b.set_state(np.concatenate([x[0] for x in init]), np.concatenate([x[1] for x in init]),
                    np.concatenate([x[2] for x in init]))
    init_s = time.perf_counter() - t0
    gpu_s = 0.0
    while True:
        b.run(1000)
        gpu_s += b.last_run_seconds()
        it, tt, rm, dn = b.status()
        print("ensemble: %d/%d reaches stopped, max step %d" % (dn.sum(), len(dn), it.max()),
              file=sys.stderr, flush=True)
        if dn.all() or it.max() >= args.max_steps:
            break
    wall = time.perf_counter() - t0
    n = np.array([c.n_elems for c in reaches])
    updates = float((n * it).sum())
    out = {
        "metric": "time-to-steady (%s start), batched ensemble" % args.start,
        "start": args.start, "init_s": init_s,
        "value": wall, "unit": "s", "n_gpus": 1, "higher_is_better": False,
        "cell_updates_per_s": updates / wall, "gpu_seconds": gpu_s,
        "reaches": len(reaches), "triangles": int(n.sum()),
        "converged": int((dn & np.isfinite(rm)).sum()), "diverged": int((~np.isfinite(rm)).sum()),
        "diverged_reaches": [int(k) for k in np.where(~np.isfinite(rm))[0]],
        "capped_reaches": [int(k) for k in np.where(~dn)[0]],
        "steps_max": int(it.max()), "steps_median": float(np.median(it)),
        "dtype": "f64", "data": "synthetic (storm_amd/cases.py ensemble_reach, seeds 4004+k)",
        "config": {"workload": "ensemble-%d" % len(reaches),
                   "scheme": "HLL %s, SSP-RK2, steady local dt CFL 0.9%s, stop RMS dh <= 1e-6"
                   % ("first order" if args.limiter == 0 else
                      "LSQ + Barth-Jespersen" if args.limiter == 2 else "LSQ unlimited",
                      ", RSAVG %d sweeps" % args.rsavg if args.rsavg else ""), "launch": best,
                   "tune_s_per_step": times, "generate_s": round(gen_s, 1),
                   "setup_s": round(setup_s, 1)},
    }
    print(json.dumps(out))


def flume(args):
    """The straight flume in full: 20,000 triangles, 2,000 steps (HLL, LSQ +
    Barth-Jespersen, SSP-RK2, CFL 0.9 with dt from the CFL reduction), from
    the same state on the GPU and, for the CPU baseline, through the
    reference Fortran on the same .run case (its solverFVT call timed)."""
    import subprocess
    import tempfile
    from storm_amd import cases, fvt
    steps = 2000
    case = cases.flume(flux="hll", istop=steps)
    s = fvt.FVTSolver(case, dict(istop=steps, graph=args.graph))
    t0 = time.perf_counter()
    it, tt, _ = s.run(steps)
    wall = time.perf_counter() - t0
    gpu_s = s.last_run_seconds()
    s.close()
    n = case.n_elems
    out = {
        "metric": "FP64 cell-updates/s, flume (2,000 steps)", "value": n * it / wall,
        "unit": "cell-updates/s", "n_gpus": 1, "higher_is_better": True, "steps": it,
        "wall_s": wall, "gpu_seconds": gpu_s, "sim_time_s": tt, "dtype": "f64",
        "data": "synthetic (storm_amd/cases.py flume, seed 1001)",
        "config": {"workload": "flume-20k", "triangles": n,
                   "scheme": "HLL, LSQ+Barth-Jespersen, SSP-RK2, CFL 0.9", "graph": args.graph},
    }
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O  # the baseline only
        if os.path.exists(O.REF_DRIVER):
            with tempfile.TemporaryDirectory() as d:
                run = cases.write_run(case, d)
                r = subprocess.run([O.REF_DRIVER, os.path.basename(run), "out.bin"], cwd=d,
                                   capture_output=True, text=True, timeout=600)
                ts = [float(l.split()[-1]) for l in r.stdout.splitlines()
                      if l.startswith("solverFVT seconds")]
            if r.returncode == 0 and ts:
                out["cpu_baseline"] = {"value": n * steps / ts[0], "unit": "cell-updates/s",
                                       "cores": 1, "kind": "reference",
                                       "sample": "the whole run through the reference Fortran "
                                                 "(flang -O2, serial), solverFVT call alone, "
                                                 "%.1f s" % ts[0]}
            else:
                print("bench.py: the reference run failed (exit %d)" % r.returncode,
                      file=sys.stderr)
    print(json.dumps(out))


def cell_areas(case):
    x, y, t = case.x, case.y, case.tri
    return 0.5 * np.abs((x[t[:, 1]] - x[t[:, 0]]) * (y[t[:, 2]] - y[t[:, 0]]) -
                        (x[t[:, 2]] - x[t[:, 0]]) * (y[t[:, 1]] - y[t[:, 0]]))


def dambreak(args):
    """The dam break in full: 1M triangles from rest until the simulated time
    reaches t_end (default 300 s), Rusanov flux, SSP-RK2, CFL 0.5 with dt from
    the CFL reduction, h_dry 1e-4, closed walls.  The run stops at the first
    step that ends at or past t_end.  Reports the wall time, the steps, the
    cell-updates/s and the change of the water volume (closed walls: only
    the depth clip at the wet/dry front changes it)."""
    from storm_amd import cases, fvt
    case = cases.dambreak()
    area = cell_areas(case)
    v0 = float((case.h * area).sum())
    t0 = time.perf_counter()
    s = fvt.FVTSolver(case, dict(istop=10 ** 9, graph=args.graph))
    setup_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    gpu_s, it, tt = 0.0, 0, 0.0
    it0, tt0 = 0, 0.0
    while tt < args.t_end:
        # steps to go at the last chunk's mean dt (dt shrinks as the flow
        # speeds up), less a margin; the last few are run one at a time
        if it == 0:
            n = 1000
        else:
            n = int(0.9 * (args.t_end - tt) / ((tt - tt0) / (it - it0)))
            n = max(1, min(n, 5000))
        it0, tt0 = it, tt
        it, tt, _ = s.run(n)
        gpu_s += s.last_run_seconds()
        print("dambreak: step %d, t = %.3f s" % (it, tt), file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    st = s.state()
    s.close()
    v1 = float((st["qh"] * area).sum())
    n = case.n_elems
    out = {
        "metric": "time to t = %g s, dam break over a dry bed" % args.t_end,
        "value": wall, "unit": "s", "n_gpus": 1, "higher_is_better": False,
        "steps": it, "sim_time_s": tt, "gpu_seconds": gpu_s,
        "cell_updates_per_s": n * it / wall, "volume_rel_change": (v1 - v0) / v0,
        "wet_fraction": float((st["h"] > case.opts["h_dry"]).mean()),
        "dtype": "f64", "data": "synthetic (storm_amd/cases.py dambreak, seed 3003)",
        "config": {"workload": "dambreak-1M", "triangles": n,
                   "scheme": "Rusanov, LSQ+Barth-Jespersen, SSP-RK2, CFL 0.5, h_dry 1e-4",
                   "setup_s": round(setup_s, 3)},
    }
    print(json.dumps(out))


def warmstart_accuracy(args):
    """Time to a fixed accuracy from each start.  The reference is each
    reach's state converged to RMS dh <= 1e-9 (from the backwater start).  Each
    start then runs, with the RMS dh test off, until every reach's RMS error
    in h against its reference is at most 1 cm and then 1 mm; the state is
    read back every 500 steps, so step counts are multiples of 500.  The GPU
    time counts the steps only."""
    import ctypes as C
    from storm_amd import fvt
    from storm_amd import warmstart as W
    n_r = args.reaches if args.reaches != 256 else 8
    ckpt = args.checkpoint or W.DEFAULT_CHECKPOINT
    reaches = [W.heldout_reach(k, args.size) for k in range(n_r)]
    t0 = time.perf_counter()
    ref, _ = W.steady_states(reaches, dict(rstop=1e-9, nonfinite_stop=True), 400000,
                             chunk=5000, init=W.starts(reaches, "gvf"),
                             log=lambda m: print("reference: %s" % m, file=sys.stderr, flush=True))
    ref_s = time.perf_counter() - t0
    keep = [r for r, st in enumerate(ref) if st[0] is not None]
    reaches = [reaches[r] for r in keep]
    href = [ref[r][0]["h"] for r in keep]
    tols = (1e-2, 1e-3)
    res = {}
    for mode in args.modes.split(","):
        if mode == "cold":
            init = W.starts(reaches, "cold")
        elif mode == "plane":
            init = W.starts(reaches, "plane")
        elif mode == "gvf":
            init = W.starts(reaches, "gvf")
        else:
            ws = W.WarmStart.load(ckpt)
            init = ws.predict(reaches)
            ws.close()
        o = dict(W.STEADY_OPTS)
        o.update(rstop=0.0, nonfinite_stop=True, istop=args.max_steps)
        b = fvt.BatchSolver(reaches, o)
        if init is not None:
            b.set_state(np.concatenate([x[0] for x in init]), np.concatenate([x[1] for x in init]),
                        np.concatenate([x[2] for x in init]))
        first = {tol: [None] * len(reaches) for tol in tols}
        at_gpu = {tol: None for tol in tols}
        gpu_s, it = 0.0, 0
        qh, qu, qv = np.zeros(b.N), np.zeros(b.N), np.zeros(b.N)
        dp = C.POINTER(C.c_double)
        while it < args.max_steps:
            b.run(500)
            gpu_s += b.last_run_seconds()
            it = int(b.status()[0].max())
            b._chk(fvt.load().storm_fvt_get_conserved(b.h, qh.ctypes.data_as(dp),
                                                       qu.ctypes.data_as(dp),
                                                       qv.ctypes.data_as(dp)))
            for r in range(len(reaches)):
                e = float(np.sqrt(np.mean((qh[b.cell_ptr[r]:b.cell_ptr[r + 1]] - href[r]) ** 2)))
                for tol in tols:
                    if first[tol][r] is None and e <= tol:
                        first[tol][r] = it
            for tol in tols:
                if at_gpu[tol] is None and all(x is not None for x in first[tol]):
                    at_gpu[tol] = gpu_s
            print("%s: step %d, reaches within 1 mm: %d" %
                  (mode, it, sum(x is not None for x in first[1e-3])), file=sys.stderr, flush=True)
            if at_gpu[tols[-1]] is not None:
                break
        b.close()
        res[mode] = {("%g m" % tol): dict(gpu_s_all=at_gpu[tol], steps=first[tol]) for tol in tols}
        print("%s: %s" % (mode, json.dumps(res[mode])), file=sys.stderr, flush=True)
    out = {
        "metric": "GPU time until every reach is within 1 cm / 1 mm RMS of its converged h",
        "unit": "s", "n_gpus": 1, "higher_is_better": False, "reaches": len(reaches),
        "triangles_per_reach": args.size, "reference_s": ref_s,
        "data": "synthetic (storm_amd/warmstart.py heldout_reach, seeds 9009+k)",
        "config": {"workload": "warmstart-accuracy-%dx%d" % (len(reaches), args.size),
                   "reference": "RMS dh <= 1e-9 from the backwater start",
                   "checkpoint": os.path.relpath(ckpt, ROOT)},
        "results": res,
    }
    print(json.dumps(out))


def warmstart(args):
    """Wall time and steps to steady state (RMS dh <= rstop) of held-out
    reaches from three starts, each run as one batched context: cold (still
    water covering every cell, warmstart.cold_state), the plane guess
    (warmstart.baseline_state, no network), and the learned warm start
    (features + network inference + upload inside the timed region)."""
    from storm_amd import warmstart as W
    n_r = args.reaches if args.reaches != 256 else 64
    ckpt = args.checkpoint or W.DEFAULT_CHECKPOINT
    t0 = time.perf_counter()
    reaches = [W.heldout_reach(k, args.size) for k in range(n_r)]
    gen_s = time.perf_counter() - t0
    opts = dict(rstop=args.rstop, nonfinite_stop=True)
    res, finals = {}, {}
    modes = args.modes.split(",")
    for mode in modes:
        t0 = time.perf_counter()
        init, pred_s = None, 0.0
        if mode == "cold":
            init = W.starts(reaches, "cold")
        elif mode == "plane":
            init = W.starts(reaches, "plane")
        elif mode == "gvf":
            init = W.starts(reaches, "gvf")
        elif mode in ("learned", "learned_discharge"):
            # learned: the network's water surface with uniform discharge
            # along the axis (WarmStart.predict's default); learned_discharge:
            # its discharge outputs too (warmstart.decode)
            ws = W.WarmStart.load(ckpt)
            init = ws.predict(reaches,
                              discharge="learned" if mode == "learned_discharge" else "uniform")
            ws.close()
        pred_s = time.perf_counter() - t0
        out, _ = W.steady_states(reaches, opts, args.max_steps, chunk=1000, init=init,
                                 log=lambda m: print("%s: %s" % (mode, m), file=sys.stderr,
                                                     flush=True))
        wall = time.perf_counter() - t0
        steps = np.array([it for _, it in out])
        ok = np.array([st is not None for st, _ in out])
        finals[mode] = out
        res[mode] = dict(wall_s=wall, init_s=pred_s, converged=int(ok.sum()),
                         steps_median=float(np.median(steps)), steps_max=int(steps.max()),
                         steps_sum=int(steps.sum()), steps=steps.tolist(),
                         cell_updates=float(sum(c.n_elems * it for c, it in zip(reaches, steps))))
        if mode != "cold":
            # the start's error against the state the solver converged to
            eh, eq = [], []
            for (st, _), (h, u, v) in zip(out, init):
                if st is None:
                    continue
                eh.append(np.sqrt(np.mean((h - st["h"]) ** 2)) / np.mean(st["h"]))
                eq.append(np.sqrt(np.mean((h * u - st["qu"]) ** 2)) / np.mean(np.abs(st["qu"])))
            res[mode]["pred_err_h_rms_rel"] = float(np.median(eh)) if eh else None
            res[mode]["pred_err_hu_rms_rel"] = float(np.median(eq)) if eq else None
        print("%s: %s" % (mode, json.dumps(res[mode])), file=sys.stderr, flush=True)
    # how far apart the converged states of the warm and cold starts are
    dd = []
    if "learned" in finals and "cold" in finals:
        for (a, _), (b, _) in zip(finals["learned"], finals["cold"]):
            if a is not None and b is not None:
                dd.append(np.sqrt(np.mean((a["h"] - b["h"]) ** 2)))
    wall = lambda m: res[m]["wall_s"] if m in res else None
    ratio = wall("cold") / wall("learned") if wall("cold") and wall("learned") else None
    out = {
        "metric": "time-to-steady, cold / learned warm start (wall-time ratio)",
        "value": ratio, "unit": "x", "n_gpus": 1, "higher_is_better": True,
        "dtype": "f64 solver, bf16 network (fp32 accumulate)",
        "data": "synthetic (storm_amd/warmstart.py heldout_reach, seeds 9009+k)",
        "config": {"workload": "warmstart-%dx%d" % (n_r, args.size), "rstop": args.rstop,
                   "scheme": "HLL first order, SSP-RK2, steady local dt CFL 0.9, stop RMS dh",
                   "checkpoint": os.path.relpath(ckpt, ROOT), "generate_s": round(gen_s, 1)},
        "cold": res.get("cold"), "plane": res.get("plane"), "gvf": res.get("gvf"),
        "learned": res.get("learned"),
        "learned_discharge": res.get("learned_discharge"),
        "plane_over_cold": wall("cold") / wall("plane") if wall("cold") and wall("plane") else None,
        "h_rms_learned_vs_cold_final_m": float(np.median(dd)) if dd else None,
    }
    print(json.dumps(out))


def warmstart_train(args):
    """Training rate of the warm-start network: `--steps` AdamW updates (after
    `--warmup` untimed ones) on batches joining 32 windows of 4,096 cells cut
    from 64 training reaches (131,072 nodes, about 390,000 directed edges per
    batch), with random targets; each update runs the forward pass, the
    backward pass and the optimizer on the device, its batch assembled on a
    host thread meanwhile."""
    from storm_amd import warmstart as W
    import torch
    reaches = [W.training_reach(k) for k in range(64)]
    frs = W.frames(reaches)
    rng = np.random.default_rng(0)
    ys = [(0.1 * rng.standard_normal((fr.node.shape[0], W.F_OUT))).astype(np.float32)
          for fr in frs]
    ws = W.WarmStart(features=W.FEATURES)
    if args.warmup:
        ws.train_windows(frs, ys, args.warmup, seed=1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ws.train_windows(frs, ys, args.steps, seed=2)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ws.close()
    print(json.dumps({
        "metric": "warm-start network AdamW updates/s", "value": args.steps / wall,
        "unit": "updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps, "higher_is_better": True,
        "dtype": "bf16 (fp32 accumulate)",
        "data": "synthetic (storm_amd/warmstart.py training_reach 0-63, random targets)",
        "config": {"workload": "warmstart-train-32x4096", "layers": ws.layers,
                   "width": 64, "features": W.FEATURES}}))


if __name__ == "__main__":
    main()
