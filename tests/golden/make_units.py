"""Writes tests/golden/units.npz: per-routine vectors of the reference.

Runs oracle/_ref/ref_units (single routines of the reference SToRM, built
from its own Fortran sources by oracle/Makefile; this container only) on
seeded inputs and stores inputs and outputs:

* flux_rusanov (flux_rusanov.f90:1-82) on random states, dry sides included,
  with the caller's wave speeds sqrt(g h) (0 on a dry side), as
  invisc_fluxes.f90:143-144 would pass them;
* flux_HLL (flux_HLL.f90:1-112) on states whose two waves run the same way
  (the left or right upwind flux): the routine's star region is defective,
  so only its wave speeds (dry-bed forms included) and upwind fluxes are
  taken from it;
* fric_terms (fric_terms.f90:22-98, Manning) on random cells, some dry;
* fvsources, part_dry and drynwet (the RDS node routines) on two small
  meshes of tests/cases_small.py with random vertex depths, some dry.

Usage (from the repository root):  python tests/golden/make_units.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from storm_amd import cases  # noqa: E402
from cases_small import small_cases  # noqa: E402

UNITS = os.path.join(ROOT, "oracle", "_ref", "ref_units")
G = 9.81
RDS_CASES = ("meander_crit", "dambreak_wetdry")


def run(mode, payload, nout, dtype, wd, run_file=None):
    inp = os.path.join(wd, mode + ".in")
    out = os.path.join(wd, mode + ".out")
    with open(inp, "wb") as f:
        for a in payload:
            f.write(np.ascontiguousarray(a).tobytes())
    cmd = [UNITS, mode, inp, out] + ([run_file] if run_file else [])
    subprocess.run(cmd, cwd=wd, check=True, stdout=subprocess.DEVNULL, timeout=120)
    return np.fromfile(out, dtype=dtype, count=nout)


def states(rng, k, dry_frac=0.2):
    a = rng.uniform(0, 2 * np.pi, k)
    nx, ny = np.cos(a), np.sin(a)
    hl = rng.uniform(0.01, 4.0, k)
    hr = rng.uniform(0.01, 4.0, k)
    hl[rng.random(k) < dry_frac] = 0.0
    hr[rng.random(k) < dry_frac] = 0.0
    ul, vl, ur, vr = (rng.normal(0.0, 2.0, k) for _ in range(4))
    return nx, ny, hl, ul, vl, hr, ur, vr


def hll_speeds(nx, ny, hl, ul, vl, hr, ur, vr, h0):
    # flux_HLL.f90:42-68, to choose states that take an upwind branch
    cl, cr = np.sqrt(G * hl), np.sqrt(G * hr)
    us = 0.5 * ((ul + ur) * nx + (vl + vr) * ny) + cl - cr
    cs = 0.5 * (cl + cr) + ((ul - ur) * nx + (vl - vr) * ny) / 4.0
    unl, unr = ul * nx + vl * ny, ur * nx + vr * ny
    sl = np.where(hl < h0, unr - 2 * cr, np.where(hr < h0, unl - cl, np.minimum(unl - cl, us - cs)))
    sr = np.where(hl < h0, unr + cr, np.where(hr < h0, unl + 2 * cl, np.maximum(unr - cr, us + cs)))
    return sl, sr


def main():
    rng = np.random.default_rng(7007)
    data = {}
    with tempfile.TemporaryDirectory() as wd:
        # flux_rusanov
        k = 400
        nx, ny, hl, ul, vl, hr, ur, vr = states(rng, k)
        cl = np.where(hl > 0, np.sqrt(G * hl), 0.0)
        cr = np.where(hr > 0, np.sqrt(G * hr), 0.0)
        rec = np.stack([nx, ny, hl, ul, vl, cl, hr, ur, vr, cr], axis=1)
        data["rusanov_in"] = rec
        data["rusanov_out"] = run("rusanov", [np.int32(k), rec], 3 * k, np.float64, wd).reshape(k, 3)
        # flux_HLL, upwind branches only
        keep = []
        while sum(len(x) for x in keep) < 400:
            nx, ny, hl, ul, vl, hr, ur, vr = states(rng, 4000, dry_frac=0.3)
            ul, vl, ur, vr = 3 * ul, 3 * vl, 3 * ur, 3 * vr   # supercritical more often
            h0 = np.full(4000, 1e-3)
            sl, sr = hll_speeds(nx, ny, hl, ul, vl, hr, ur, vr, h0)
            both_dry = (hl < h0) & (hr < h0)
            ok = ((sl > 1e-6) | (sr < -1e-6)) & ~((sl > 0) & (sr < 0)) & ~both_dry
            keep.append(np.stack([nx, ny, hl, ul, vl, hr, ur, vr, h0], axis=1)[ok])
        rec = np.concatenate(keep)[:400]
        k = len(rec)
        data["hll_in"] = rec
        data["hll_out"] = run("hll", [np.int32(k), rec], 3 * k, np.float64, wd).reshape(k, 3)
        # fric_terms
        k = 300
        dry = (rng.random(k) < 0.2).astype(np.int32)
        cols = np.stack([rng.uniform(0.02, 0.06, k), rng.uniform(0.005, 5.0, k),
                         rng.uniform(0.0, 3.0, k), rng.normal(0, 1.5, k), rng.normal(0, 1.5, k),
                         rng.uniform(0.5, 50.0, k), rng.normal(0, 10.0, k),
                         rng.normal(0, 10.0, k)], axis=1)
        data["fric_dry"] = dry
        data["fric_in"] = cols
        out = run("fric", [np.int32(k), dry, cols.T.copy()], 2 * k,
                  np.float64, wd)
        data["fric_out"] = out.reshape(2, k)
        # the RDS node routines on two meshes
        for name in RDS_CASES:
            case, _ = small_cases()[name]
            runf = cases.write_run(case, os.path.join(wd, name))
            P, N = case.n_pts, case.n_elems
            h_dry = float(case.opts["h_dry"])
            hv = rng.uniform(0.0, 1.5, P)
            dryv = rng.random(P) < 0.3
            hv[dryv] = rng.uniform(0.0, 0.9 * h_dry, int(dryv.sum()))
            out = run("rds", [np.float64(h_dry), case.zvtx, hv], 3 * P + 3 * N, np.float64,
                      os.path.join(wd, name), run_file=os.path.basename(runf))
            with open(os.path.join(wd, name, "rds.out"), "rb") as f:
                f.seek(8 * (3 * P + 3 * N))
                pd = np.frombuffer(f.read(4 * 4 * N), dtype=np.int32).reshape(N, 4)
            data[f"rds_{name}_h"] = hv
            data[f"rds_{name}_source"] = out[:3 * P].reshape(3, P)
            data[f"rds_{name}_zout"] = out[3 * P:].reshape(N, 3)
            data[f"rds_{name}_pd"] = pd
    data["meta"] = np.array(repr({"seed": 7007, "rds_cases": RDS_CASES,
                                  "harness": "oracle/ref_units.f90"}))
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "units.npz"),
                        **data)
    print("units.npz:", ", ".join(sorted(k for k in data if not k.endswith("_in"))))


if __name__ == "__main__":
    main()
