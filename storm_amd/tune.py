"""Autotuner of the FVT step's launch configuration.

The tunables are the stage kernel's tile capacity (`block`: threads per
block, which sets the tile size and the halo share) and whole-step HIP graph
replay (`graph`).  `tune` times every candidate on the caller's own mesh or
ensemble for a few steps, keeps the fastest, and caches the choice per GPU
architecture and workload shape in a JSON file ($STORM_TUNE_CACHE, default
~/.cache/storm_amd/tune.json).  The tunables change only the launch shape:
results are bit-identical whichever is chosen.
"""
from __future__ import annotations

import json
import math
import os
from typing import Dict, List, Optional, Sequence

from . import fvt

CANDIDATES: List[Dict[str, int]] = [dict(block=b, graph=g) for b in (256, 384, 512) for g in (0, 1)]


def _cache_path() -> str:
    return os.environ.get("STORM_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "storm_amd", "tune.json"))


def _arch() -> str:
    try:
        import torch
        if torch.cuda.is_available():
            return str(torch.cuda.get_device_properties(0).gcnArchName).split(":")[0]
    except Exception:
        pass
    return "unknown"


def signature(cases_: Sequence, opts: Dict[str, object]) -> str:
    """Workload shape: reach count, total cells (to a factor of 2) and the
    options that change the kernels' work."""
    n = sum(c.n_elems for c in cases_)
    return "r%d-n2^%d-rk%s-%s-%s-f%s-l%s-s%s" % (
        len(cases_), int(round(math.log2(max(n, 1)))), opts.get("rkorder"), opts.get("flux"),
        opts.get("gradient"), opts.get("friction"), opts.get("limiter"),
        int(opts.get("rsavg_n", 0)) if opts.get("rsavg") else 0)


def _make(cases_, opts):
    if len(cases_) == 1:
        return fvt.FVTSolver(cases_[0], opts)
    return fvt.BatchSolver(list(cases_), opts)


def tune(case_or_cases, opts: Optional[Dict[str, object]] = None, steps: int = 6,
         use_cache: bool = True, candidates: Optional[List[Dict[str, int]]] = None):
    """Return (best candidate, {candidate key: seconds per step})."""
    cases_ = list(case_or_cases) if isinstance(case_or_cases, (list, tuple)) else [case_or_cases]
    base = dict(cases_[0].opts)
    if opts:
        base.update(opts)
    key = "%s/%s" % (_arch(), signature(cases_, base))
    path = _cache_path()
    cache = {}
    if use_cache and os.path.exists(path):
        try:
            with open(path) as f:
                cache = json.load(f)
        except (OSError, ValueError):
            cache = {}
    if use_cache and key in cache:
        return dict(cache[key]["best"]), dict(cache[key]["times"])
    times = {}
    best, best_t = None, float("inf")
    if candidates is None:
        candidates = CANDIDATES
        if base.get("rsavg") or str(base.get("fluxes", "std")) == "xsf":
            candidates = [c for c in CANDIDATES if c["block"] == 256]  # their only tile size
    for cand in candidates:
        o = dict(base)
        o.update(cand)
        o["istop"] = 10 ** 9
        o["steady"] = bool(base.get("steady", False))
        s = _make(cases_, o)
        s.run(2)                      # warm-up (first touch, graph capture)
        s.run(steps)
        t = s.last_run_seconds() / steps
        s.close()
        times["block=%d,graph=%d" % (cand["block"], cand["graph"])] = t
        if t < best_t:
            best, best_t = dict(cand), t
    if use_cache:
        cache[key] = dict(best=best, times=times)
        os.makedirs(os.path.dirname(path), exist_ok=True)
        with open(path, "w") as f:
            json.dump(cache, f, indent=1)
    return best, times
