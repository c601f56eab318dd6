"""Train the warm-start network on steady states the batched FP64 solver
generates.

    python -m storm_amd.train_warmstart --reaches 512 --out <path.npz>

1. Training reaches k = 0..R-1 come from the ensemble generator at 20k-100k
   triangles (seeds 5005+k).  The batched solver runs them together, from
   the 1-D backwater profile of storm_amd.warmstart.gvf_state (or, with
   --start plane, the plane guess of baseline_state), to RMS dh <= rstop
   (default 1e-8: at 1e-6 the RMS-change test still leaves centimetre errors
   in h).  A reach that diverges or does not converge is left out.
2. Features and targets (storm_amd.warmstart.frame / targets); a held-out
   validation share.
3. AdamW on batches that join windows of a few thousand cells cut from many
   different reaches (the prediction depends mostly on each reach's slope,
   roughness, discharge and stage, so a batch must span many reaches), all
   on the GPU in the network's own forward, backward and AdamW kernels.
4. The checkpoint (FP32 parameters + a JSON record of the data and the loss
   history) is written with WarmStart.save.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import warmstart as W


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def generate(n_reaches: int, gen_batch: int, rstop: float, max_steps: int,
             gen_time: float = 0.0, first: int = 0, start: str = "gvf"):
    """Steady states of training reaches first..first+n-1 (None where not
    converged).  gen_time > 0: no further batch is started once that many
    seconds have passed, so fewer than n reaches may come back."""
    states, steps, walls = [], [], []
    t0 = time.perf_counter()
    for s in range(first, first + n_reaches, gen_batch):
        if gen_time > 0 and time.perf_counter() - t0 > gen_time:
            log("generation time limit reached after %d reaches" % (s - first))
            break
        ks = list(range(s, min(first + n_reaches, s + gen_batch)))
        rs = [W.training_reach(k) for k in ks]
        init = [W.gvf_state(c) if start == "gvf" else W.baseline_state(c) for c in rs]
        out, wall = W.steady_states(rs, dict(rstop=rstop, nonfinite_stop=True), max_steps,
                                    chunk=5000, init=init, log=log)
        walls.append(wall)
        for (st, it) in out:
            states.append(st)
            steps.append(it)
        log("generated reaches %d-%d: %d converged, %.1f s" %
            (ks[0], ks[-1], sum(o[0] is not None for o in out), wall))
    return states, steps, walls


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--reaches", type=int, default=512)
    ap.add_argument("--gen-batch", type=int, default=128)
    ap.add_argument("--rstop", type=float, default=1e-8)
    ap.add_argument("--max-steps", type=int, default=150000)
    ap.add_argument("--gen-time", type=float, default=0.0,
                    help="start no new generation batch after this many seconds (0: no limit)")
    ap.add_argument("--train-time", type=float, default=0.0,
                    help="fit the AdamW updates into this many seconds (0: no limit)")
    ap.add_argument("--steps", type=int, default=4000, help="AdamW updates")
    ap.add_argument("--reaches-per-step", type=int, default=32)
    ap.add_argument("--window", type=int, default=4096, help="cells cut from each reach")
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--lr-final", type=float, default=1e-5, help="end of the cosine decay")
    ap.add_argument("--val", type=float, default=0.05, help="validation share")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--features", default=str(W.FEATURES),
                    help="node feature set(s), comma-separated: one network per set, "
                         "trained on the same data (out gets a _f<set> suffix when several)")
    ap.add_argument("--first", type=int, default=0,
                    help="first training reach (a later phase takes reaches no earlier one saw)")
    ap.add_argument("--init", default="",
                    help="continue training this checkpoint (a later phase on new reaches)")
    ap.add_argument("--start", choices=["gvf", "plane"], default="gvf",
                    help="the solver's start in generation: the 1-D backwater profile "
                         "(warmstart.gvf_state) or the plane (baseline_state)")
    ap.add_argument("--out", required=True)
    a = ap.parse_args(argv)

    t0 = time.perf_counter()
    states, steps, walls = generate(a.reaches, a.gen_batch, a.rstop, a.max_steps, a.gen_time,
                                    a.first, a.start)
    gen_s = time.perf_counter() - t0
    # keys: reach numbers of the converged steady states
    ks = [a.first + j for j, st in enumerate(states) if st is not None]
    sets = [int(f) for f in a.features.split(",")]
    prev = None
    if a.init:
        prev = W.WarmStart.load(a.init)
        sets = [prev.features]
    for fs in sets:
        frames, ys = [], []
        for k in ks:
            c = W.training_reach(k)
            fr = W.frame(c, fs)
            st = states[k - a.first]
            frames.append(fr)
            ys.append(W.targets(fr, st["qh"], st["qu"], st["qv"]))
        n_val = max(1, int(round(a.val * len(ks))))
        tr_f, tr_y = frames[:-n_val], ys[:-n_val]
        va_f, va_y = frames[-n_val:], ys[-n_val:]
        log("feature set %d: training on %d reaches, validating on %d" % (fs, len(tr_f), len(va_f)))
        if prev is not None:
            ws = W.WarmStart(layers=prev.layers, lr=a.lr, seed=a.seed, params=prev.net.params(),
                             features=fs)
            ws.history = list(prev.history)
            prev.close()
        else:
            ws = W.WarmStart(layers=6, lr=a.lr, seed=a.seed, features=fs)
        val0 = ws.loss(va_f, va_y)
        t1 = time.perf_counter()
        ws.train_windows(tr_f, tr_y, a.steps, reaches_per_step=a.reaches_per_step,
                         size=a.window, seed=a.seed, log=log, lr=a.lr, lr_final=a.lr_final,
                         time_limit=a.train_time or None)
        train_s = time.perf_counter() - t1
        val1 = ws.loss(va_f, va_y)
        zero = float(np.mean([float((y ** 2).mean()) for y in va_y]))
        meta = dict(
            first=a.first, reaches=len(states), converged=len(ks), rstop=a.rstop,
            solver=W.STEADY_OPTS,
            train_reaches=len(tr_f), val_reaches=len(va_f), val_keys=ks[-n_val:],
            steps_median=float(np.median([steps[k - a.first] for k in ks])),
            generate_s=gen_s, train_s=train_s, steps=ws.history[-1]["step"],
            reaches_per_step=a.reaches_per_step, window=a.window, lr=a.lr,
            lr_final=a.lr_final, val_loss_initial=val0, val_loss=val1,
            val_loss_of_zero_output=zero,
            generator="storm_amd.warmstart.training_reach (seeds 5005+k, 20k-100k triangles)",
            start=a.start)
        # a later phase keeps the record of the earlier ones
        phases = []
        if a.init:
            pm = W.WarmStart.load(a.init)
            old = dict(pm.meta)
            pm.close()
            phases = old.pop("phases", None) or [{k: v for k, v in old.items()
                                                  if k not in ("history", "val_keys")}]
        phases.append({k: v for k, v in meta.items() if k != "val_keys"})
        meta["phases"] = phases
        meta["steady_states_total"] = int(sum(ph["converged"] for ph in phases))
        out = a.out if len(sets) == 1 else a.out.replace(".npz", "_f%d.npz" % fs)
        ws.save(out, meta)
        log("saved %s: val loss %.5f (untrained %.5f, zero output %.5f)" % (out, val1, val0, zero))
        print(json.dumps({k: v for k, v in meta.items() if k not in ("val_keys",)}))
        ws.close()


if __name__ == "__main__":
    main()
