"""Writes the golden reference vectors under tests/golden/.

Runs the reference SToRM FVT path -- built from the reference's own Fortran
sources by oracle/Makefile into oracle/_ref/ref_driver (this container only)
-- on the small seeded cases of tests/cases_small.py and stores its final
state (h, u, v, rk*(:,rkorder), phi, drycell, hvtx, zetavtx, iter, ttime),
the post-loop fields (uvtx, vvtx, cdvtx, zvtx, delh, e_est) and the module
copies dataprepFVT leaves (prep_*) as .npz fixtures.  The fixtures are data: inputs are regenerated from the case
generator and seed recorded with them.

The full-length runs of tests/cases_small.py:large_cases() (final h, hu,
hv, dry mask and vertex depths only) are written when named.

Usage (from the repository root):  python tests/golden/make_golden.py [name ...]
"""
import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import oracle as O  # noqa: E402
from storm_amd import cases  # noqa: E402
from cases_small import golden_cases, large_cases  # noqa: E402

# the full-length runs keep their final state only
LARGE_KEYS = ("h", "qh", "qu", "qv", "dry", "hvtx")
KEYS = ("h", "u", "v", "qh", "qu", "qv", "phi1", "phi2", "phi3", "dry", "hvtx", "zetavtx",
        "uvtx", "vvtx", "cdvtx", "zvtx", "delhx", "delhy", "e_est")


def main():
    out_dir = os.path.dirname(os.path.abspath(__file__))
    only = set(sys.argv[1:])
    large = large_cases()
    todo = dict(golden_cases())
    if only:
        todo.update(large)
    for name, (case, n) in todo.items():
        if only and name not in only:
            continue
        assert int(case.opts["istop"]) == n
        with tempfile.TemporaryDirectory() as wd:
            ref = O.run_reference(case, wd, timeout=3600.0)
        if name in large:
            data = {k: ref[k] for k in LARGE_KEYS}
        else:
            data = {k: ref[k] for k in KEYS}
            data.update({"prep_" + k: ref["prep_" + k] for k, _ in cases.PREP_KEYS})
        data["iter"] = np.int64(ref["iter"])
        data["ttime"] = np.float64(ref["ttime"])
        data["meta"] = np.array(repr(dict(case.meta)))
        np.savez_compressed(os.path.join(out_dir, f"{name}.npz"), **data)
        print(name, case.n_elems, "cells", n, "steps")


if __name__ == "__main__":
    main()
