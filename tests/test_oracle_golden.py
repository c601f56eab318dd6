"""The C oracle against the reference's own outputs (golden vectors written
by tests/golden/make_golden.py from the flang-built reference)."""
import os

import numpy as np
import pytest

from cases_small import golden_cases

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
KEYS = ("h", "u", "v", "qh", "qu", "qv", "phi1", "phi2", "phi3", "dry", "hvtx", "zetavtx")
# the post-loop fields of solverFVT.f90:483-487 (to_vertices with enforce_bcs,
# ctr2vtx of cd, error_est(2) on the last stage's delh)
OUT_KEYS = ("uvtx", "vvtx", "cdvtx", "e_est", "delhx", "delhy")


@pytest.mark.parametrize("name", list(golden_cases().keys()))
def test_oracle_matches_reference_bitwise(name):
    import oracle as O
    case, n = golden_cases()[name]
    g = np.load(os.path.join(GOLDEN, f"{name}.npz"))
    o = O.Oracle(case)
    r = o.run(n)
    o.finalize()
    s = o.state()
    assert r["iter"] == int(g["iter"])
    assert r["ttime"] == float(g["ttime"])
    for k in KEYS:
        a = g[k]
        b = s[k]
        # bit-for-bit (signed zeros compare equal)
        assert np.array_equal(a, b.astype(a.dtype)), (k, np.abs(a - b).max())
    out = o.output()
    for k in OUT_KEYS:
        assert np.array_equal(g[k], out[k]), (k, np.abs(g[k] - out[k]).max())


@pytest.mark.parametrize("name", list(golden_cases().keys()))
def test_oracle_dataprep_copies_match_reference_bitwise(name):
    """The module copies dataprepFVT.f90:69-140 leaves (ctr2vtx type 3 and
    vtx2ctr of the inputs given at the other location), before any step."""
    import oracle as O
    from storm_amd import cases
    case, _ = golden_cases()[name]
    g = np.load(os.path.join(GOLDEN, f"{name}.npz"))
    o = O.Oracle(case)
    p = o.prep()
    s = o.state()
    for k, _ in cases.PREP_KEYS:
        b = p[k] if k in p else s[k]
        assert np.array_equal(g["prep_" + k], b), (k, np.abs(g["prep_" + k] - b).max())
