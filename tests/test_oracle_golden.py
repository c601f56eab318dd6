"""The C oracle against the reference's own outputs (golden vectors written
by tests/golden/make_golden.py from the flang-built reference)."""
import os

import numpy as np
import pytest

from cases_small import small_cases

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
KEYS = ("h", "u", "v", "qh", "qu", "qv", "phi1", "phi2", "phi3", "dry", "hvtx", "zetavtx")


@pytest.mark.parametrize("name", list(small_cases().keys()))
def test_oracle_matches_reference_bitwise(name):
    import oracle as O
    case, n = small_cases()[name]
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
