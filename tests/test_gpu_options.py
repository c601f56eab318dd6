"""HIP path vs the CPU oracle for the options whose reference routine is
defective or off the reference's active path (corrected HLL and BGNVC,
Rusanov, Green-Gauss gradients, explicit friction, the corrected CFL time
step, stopping test 2).  The oracle's forms of these are pinned by the closed-form tests in
test_oracle_closed_forms.py."""
import numpy as np
import pytest

from cases_small import option_cases

pytestmark = pytest.mark.gpu

RTOL = 1e-10


def _cmp(a, b):
    scale = max(np.abs(a).max(), 1e-300)
    return np.abs(a - b).max() / scale


@pytest.mark.parametrize("name", list(option_cases().keys()))
def test_option_parity(name):
    import oracle as O
    from storm_amd import fvt
    case, n = option_cases()[name]
    print(f"[gpu options] {name} N={case.n_elems} steps={n}", flush=True)
    orc = O.Oracle(case)
    r_o = orc.run(n)
    orc.finalize()
    so = orc.state()
    s = fvt.FVTSolver(case)
    it, tt, rm = s.run(n)
    s.finalize()
    sg = s.state()
    assert it == r_o["iter"]
    assert tt == pytest.approx(r_o["ttime"], rel=1e-12)
    assert rm == pytest.approx(r_o["rmax"], rel=1e-8)
    for k in ("h", "qh", "qu", "qv", "u", "v"):
        assert _cmp(so[k], sg[k]) <= RTOL, (k, _cmp(so[k], sg[k]))
    assert np.array_equal(so["dry"], sg["dry"])
