"""HIP path vs the CPU oracle on the same seeded cases (full time loop)."""
import numpy as np
import pytest

from cases_small import small_cases

pytestmark = pytest.mark.gpu

# relative L-infinity bound on h, hu, hv after N steps (north-star bound 1e-10)
RTOL = 1e-10


def _cmp(a, b):
    scale = max(np.abs(a).max(), 1e-300)
    return np.abs(a - b).max() / scale


@pytest.mark.parametrize("name", list(small_cases().keys()))
def test_step_parity(name):
    import oracle as O
    from storm_amd import fvt
    case, n = small_cases()[name]
    print(f"[gpu parity] {name} N={case.n_elems} steps={n}", flush=True)
    orc = O.Oracle(case)
    r_o = orc.run(n)
    orc.finalize()
    so = orc.state()
    s = fvt.FVTSolver(case)
    it, tt, rm = s.run(n)
    s.finalize()
    sg = s.state()
    assert it == r_o["iter"]
    assert tt == pytest.approx(r_o["ttime"], rel=1e-13)
    for k in ("h", "qh", "qu", "qv", "u", "v"):
        assert _cmp(so[k], sg[k]) <= RTOL, (k, _cmp(so[k], sg[k]))
    for k in ("hvtx", "zetavtx", "phi1", "phi2", "phi3"):
        assert _cmp(so[k], sg[k]) <= 1e-9, (k, _cmp(so[k], sg[k]))
    assert np.array_equal(so["dry"], sg["dry"])
