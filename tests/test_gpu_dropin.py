"""The drop-in end to end: the reference program's own input path
(read_mfa and the mesh set-up) with dataprepFVT/solverFVT replaced by the
Fortran shim over libstorm_fvt (oracle/_ref/dropin_driver, linked by
`make -C oracle dropin`), run on the .run form of each small case (the wind, culvert and residual
smoothing options among them: the shim reads the wind file and passes
ioculvert's data to the library) and
compared with the reference's golden outputs, the post-loop fields
(vertex velocities, cdvtx, e_est) included."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from cases_small import small_cases, source_cases

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "oracle", "_ref", "dropin_driver")
GOLDEN = os.path.join(ROOT, "tests", "golden")


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="drop-in driver not built")
def _cases():
    out = dict(small_cases())
    out.update(source_cases())
    return out


@pytest.mark.parametrize("name", list(_cases().keys()))
def test_dropin_matches_reference(name):
    from storm_amd import cases
    case, n = _cases()[name]
    g = np.load(os.path.join(GOLDEN, f"{name}.npz"))
    print(f"[dropin] {name} N={case.n_elems}", flush=True)
    with tempfile.TemporaryDirectory() as wd:
        run = cases.write_run(case, wd)
        subprocess.run([DRIVER, os.path.basename(run), "out.bin"], cwd=wd, check=True,
                       stdout=subprocess.DEVNULL, stderr=subprocess.STDOUT, timeout=120)
        out = cases.read_ref_output(os.path.join(wd, "out.bin"))
    assert int(out["iter"]) == int(g["iter"])
    assert out["ttime"] == pytest.approx(float(g["ttime"]), rel=1e-13)
    for k in ("h", "qh", "qu", "qv", "u", "v"):
        a, b = g[k], out[k]
        assert np.abs(a - b).max() <= 1e-10 * max(np.abs(a).max(), 1e-300), k
    for k in ("phi1", "phi2", "phi3", "hvtx", "zetavtx"):
        a, b = g[k], out[k]
        assert np.abs(a - b).max() <= 1e-9 * max(np.abs(a).max(), 1e-300), k
    assert np.array_equal(g["dry"], out["dry"])
    # the post-loop fields the reference's writers read: vertex velocities and
    # drag, the bed at the vertices, delh and the error estimates
    for k in ("uvtx", "vvtx", "cdvtx", "zvtx", "delhx", "delhy", "e_est"):
        a, b = g[k], out[k]
        assert np.abs(a - b).max() <= 1e-9 * max(np.abs(a).max(), 1e-300), k
