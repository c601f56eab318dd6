"""The autotuner picks a candidate from its list, caches it per architecture
and workload, and its choice does not change results."""
import numpy as np
import pytest

from storm_amd import cases, fvt, tune

pytestmark = pytest.mark.gpu


def test_tune_picks_and_caches(tmp_path, monkeypatch):
    monkeypatch.setenv("STORM_TUNE_CACHE", str(tmp_path / "tune.json"))
    rs = [cases.ensemble_reach(k, n_min=300, n_max=3000) for k in range(3)]
    opts = dict(dt_op=1, cfl=0.5, steady=False)
    print("[gpu tune] 3 reaches", flush=True)
    best, times = tune.tune(rs, opts, steps=4)
    assert best in tune.CANDIDATES
    assert len(times) == len(tune.CANDIDATES) and all(t > 0 for t in times.values())
    again, times2 = tune.tune(rs, opts, steps=4)
    assert again == best and times2 == times  # served from the cache
    o = dict(opts, istop=20)
    ref = fvt.BatchSolver(rs, dict(o, block=256, graph=0))
    ref.run(20)
    a = ref.state()
    ref.close()
    b = fvt.BatchSolver(rs, dict(o, **best))
    b.run(20)
    bb = b.state()
    b.close()
    for k in ("h", "qh", "qu", "qv", "dry"):
        assert np.array_equal(a[k], bb[k]), k
