"""The learned warm start on the GPU: the trained network (the committed
checkpoint, trained by storm_amd.train_warmstart on steady states of the
batched solver) predicts held-out reaches better than both non-learned
starts, the plane and the 1-D backwater profile it refines (what the network
gains over the method without it), and the FP64 solver started from its
prediction reaches the steady criterion in fewer steps than from a cold
start, at a state as close to the tightly converged one."""
import os

import numpy as np
import pytest

from storm_amd import warmstart as W

pytestmark = pytest.mark.gpu

N_REACH, SIZE = 4, 200000  # the held-out evaluation's reach size


def _reaches():
    return [W.heldout_reach(k, SIZE) for k in range(N_REACH)]


@pytest.fixture(scope="module")
def net():
    if not os.path.exists(W.DEFAULT_CHECKPOINT):
        pytest.fail("trained checkpoint missing: %s" % W.DEFAULT_CHECKPOINT)
    ws = W.WarmStart.load(W.DEFAULT_CHECKPOINT)
    yield ws
    ws.close()


@pytest.fixture(scope="module")
def steady():
    """Tightly converged states (RMS dh <= 1e-9) of the held-out reaches."""
    rs = _reaches()
    print("[gpu warmstart] reference steady states, %d x %d" % (N_REACH, SIZE), flush=True)
    out, _ = W.steady_states(rs, dict(rstop=1e-9, nonfinite_stop=True), 300000, chunk=5000,
                             init=[W.gvf_state(c) for c in rs])
    assert all(st is not None for st, _ in out)
    return rs, [st for st, _ in out]


def _rms(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)))


def test_prediction_is_closer_to_steady_than_the_non_learned_starts(net, steady):
    rs, ref = steady
    pred = net.predict(rs)
    plane = [W.baseline_state(c) for c in rs]
    gvf = [W.gvf_state(c) for c in rs]
    e_net = [_rms(p[0], st["h"]) for p, st in zip(pred, ref)]
    e_plane = [_rms(p[0], st["h"]) for p, st in zip(plane, ref)]
    e_gvf = [_rms(p[0], st["h"]) for p, st in zip(gvf, ref)]
    print("[gpu warmstart] RMS h error, network %s, backwater %s, plane %s"
          % (e_net, e_gvf, e_plane), flush=True)
    assert np.median(e_net) < 0.5 * np.median(e_plane)
    assert np.median(e_net) < 0.8 * np.median(e_gvf)
    # discharge along the reach: within 2% of the steady one
    for (h, u, _), st in zip(pred, ref):
        assert _rms(h * u, st["qu"]) < 0.02 * np.mean(np.abs(st["qu"]))


def test_warm_start_needs_fewer_steps_than_cold(net, steady):
    rs, ref = steady
    opts = dict(rstop=1e-6, nonfinite_stop=True)
    cold, _ = W.steady_states(rs, opts, 200000, chunk=5000, init=[W.cold_state(c) for c in rs])
    warm, _ = W.steady_states(rs, opts, 200000, chunk=5000, init=net.predict(rs))
    s_cold = np.array([it for _, it in cold])
    s_warm = np.array([it for _, it in warm])
    print("[gpu warmstart] steps cold %s warm %s" % (s_cold.tolist(), s_warm.tolist()), flush=True)
    assert all(st is not None for st, _ in warm)
    assert s_warm.sum() * 2 < s_cold.sum()
    # the warm run stops at least as close to the converged state as the cold one
    for (a, _), (b, _), st in zip(warm, cold, ref):
        if b is not None:
            assert _rms(a["h"], st["h"]) <= 1.5 * _rms(b["h"], st["h"]) + 1e-3


def test_outputs_do_not_depend_on_the_batch(net):
    """A forward pass over several reaches (their disjoint union) gives each
    reach the bits it gets alone."""
    rs = [W.heldout_reach(k, 20000) for k in range(3)]
    frs = W.frames(rs, net.features)
    print("[gpu warmstart] batch independence, 3 reaches", flush=True)
    alone = net.predict_outputs(frs, batch=1)
    together = net.predict_outputs(frs, batch=3)
    for a, b in zip(alone, together):
        assert np.array_equal(a, b)
