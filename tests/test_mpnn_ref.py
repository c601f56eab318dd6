"""The FP64 restatement of the warm-start network: its gradients equal
central finite differences of its loss."""
import numpy as np

from mpnn_ref import Ref, toy_graph


def test_reference_gradients_match_finite_differences():
    ref = Ref(layers=2)
    rng = np.random.default_rng(3)
    p = rng.normal(scale=0.2, size=ref.n)
    src, dst, X, EF, T = toy_graph(n=12, seed=1)
    w = rng.uniform(0.5, 1.5, size=12)
    g = ref.grads(p, src, dst, X, EF, T, w)
    idx = rng.choice(ref.n, size=60, replace=False)
    for i in idx:
        e = np.zeros(ref.n)
        e[i] = 1e-6
        fd = (ref.loss(p + e, src, dst, X, EF, T, w) - ref.loss(p - e, src, dst, X, EF, T, w)) / 2e-6
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_bf16_rounding_is_nearest_even():
    from mpnn_ref import bf16
    eps = 2.0 ** -7  # BF16 spacing at 1
    x = np.array([1.0, 1 + eps / 2, 1 + 3 * eps / 2, 1 + eps / 2 + 2 ** -20, -1 - eps / 4,
                  3.140625, 1e-40, np.inf])
    want = np.array([1.0, 1.0, 1 + 2 * eps, 1 + eps, -1.0, 3.140625, bf16(1e-40)[()], np.inf])
    assert np.array_equal(bf16(x), want)
    # every BF16 value is an FP32 value whose low 16 bits are zero
    r = bf16(np.random.default_rng(0).normal(size=1000)).astype(np.float32)
    assert not (r.view(np.uint32) & 0xFFFF).any()


def test_bf16_reference_tracks_fp64_to_bf16_precision():
    rng = np.random.default_rng(4)
    src, dst, X, EF, T = toy_graph(n=50, seed=3)
    r64, r16 = Ref(layers=2), Ref(layers=2, bf16=True)
    p = rng.normal(scale=0.1, size=r64.n)
    y64, _ = r64.forward(p, src, dst, X, EF)
    y16, _ = r16.forward(p, src, dst, X, EF)
    rel = np.linalg.norm(y16 - y64) / np.linalg.norm(y64)
    assert 1e-4 < rel < 3e-2
    g64, g16 = r64.grads(p, src, dst, X, EF, T), r16.grads(p, src, dst, X, EF, T)
    assert np.linalg.norm(g16 - g64) / np.linalg.norm(g64) < 5e-2
