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
