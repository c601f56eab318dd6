"""The BF16 MFMA network on the GPU against its NumPy restatement (BF16
operand rounding emulated, and plain FP64): forward pass, every parameter
gradient, determinism, and a loss that falls under AdamW."""
import numpy as np
import pytest

from mpnn_ref import Ref, toy_graph

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)


def test_forward_and_gradients_match_the_reference():
    """Against the NumPy restatement with every product's operands rounded to
    BF16 as the kernels round them (the rest differs only by FP32
    accumulation and storage), and against plain FP64 (looser: the BF16
    rounding itself, about 2^-9 per operand, compounds through the layers)."""
    from storm_amd import mpnn
    print("[gpu mpnn] gradients vs FP64 and BF16-rounded references", flush=True)
    src, dst, X, EF, T = toy_graph(n=300, seed=2)
    net = mpnn.Net(layers=3, seed=7)
    ref = Ref(layers=3, bf16=True)
    r64 = Ref(layers=3)
    p = net.params()
    # the decoder starts at zero (an untrained network outputs zero, and every
    # other gradient vanishes): give it weights so every gradient is checked
    o, n = ref.off["dec_w"]
    assert not p[o:o + n].any()
    p[o:o + n] = np.random.default_rng(5).uniform(-0.3, 0.3, n)
    net.set_params(p)
    net.set_graph(src, dst, X, EF)
    y = net.forward()
    yr, _ = ref.forward(p, src, dst, X, EF)
    assert _rel(y, yr) < 1e-3
    assert _rel(y, r64.forward(p, src, dst, X, EF)[0]) < 2e-2
    loss = net.backward(T)
    assert abs(loss - ref.loss(p, src, dst, X, EF, T)) <= 1e-4 * abs(loss)
    g = net.grads()
    gr = ref.grads(p, src, dst, X, EF, T)
    g64 = r64.grads(p, src, dst, X, EF, T)
    # FP32-level differences (accumulation order) flip a few BF16 roundings,
    # and a flip moves an operand by 2^-8; perturbing the emulation's operands
    # by 3e-7 before rounding moves the first layers' gradients by up to 3e-3
    # on this graph, so the bound against the emulation is 1e-2
    for name, (o, n) in ref.off.items():
        err = _rel(g[o:o + n], gr[o:o + n])
        assert err < 1e-2, (name, err)
        err64 = _rel(g[o:o + n], g64[o:o + n])
        assert err64 < 1e-1, (name, err64)
    # bit-repeatable
    net.forward()
    net.backward(T)
    assert np.array_equal(net.grads(), g)
    net.close()


def test_adamw_lowers_the_loss():
    from storm_amd import mpnn
    print("[gpu mpnn] training", flush=True)
    src, dst, X, EF, T = toy_graph(n=200, seed=4)
    T = np.tanh(X[:, :16] * 0.5).astype(np.float32)  # a learnable target
    net = mpnn.Net(layers=2, lr=3e-3, seed=1)
    net.set_graph(src, dst, X, EF)
    first = None
    for it in range(150):
        net.forward()
        loss = net.backward(T)
        first = loss if first is None else first
        net.step()
    assert loss < 0.2 * first, (first, loss)
    net.close()


def test_gradients_of_a_larger_graph():
    """Thousands of nodes and edges: the weight gradients go through many
    row-block partials and both levels of their fixed-order reduction."""
    from storm_amd import mpnn
    print("[gpu mpnn] gradients, 6,000 nodes", flush=True)
    src, dst, X, EF, T = toy_graph(n=6000, seed=3)
    assert src.shape[0] > 32 * 128  # more than one reduction chunk of edge partials
    net = mpnn.Net(layers=2, seed=9)
    ref = Ref(layers=2, bf16=True)
    p = net.params()
    o, n = ref.off["dec_w"]
    p[o:o + n] = np.random.default_rng(6).uniform(-0.3, 0.3, n)
    net.set_params(p)
    net.set_graph(src, dst, X, EF)
    net.forward()
    net.backward(T)
    g = net.grads()
    gr = ref.grads(p, src, dst, X, EF, T)
    for name, (o, n) in ref.off.items():
        err = _rel(g[o:o + n], gr[o:o + n])
        assert err < 1e-2, (name, err)
    net.forward()
    net.backward(T)
    assert np.array_equal(net.grads(), g)
    net.close()
