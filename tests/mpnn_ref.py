"""FP64 NumPy restatement of the warm-start network (include/storm_mpnn.h):
the same layers, parameter layout and loss, for checking the GPU forward
pass and gradients.  Test infrastructure only.

With bf16=True every product takes its two operands rounded to BF16 (round
to nearest even), as the GPU's MFMA GEMMs do, and sums them exactly; every
other operation (bias, ReLU, edge sums, the loss) stays FP64.  That mirrors
the kernels' arithmetic up to their FP32 accumulation and storage."""
import numpy as np

H = 64


def bf16(a):
    """Round to the nearest BF16 value (ties to even), returned as FP64."""
    x = np.asarray(a, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    r = np.where(np.isfinite(x), r, x.view(np.uint32))
    return r.view(np.float32).astype(np.float64)


def layout(layers, f_node, f_edge, f_out):
    off, o = {}, 0

    def take(name, n):
        nonlocal o
        off[name] = (o, n)
        o += n
    take("enc_w", f_node * H)
    take("enc_b", H)
    for l in range(layers):
        take(f"e_w{l}", (2 * H + f_edge) * H)
        take(f"e_b{l}", H)
        take(f"n_w{l}", 2 * H * H)
        take(f"n_b{l}", H)
    take("dec_w", H * f_out)
    take("dec_b", f_out)
    return off, o


class Ref:
    def __init__(self, layers=6, f_node=16, f_edge=16, f_out=16, bf16=False):
        self.L, self.fn, self.fe, self.fo = layers, f_node, f_edge, f_out
        self.off, self.n = layout(layers, f_node, f_edge, f_out)
        self.r = bf16 if not bf16 else globals()["bf16"]

    def mm(self, a, b):
        """A product as the network forms it (operands rounded if bf16)."""
        if self.r is False:
            return a @ b
        return self.r(a) @ self.r(b)

    def _get(self, p, name, shape):
        o, n = self.off[name]
        return p[o:o + n].reshape(shape)

    def forward(self, p, src, dst, X, EF):
        p = np.asarray(p, dtype=np.float64)
        c = {"X": X.astype(np.float64), "EF": EF.astype(np.float64)}
        N = X.shape[0]
        h = np.maximum(self.mm(c["X"], self._get(p, "enc_w", (self.fn, H))) + self._get(p, "enc_b", (H,)), 0)
        c["h0"] = h
        for l in range(self.L):
            xe = np.concatenate([h[src], h[dst], c["EF"]], axis=1)
            m = np.maximum(self.mm(xe, self._get(p, f"e_w{l}", (2 * H + self.fe, H))) +
                           self._get(p, f"e_b{l}", (H,)), 0)
            a = np.zeros((N, H))
            np.add.at(a, dst, m)
            xn = np.concatenate([h, a], axis=1)
            u = np.maximum(self.mm(xn, self._get(p, f"n_w{l}", (2 * H, H))) + self._get(p, f"n_b{l}", (H,)), 0)
            c[f"xe{l}"], c[f"m{l}"], c[f"xn{l}"], c[f"u{l}"], c[f"hin{l}"] = xe, m, xn, u, h
            h = h + u
        c["hL"] = h
        y = self.mm(h, self._get(p, "dec_w", (H, self.fo))) + self._get(p, "dec_b", (self.fo,))
        return y, c

    def loss(self, p, src, dst, X, EF, T, w=None):
        y, _ = self.forward(p, src, dst, X, EF)
        w = np.ones(X.shape[0]) if w is None else w
        return float((w[:, None] * (y - T) ** 2).sum() / y.size)

    def grads(self, p, src, dst, X, EF, T, w=None):
        p = np.asarray(p, dtype=np.float64)
        y, c = self.forward(p, src, dst, X, EF)
        N = X.shape[0]
        w = np.ones(N) if w is None else w
        g = np.zeros_like(p)

        def put(name, val):
            o, n = self.off[name]
            g[o:o + n] = val.reshape(-1)
        dy = 2 * w[:, None] * (y - T) / y.size
        put("dec_w", self.mm(c["hL"].T, dy))
        put("dec_b", dy.sum(0))
        dh = self.mm(dy, self._get(p, "dec_w", (H, self.fo)).T)
        for l in reversed(range(self.L)):
            du = dh * (c[f"u{l}"] > 0)
            put(f"n_w{l}", self.mm(c[f"xn{l}"].T, du))
            put(f"n_b{l}", du.sum(0))
            dxn = self.mm(du, self._get(p, f"n_w{l}", (2 * H, H)).T)
            dm = dxn[:, H:][dst] * (c[f"m{l}"] > 0)
            put(f"e_w{l}", self.mm(c[f"xe{l}"].T, dm))
            put(f"e_b{l}", dm.sum(0))
            dxe = self.mm(dm, self._get(p, f"e_w{l}", (2 * H + self.fe, H)).T)
            dh_new = dh + dxn[:, :H]
            np.add.at(dh_new, src, dxe[:, :H])
            np.add.at(dh_new, dst, dxe[:, H:2 * H])
            dh = dh_new
        dz = dh * (c["h0"] > 0)
        put("enc_w", self.mm(c["X"].T, dz))
        put("enc_b", dz.sum(0))
        return g


def toy_graph(n=40, seed=0, f_node=16, f_edge=16, f_out=16):
    """A small random graph with both directions of every edge."""
    rng = np.random.default_rng(seed)
    pairs = set()
    for i in range(1, n):
        pairs.add((int(rng.integers(0, i)), i))
    for _ in range(n):
        a, b = rng.integers(0, n, 2)
        if a != b:
            pairs.add((int(min(a, b)), int(max(a, b))))
    e = np.array(sorted(pairs))
    src = np.concatenate([e[:, 0], e[:, 1]]).astype(np.int32)
    dst = np.concatenate([e[:, 1], e[:, 0]]).astype(np.int32)
    X = rng.normal(size=(n, f_node)).astype(np.float32)
    EF = rng.normal(size=(src.shape[0], f_edge)).astype(np.float32)
    T = rng.normal(size=(n, f_out)).astype(np.float32)
    return src, dst, X, EF, T
