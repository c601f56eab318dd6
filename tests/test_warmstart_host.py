"""Host side of the learned warm start: the reach frame, the graph, and the
target/decode pair (no device needed)."""
import numpy as np

from storm_amd import cases, warmstart as W


def _reach():
    return cases.ensemble_reach(3, n_min=2000, n_max=2000, base_seed=77)


def test_graph_is_the_cell_adjacency_both_ways():
    c = _reach()
    fr = W.frame(c)
    inter = c.edge_e[:, 1] >= 0
    assert fr.src.shape[0] == 2 * inter.sum()
    pairs = set(zip(fr.src.tolist(), fr.dst.tolist()))
    assert all((b, a) in pairs for a, b in pairs)
    assert fr.node.shape == (c.n_elems, W.F_NODE) and fr.edge.shape == (fr.src.shape[0], W.F_EDGE)
    assert np.isfinite(fr.node).all() and np.isfinite(fr.edge).all()
    # inflow / outflow flags sit at x = 0 and x = L
    xc = c.x[c.tri].mean(axis=1)
    assert xc[fr.node[:, 8] > 0].max() < 30.0
    assert xc[fr.node[:, 9] > 0].min() > 970.0


def test_frame_of_a_flume():
    c = _reach()
    fr = W.frame(c)
    assert np.allclose(fr.e, [1.0, 0.0], atol=1e-3)
    assert fr.H == c.outflows[0].v[0]
    # q: discharge over the inflow string's length (100 m)
    assert abs(fr.q - c.inflows[0].v[0] / 100.0) < 1e-9 * fr.q
    # the fitted bed slope recovers the generator's slope (noise sd 0.02 m)
    assert abs(fr.S - c.meta["slope"]) < 0.05 * c.meta["slope"] + 2e-5


def test_decode_inverts_targets():
    c = _reach()
    fr = W.frame(c)
    rng = np.random.default_rng(0)
    h = np.maximum(fr.H + fr.S * fr.dist - fr.zc + rng.normal(0, 0.05, c.n_elems), 0.0)
    h[:5] = 0.0  # some dry cells
    qu = fr.q * (1 + rng.normal(0, 0.01, c.n_elems))
    qv = fr.q * rng.normal(0, 0.01, c.n_elems)
    T = W.targets(fr, h, qu, qv).astype(np.float64)
    hd, ud, vd = W.decode(fr, T)
    wet = h > 1e-3
    # float32 targets: 0.1 m * 2^-24 relative on O(1) values
    assert np.abs(hd - h).max() < 1e-6
    assert np.allclose(ud[wet] * hd[wet], qu[wet], rtol=1e-5, atol=1e-6 * fr.q)
    assert np.allclose(vd[wet] * hd[wet], qv[wet], rtol=1e-5, atol=1e-6 * fr.q)
    assert (ud[~wet] == 0).all() and (vd[~wet] == 0).all()


def test_baseline_is_the_plane_with_uniform_discharge():
    c = _reach()
    fr = W.frame(c)
    h, u, v = W.baseline_state(c)
    assert np.allclose(h, np.maximum(fr.H + fr.S * fr.dist - fr.zc, 0.0))
    wet = h > 1e-3
    assert np.allclose(u[wet] * h[wet], fr.q * fr.e[0])
    assert np.allclose(v[wet] * h[wet], fr.q * fr.e[1], atol=1e-12 + 1e-3 * fr.q)


def test_union_offsets_the_graphs():
    frs = [W.frame(cases.ensemble_reach(k, n_min=500, n_max=900, base_seed=5)) for k in range(3)]
    src, dst, X, EF, T = W._union(frs, [np.zeros((f.node.shape[0], 16), np.float32) for f in frs])
    n = [f.node.shape[0] for f in frs]
    assert X.shape[0] == sum(n) and T.shape[0] == sum(n)
    assert src.max() < sum(n) and dst.max() < sum(n)
    o = n[0]
    e0 = frs[0].src.shape[0]
    assert (src[e0:e0 + frs[1].src.shape[0]] >= o).all()


def test_window_cut_and_loss_weights():
    c = cases.ensemble_reach(1, n_min=3000, n_max=3000, base_seed=9)
    fr = W.frame(c)
    y = np.arange(c.n_elems * 16, dtype=np.float32).reshape(-1, 16)
    size, a, hops = 800, 1000, 3
    src, dst, X, EF, T, w = W.window(fr, y, a, size, hops)
    cells = fr.order[a:a + size]
    # the window is a run along the axis (upstream first) with its own edges
    assert np.all(np.diff(fr.dist[fr.order]) <= 0)
    assert np.array_equal(X, fr.node[cells]) and np.array_equal(T, y[cells])
    assert src.min() >= 0 and max(src.max(), dst.max()) < size
    pairs = set(zip(cells[src].tolist(), cells[dst].tolist()))
    full = set(zip(fr.src.tolist(), fr.dst.tolist()))
    inside = set(cells.tolist())
    assert pairs == {(i, j) for i, j in full if i in inside and j in inside}
    assert EF.shape[0] == src.shape[0]
    # in the reach's edge order, each with its own features
    k = np.nonzero(np.isin(fr.src, cells) & np.isin(fr.dst, cells))[0]
    assert np.array_equal(cells[src], fr.src[k]) and np.array_equal(cells[dst], fr.dst[k])
    assert np.array_equal(EF, fr.edge[k])
    # weight zero exactly within `hops` edges of a cell that has an outside
    # neighbour (brute-force breadth-first distance)
    nb = {}
    for i, j in full:
        nb.setdefault(i, []).append(j)
    frontier = {i for i in inside if any(j not in inside for j in nb.get(i, []))}
    dist = {i: 1 for i in frontier}
    cur = frontier
    for d in range(2, hops + 1):
        cur = {j for i in cur for j in nb.get(i, []) if j in inside and j not in dist}
        for j in cur:
            dist[j] = d
    want = np.array([0.0 if int(cid) in dist else 1.0 for cid in cells], dtype=np.float32)
    assert np.array_equal(w, want)
    assert 0 < w.sum() < size
    # a window that takes the whole reach has no cut
    *_, w_all = W.window(fr, y, 0, c.n_elems, hops)
    assert w_all.all()


def test_friction_multiplier_of_identical_cells_is_one():
    """Every cell of an unjittered grid has the last cell's area over
    perimeter, so the steady friction factor is exactly 1; jitter moves it."""
    c = cases.flume(nx=20, ny=6, length=100.0, width=30.0, seed=5)
    c.x = np.round(c.x / 5.0) * 5.0  # the grid before jitter (5 m spacing)
    c.y = np.round(c.y / 5.0) * 5.0
    assert abs(W.friction_multiplier(c) - 1.0) < 1e-12
    assert abs(W.friction_multiplier(_reach()) - 1.0) > 1e-3


def _gvf_frame(h_out, n=0.03, q=2.0, S=1e-3, span=1000.0):
    dist = np.linspace(0.0, span, 501)
    fr = W.Frame(src=None, dst=None, node=None, edge=None, zc=None, dist=dist,
                 e=np.array([1.0, 0.0]), H=10.0 + h_out, q=q, S=S, a=10.0)
    return fr


def test_backwater_profile_closed_forms():
    """gvf_depth against what fixes it: at the normal depth the flow stays
    uniform; above it the M1 curve falls monotonically upstream towards the
    normal depth; below it the M2 curve rises towards it; the profile obeys
    (1 - Fr^2) dh/dx = S - Sf (x downstream) wherever it is sampled."""
    n, q, S = 0.03, 2.0, 1e-3
    hn = (n * q / np.sqrt(S)) ** 0.6  # n^2 q^2 / hn^(10/3) = S
    h = W.gvf_depth(_gvf_frame(hn, n, q, S), n)
    assert np.abs(h - hn).max() < 1e-9 * hn
    for h_out in (hn + 0.3, hn - 0.2):
        fr = _gvf_frame(h_out, n, q, S)
        h = W.gvf_depth(fr, n)
        d = np.diff(h)
        if h_out > hn:
            assert (d <= 0).all() and h[-1] > hn and abs(h[0] - h_out) < 1e-12
        else:
            assert (d >= 0).all() and h[-1] < hn and abs(h[0] - h_out) < 1e-12
        # the ODE at interior samples (dist points upstream: dh/ddist = -dh/dx)
        s = fr.dist
        dh = (h[2:] - h[:-2]) / (s[2:] - s[:-2])
        hm = h[1:-1]
        rhs = -(S - n * n * q * q / hm ** (10.0 / 3.0)) / (1.0 - q * q / (9.81 * hm ** 3))
        assert np.abs(dh - rhs).max() < 1e-4 * np.abs(rhs).max() + 1e-12


def _gvf_steps_in_python(q, S, n, h0, h_floor, ds, n_steps):
    """The integration as the network's training data was made with it:
    RK4 in Python floats (libm pow)."""
    def rate(h):
        h = max(h, h_floor)
        fr2 = q * q / (9.81 * h ** 3)
        sf = n * n * q * q / h ** (10.0 / 3.0)
        return -(S - sf) / (1.0 - fr2)
    h = [h0]
    for _ in range(n_steps):
        y = h[-1]
        k1 = rate(y)
        k2 = rate(y + 0.5 * ds * k1)
        k3 = rate(y + 0.5 * ds * k2)
        k4 = rate(y + ds * k3)
        h.append(max(y + ds * (k1 + 2 * k2 + 2 * k3 + k4) / 6.0, h_floor))
    return np.array(h)


def test_gvf_profile_is_bit_identical_to_the_python_steps():
    """storm_mpnn_gvf_profile (host C++) gives the same bits as the Python
    RK4 steps, M1 and M2 curves and the critical-depth floor included, so
    the starts and the network's features are unchanged by it."""
    from storm_amd import mpnn
    rng = np.random.default_rng(5)
    for _ in range(40):
        q = float(rng.uniform(0.3, 8.0))
        S = float(10 ** rng.uniform(-4.5, -2.5))
        n = float(rng.uniform(0.02, 0.08))
        hc = (q * q / 9.81) ** (1.0 / 3.0)
        hn = (n * q / np.sqrt(S)) ** 0.6
        h0 = float(max(hn * rng.uniform(0.3, 2.5), 1.1 * hc))
        ds = float(rng.uniform(0.05, 5.0))
        want = _gvf_steps_in_python(q, S, n, h0, 1.1 * hc, ds, 300)
        got = mpnn.gvf_profile(q, S, n, h0, 1.1 * hc, ds, 300)
        assert np.array_equal(got, want)
    assert mpnn.gvf_profile(1.0, 1e-3, 0.03, 1.0, 0.5, 1.0, 0).tolist() == [1.0]


def test_parallel_builders_equal_the_serial_ones():
    cs = [cases.ensemble_reach(k, n_min=2000, n_max=4000, base_seed=91) for k in range(3)]
    for kind, fn in (("plane", W.baseline_state), ("gvf", W.gvf_state), ("cold", W.cold_state)):
        for a, b in zip(W.starts(cs, kind), [fn(c) for c in cs]):
            assert all(np.array_equal(x, y) for x, y in zip(a, b)), kind
    for a, b in zip(W.frames(cs), [W.frame(c) for c in cs]):
        assert np.array_equal(a.node, b.node) and np.array_equal(a.edge, b.edge)
        assert np.array_equal(a.base, b.base)
