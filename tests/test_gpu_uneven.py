"""Uneven batch: every graph of a block-diagonal batch against the oracle's
pagerank_with / adjoint on that graph alone.

Forward bar: each graph's sweep count within one of the oracle's at the same
epsilon, and its ranks within L1 <= 1e-6 of the oracle's iterate after the
same number of sweeps (the FP32 form against FP64: rounding only)."""
import numpy as np
import pytest

import gt_hip
import oracle

pytestmark = pytest.mark.gpu


def _batch(G, seed, n_max=1500):
    voff, eoff, d, s, t, w = gt_hip.gen_chung_lu_batch(G, seed, n_min=64, n_max=n_max)
    n = int(voff[-1])
    dev = gt_hip.DeviceGraph.from_edges(n, s, t)
    return voff, eoff, d, s, t, w, n, dev


def check_forward(o, got, it, weight, damping, eps=1e-6):
    """Sweeps within one of the oracle's at `eps`, or else a stop where the
    oracle's own L1 change is within 10 % of eps (below 1.1 eps at sweep
    `it`, at least 0.9 eps at sweep it - 1); L1 <= 1e-6 against the oracle's
    iterate after exactly `it` sweeps.

    The second form is for slowly converging graphs: an FP32 iterate's L1
    change differs from the FP64 one by rounding.  On the damping-0.95 hub
    graph of the edge-case batch the change falls by 5 % a sweep, and the
    FP32 iterate -- emulated on the host with exact row sums -- stops at
    sweep 281 where the FP64 iteration stops at 283 (its change at 281:
    1.099e-6)."""
    _, ito = o.pagerank(weight=weight, damping=damping, epsilon=eps, grain=64)
    want, _ = o.pagerank(weight=weight, damping=damping, epsilon=1e-300, max_iter=it, grain=64)
    if abs(it - ito) > 1:
        assert it >= 2, (it, ito)
        w1, _ = o.pagerank(weight=weight, damping=damping, epsilon=1e-300, max_iter=it - 1,
                           grain=64)
        w2, _ = o.pagerank(weight=weight, damping=damping, epsilon=1e-300, max_iter=it - 2,
                           grain=64)
        d_it = float(np.abs(want - w1).sum())
        d_prev = float(np.abs(w1 - w2).sum())
        assert d_it < 1.1 * eps and d_prev >= 0.9 * eps, (it, ito, d_it, d_prev)
    l1 = float(np.abs(got - want).sum())
    assert l1 <= 1e-6, (l1, it)
    return l1


def _graph(k, voff, eoff, s, t, w):
    lo, hi = int(eoff[k]), int(eoff[k + 1])
    nk = int(voff[k + 1] - voff[k])
    return nk, s[lo:hi] - voff[k], t[lo:hi] - voff[k], w[lo:hi].astype(np.float64)


@pytest.mark.parametrize("weighted", [False, True])
def test_uneven_forward_matches_each_graph(gpu, weighted):
    voff, eoff, d, s, t, w, n, dev = _batch(24, 101)
    wm = w if weighted else None
    x, it = gt_hip.pagerank_uneven(dev, voff, d, weight=wm, epsilon=1e-6)
    for k in range(24):
        nk, sk, tk, wk = _graph(k, voff, eoff, s, t, w)
        o = oracle.Adj(nk, sk, tk)
        check_forward(o, x[voff[k]:voff[k + 1]], int(it[k]), wk if weighted else None,
                      float(d[k]))


def test_uneven_max_iter_caps_every_graph(gpu):
    voff, eoff, d, s, t, w, n, dev = _batch(9, 7)
    x, it = gt_hip.pagerank_uneven(dev, voff, d, epsilon=1e-12, max_iter=3)
    assert np.all(it == 3)
    for k in range(9):
        nk, sk, tk, _ = _graph(k, voff, eoff, s, t, w)
        want, ito = oracle.Adj(nk, sk, tk).pagerank(damping=float(d[k]), epsilon=1e-12,
                                                    max_iter=3)
        assert ito == 3
        assert np.abs(x[voff[k]:voff[k + 1]] - want).sum() <= 1e-6


def test_uneven_rejects_edges_across_graphs(gpu):
    voff = np.array([0, 4, 8], np.uint32)
    s = np.array([0, 1, 3], np.uint32)
    t = np.array([1, 2, 5], np.uint32)  # 3 -> 5 joins graph 0 to graph 1
    dev = gt_hip.DeviceGraph.from_edges(8, s, t)
    with pytest.raises(ValueError, match="joins two graphs"):
        gt_hip.pagerank_uneven(dev, voff, [0.85, 0.85])
    with pytest.raises(ValueError):
        gt_hip.pagerank_uneven(dev, np.array([0, 4, 7], np.uint32), [0.85, 0.85])


def test_uneven_backward_matches_each_graph(gpu):
    G = 16
    voff, eoff, d, s, t, w, n, dev = _batch(G, 303, n_max=1200)
    rng = np.random.default_rng(5)
    c = rng.standard_normal(n)
    xs = np.zeros(n)
    want = []
    for k in range(G):
        nk, sk, tk, wk = _graph(k, voff, eoff, s, t, w)
        o = oracle.Adj(nk, sk, tk)
        xk, _ = o.pagerank(weight=wk, damping=float(d[k]), epsilon=1e-15, max_iter=10_000)
        xs[voff[k]:voff[k + 1]] = xk
        want.append(o.pagerank_grad(xk, c[voff[k]:voff[k + 1]], weight=wk, damping=float(d[k]),
                                    epsilon=1e-12))
    gw, gp, gd, it = gt_hip.pagerank_uneven_grad(dev, voff, d, xs, c, weight=w, epsilon=1e-9)
    for k in range(G):
        gw_o, gp_o, gd_o, _, _ = want[k]
        lo, hi = int(eoff[k]), int(eoff[k + 1])
        assert np.abs(gw[lo:hi] - gw_o).max() <= 2e-4 * np.abs(gw_o).max(), k
        gpk = gp[voff[k]:voff[k + 1]]
        assert np.abs(gpk - gp_o).max() <= 2e-4 * np.abs(gp_o).max(), k
        assert abs(gd[k] - gd_o) <= 2e-4 * max(1.0, abs(gd_o)), k
    assert np.all(it >= 1)


def _edge_case_batch():
    """Six graphs for the sweep's list of still-iterating tiles: empty graphs
    first, between the others and last; a hub whose in-row of 3,000 entries is
    split into chunks, with damping 0.95 so that it iterates far longer than
    the rest; a graph with damping 0.3 that stops after a few sweeps."""
    rng = np.random.default_rng(17)
    nh = 3001
    hub_s = np.concatenate([np.arange(1, nh), np.zeros(40, np.int64)])
    hub_t = np.concatenate([np.zeros(nh - 1, np.int64), rng.integers(1, nh, 40)])
    parts = [(0, np.zeros(0, np.int64), np.zeros(0, np.int64), 0.85),
             (nh, hub_s, hub_t, 0.95),
             (0, np.zeros(0, np.int64), np.zeros(0, np.int64), 0.5),
             (200, rng.integers(0, 200, 900), rng.integers(0, 200, 900), 0.3),
             (700, rng.integers(0, 700, 5000), rng.integers(0, 700, 5000), 0.85),
             (0, np.zeros(0, np.int64), np.zeros(0, np.int64), 0.7)]
    voff = np.zeros(len(parts) + 1, np.uint32)
    eoff = np.zeros(len(parts) + 1, np.int64)
    s, t = [], []
    for k, (nk, sk, tk, _) in enumerate(parts):
        s.append(sk + voff[k])
        t.append(tk + voff[k])
        voff[k + 1] = voff[k] + nk
        eoff[k + 1] = eoff[k] + len(sk)
    s = np.concatenate(s).astype(np.uint32)
    t = np.concatenate(t).astype(np.uint32)
    w = (1.0 - 0.9 * rng.random(len(s))).astype(np.float32)
    d = np.array([p[3] for p in parts])
    n = int(voff[-1])
    return voff, eoff, d, s, t, w, n, gt_hip.DeviceGraph.from_edges(n, s, t)


# Block tiles cut at the graphs' bounds (the default), small tiles with many
# long rows and chunk tiles, and the wave engine.
TUNINGS = {"blocks": {}, "small_blocks": {"block_tile_nnz": 1024, "block_long_min": 64},
           "wave": {"block_tile_nnz": -1}}


@pytest.mark.parametrize("tune", list(TUNINGS))
def test_uneven_list_edge_cases(gpu, tune):
    """Empty graphs, a split hub row and sweep counts from a few to hundreds,
    forward and adjoint, each non-empty graph against the oracle alone."""
    voff, eoff, d, s, t, w, n, dev = _edge_case_batch()
    dev.set_tuning(**TUNINGS[tune])
    G = len(d)
    live = [k for k in range(G) if voff[k + 1] > voff[k]]
    x, it = gt_hip.pagerank_uneven(dev, voff, d, weight=w, epsilon=1e-6)
    lay = dev.layout()
    assert (lay["block_tiles"] > 0) == (tune != "wave"), lay
    rng = np.random.default_rng(3)
    c = rng.standard_normal(n)
    xs = np.zeros(n)
    grads = {}
    for k in live:
        nk, sk, tk, wk = _graph(k, voff, eoff, s, t, w)
        o = oracle.Adj(nk, sk, tk)
        check_forward(o, x[voff[k]:voff[k + 1]], int(it[k]), wk, float(d[k]))
        xk, _ = o.pagerank(weight=wk, damping=float(d[k]), epsilon=1e-15, max_iter=100_000)
        xs[voff[k]:voff[k + 1]] = xk
        grads[k] = o.pagerank_grad(xk, c[voff[k]:voff[k + 1]], weight=wk, damping=float(d[k]),
                                   epsilon=1e-12)
    # The hub graph iterates long after the damping-0.3 graph has stopped.
    assert int(it[1]) > 10 * int(it[3]), (int(it[1]), int(it[3]))
    gw, gp, gd, _ = gt_hip.pagerank_uneven_grad(dev, voff, d, xs, c, weight=w, epsilon=1e-9)
    for k in live:
        gw_o, gp_o, gd_o, _, _ = grads[k]
        lo, hi = int(eoff[k]), int(eoff[k + 1])
        assert np.abs(gw[lo:hi] - gw_o).max() <= 2e-4 * np.abs(gw_o).max(), k
        assert np.abs(gp[voff[k]:voff[k + 1]] - gp_o).max() <= 2e-4 * np.abs(gp_o).max(), k
        assert abs(gd[k] - gd_o) <= 2e-4 * max(1.0, abs(gd_o)), k


@pytest.mark.parametrize("weighted", [False, True])
def test_uneven_blocks_agree_with_wave_engine(gpu, weighted):
    """The batch on block tiles and on the wave engine: each graph's forward
    sweep count within one (adjoint: within 2 %), its ranks within L1 2e-6
    (both are FP32 iterates within 1e-6 of the oracle's), and the adjoint's
    gradients within 2e-4 of the largest."""
    G = 20
    voff, eoff, d, s, t, w, n, dev = _batch(G, 909)
    wm = w if weighted else None
    rng = np.random.default_rng(11)
    c = rng.standard_normal(n)
    out = {}
    for tune in ("blocks", "wave"):
        dev.set_tuning(**TUNINGS[tune])
        x, it = gt_hip.pagerank_uneven(dev, voff, d, weight=wm, epsilon=1e-6)
        assert (dev.layout()["block_tiles"] > 0) == (tune == "blocks")
        gw, gp, gd, ita = gt_hip.pagerank_uneven_grad(dev, voff, d, x.astype(np.float64), c,
                                                      weight=wm, epsilon=1e-9)
        out[tune] = (x, it, gw, gp, gd, ita)
    xb, itb, gwb, gpb, gdb, itab = out["blocks"]
    xw, itw, gww, gpw, gdw, itaw = out["wave"]
    assert np.all(np.abs(itb.astype(np.int64) - itw.astype(np.int64)) <= 1)
    # The adjoint stops at 2^-22 of ||dL/dx_g||_1, where the FP32 iterate's
    # change is within rounding of the threshold for a few sweeps: counts of
    # 150-200 sweeps differ by up to 2 between the two layouts' roundings.
    ita_b, ita_w = itab.astype(np.int64), itaw.astype(np.int64)
    assert np.all(np.abs(ita_b - ita_w) <= np.maximum(1, ita_w // 50)), (ita_b, ita_w)
    for k in range(G):
        lo, hi = int(voff[k]), int(voff[k + 1])
        assert np.abs(xb[lo:hi] - xw[lo:hi]).sum() <= 2e-6, k
        elo, ehi = int(eoff[k]), int(eoff[k + 1])
        if weighted:
            assert np.abs(gwb[elo:ehi] - gww[elo:ehi]).max() <= 2e-4 * np.abs(gww[elo:ehi]).max(), k
        assert np.abs(gpb[lo:hi] - gpw[lo:hi]).max() <= 2e-4 * np.abs(gpw[lo:hi]).max(), k
        assert abs(gdb[k] - gdw[k]) <= 2e-4 * max(1.0, abs(gdw[k])), k


def test_uneven_negative_weight_runs_on_the_wave_engine(gpu):
    """A negative weight breaks the block sweeps' bound on their fixed-point
    sums (a source's terms sum to at most its |x|): the batch runs on the
    wave engine, and every graph still matches the oracle alone."""
    G = 6
    voff, eoff, d, s, t, w, n, dev = _batch(G, 515, n_max=600)
    w = w.copy()
    w[int(eoff[2]) + 3] = -0.25
    x, it = gt_hip.pagerank_uneven(dev, voff, d, weight=w, epsilon=1e-6)
    assert dev.layout()["block_tiles"] == 0
    for k in range(G):
        nk, sk, tk, wk = _graph(k, voff, eoff, s, t, w)
        check_forward(oracle.Adj(nk, sk, tk), x[voff[k]:voff[k + 1]], int(it[k]), wk, float(d[k]))
