"""Uneven batch: every graph of a block-diagonal batch against the oracle's
pagerank_with / adjoint on that graph alone."""
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
        want, ito = o.pagerank(weight=wk if weighted else None, damping=float(d[k]),
                               epsilon=1e-6, grain=64)
        got = x[voff[k]:voff[k + 1]]
        assert np.abs(got - want).sum() <= 1e-5, k
        assert abs(int(it[k]) - ito) <= 1, (k, int(it[k]), ito)


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
