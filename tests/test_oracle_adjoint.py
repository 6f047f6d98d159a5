"""Pinning the oracle's adjoint (CPU): central finite differences of the FP64
forward solve, the row-scale identity, and the closed forms of a directed
cycle.  A dropped sink term, a transposed edge or a missing 1/deg fails one."""
import numpy as np
import pytest

import oracle
from graphs import Rng


def _random(n, m, seed, zero_frac=0.0):
    r = Rng(seed)
    s = np.array([r.below(n) for _ in range(m)], np.uint32)
    t = np.array([r.below(n) for _ in range(m)], np.uint32)
    w = np.array([0.1 + (r.next() % 1000) / 1000.0 for _ in range(m)])
    c = np.array([((r.next() % 2001) - 1000) / 1000.0 for _ in range(n)])
    p = np.array([1.0 + (r.next() % 9) for _ in range(n)], np.float64)
    return s, t, w, c, p / p.sum()


def _loss(g, w, p, d, c):
    x, _ = g.pagerank(weight=w, pers=p, damping=d, epsilon=1e-15, max_iter=100_000)
    return float(c @ x), x


@pytest.mark.parametrize("seed", [1, 2])
def test_adjoint_matches_central_differences(seed):
    n, m = 40, 240
    s, t, w, c, p = _random(n, m, seed)
    g = oracle.Adj(n, s, t)
    d = 0.85
    L0, x = _loss(g, w, p, d, c)
    gw, gp, gd, lam, it = g.pagerank_grad(x, c, weight=w, pers=p, damping=d, epsilon=1e-15)
    assert it > 5
    h = 1e-4
    scale = np.abs(gw).max()
    for e in range(0, m, 7):
        wp, wm = w.copy(), w.copy()
        wp[e] += h
        wm[e] -= h
        fd = (_loss(g, wp, p, d, c)[0] - _loss(g, wm, p, d, c)[0]) / (2 * h)
        assert abs(fd - gw[e]) <= 1e-7 * scale + 1e-6 * abs(fd), (e, fd, gw[e])
    for v in range(0, n, 5):
        pp, pm = p.copy(), p.copy()
        pp[v] += h
        pm[v] -= h
        fd = (_loss(g, w, pp, d, c)[0] - _loss(g, w, pm, d, c)[0]) / (2 * h)
        assert abs(fd - gp[v]) <= 1e-7 * np.abs(gp).max() + 1e-6 * abs(fd), (v, fd, gp[v])
    fd = (_loss(g, w, p, d + h, c)[0] - _loss(g, w, p, d - h, c)[0]) / (2 * h)
    assert abs(fd - gd) <= 1e-6 * max(1.0, abs(fd)), (fd, gd)


def test_row_scale_identity():
    n, m = 60, 400
    s, t, w, c, p = _random(n, m, 5)
    g = oracle.Adj(n, s, t)
    _, x = _loss(g, w, p, 0.85, c)
    gw, *_ = g.pagerank_grad(x, c, weight=w, pers=p, damping=0.85)
    per_src = np.zeros(n)
    np.add.at(per_src, s.astype(np.int64), w * gw)
    assert np.abs(per_src).max() <= 1e-12 * max(1.0, np.abs(w * gw).max())


def test_directed_cycle_has_no_weight_or_damping_gradient():
    n = 9
    s = np.arange(n, dtype=np.uint32)
    t = ((np.arange(n) + 1) % n).astype(np.uint32)
    w = np.linspace(0.5, 2.0, n)
    c = np.linspace(-1.0, 1.0, n)
    g = oracle.Adj(n, s, t)
    x, _ = g.pagerank(weight=w, epsilon=1e-15, max_iter=100_000)
    gw, gp, gd, _, _ = g.pagerank_grad(x, c, weight=w)
    assert np.abs(gw).max() < 1e-12
    assert abs(gd) < 1e-12


def test_unweighted_graph_with_sinks_matches_differences():
    n, m = 30, 90
    s, t, w, c, p = _random(n, m, 8)
    keep = s % 4 != 1  # vertices 1, 5, 9, ... lose their out-edges: sinks
    s, t, w = s[keep], t[keep], w[keep]
    g = oracle.Adj(n, s, t)
    d = 0.7
    L0, x = _loss(g, w, p, d, c)
    gw, gp, gd, _, _ = g.pagerank_grad(x, c, weight=w, pers=p, damping=d)
    h = 1e-4
    for e in range(0, len(w), 5):
        wp, wm = w.copy(), w.copy()
        wp[e] += h
        wm[e] -= h
        fd = (_loss(g, wp, p, d, c)[0] - _loss(g, wm, p, d, c)[0]) / (2 * h)
        assert abs(fd - gw[e]) <= 1e-7 * np.abs(gw).max() + 1e-6 * abs(fd)
    fd = (_loss(g, w, p, d + h, c)[0] - _loss(g, w, p, d - h, c)[0]) / (2 * h)
    assert abs(fd - gd) <= 1e-6 * max(1.0, abs(fd))
