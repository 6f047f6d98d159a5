"""Batched personalised PageRank (one SpMM per sweep) against the oracle run
once per column.  Column c must equal pagerank_with with the one-hot
personalisation on seeds[c]: the FP32 state keeps L1 within 2e-6 of the FP64
oracle per column (FP32 rounding plus at most one sweep of convergence error)
and the per-column sweep count within one of the oracle's."""
import numpy as np
import pytest

import gt_hip
import oracle
from graphs import Rng

pytestmark = pytest.mark.gpu


def _graph(scale, seed):
    n = 1 << scale
    s, t = oracle.gen_rmat(scale, 16 * n, 0.57, 0.19, 0.19, seed)
    rng = np.random.default_rng(seed)
    w = np.exp(rng.standard_normal(len(s))).astype(np.float32)
    return n, s, t, w


@pytest.mark.parametrize("k", [64, 5, 70])
def test_batched_columns_match_single_runs(gpu, k):
    n, s, t, w = _graph(11, 3)
    g = gt_hip.AdjList.with_vertices(n)
    g.add_edges(s, t)
    r = Rng(k)
    seeds = []
    while len(seeds) < k:
        v = r.below(n)
        if v not in seeds:
            seeds.append(v)
    params = gt_hip.PowerIteration(1e-6, 1000, gt_hip.Plan(n, 64))
    X, iters = gt_hip.pagerank_batched(g, gt_hip.DenseProp.from_vec(g.graph_id, w), seeds,
                                       0.85, params)
    o = oracle.Adj(n, s, t)
    w64 = w.astype(np.float64)
    for c in range(k):
        pers = np.zeros(n)
        pers[seeds[c]] = 1.0
        want, it_o = o.pagerank(weight=w64, pers=pers, epsilon=1e-6, max_iter=1000)
        l1 = float(np.abs(X[:, c].astype(np.float64) - want).sum())
        assert abs(int(iters[c]) - it_o) <= 1, (c, iters[c], it_o)
        assert l1 <= 2e-6, (c, l1)


def test_batched_unweighted_on_device(gpu):
    import torch
    n, s, t, _ = _graph(12, 9)
    g = gt_hip.AdjList.with_vertices(n)
    g.add_edges(s, t)
    seeds = list(range(0, 64 * 7, 7))
    out = torch.zeros((n, 64), dtype=torch.float32, device="cuda")
    params = gt_hip.PowerIteration(1e-7, 1000, gt_hip.Plan(n, 64))
    X, iters = gt_hip.pagerank_batched(g, gt_hip.Unity.NEW, seeds, 0.85, params, out=out)
    Xh = X.cpu().numpy()
    o = oracle.Adj(n, s, t)
    for c in (0, 17, 63):
        pers = np.zeros(n)
        pers[seeds[c]] = 1.0
        want, it_o = o.pagerank(pers=pers, epsilon=1e-7, max_iter=1000)
        assert abs(int(iters[c]) - it_o) <= 1
        assert np.abs(Xh[:, c] - want).sum() <= 2e-6


def test_batched_max_iter_caps_every_column(gpu):
    n, s, t, w = _graph(10, 4)
    g = gt_hip.AdjList.with_vertices(n)
    g.add_edges(s, t)
    params = gt_hip.PowerIteration(1e-12, 3, gt_hip.Plan(n, 64))
    X, iters = gt_hip.pagerank_batched(g, gt_hip.Unity.NEW, [1, 2, 3], 0.85, params)
    assert (iters == 3).all()
    o = oracle.Adj(n, s, t)
    pers = np.zeros(n)
    pers[2] = 1.0
    want, it_o = o.pagerank(pers=pers, epsilon=1e-12, max_iter=3)
    assert it_o == 3
    assert np.abs(X[:, 1] - want).sum() <= 1e-6


def test_batched_rejects_bad_seed(gpu):
    g = gt_hip.AdjList.with_vertices(4)
    g.add_edge(0, 1)
    with pytest.raises(ValueError):
        gt_hip.pagerank_batched(g, gt_hip.Unity.NEW, [5], 0.85,
                                gt_hip.PowerIteration(1e-6, 10, gt_hip.Plan(4, 64)))


@pytest.mark.parametrize("weighted", [False, True])
def test_batched_split_rows_and_empty_rows(gpu, weighted):
    """256-entry SpMM tiles: a hub row of 3,000 in-entries is summed in
    twelve chunks, rows with no in-entries (isolated vertices) sit between
    the tiles' rows, and the columns still equal the single-vector runs."""
    n, s, t, _ = _graph(10, 21)
    n_all = n + 300  # the last 300 vertices have no edges
    rng = np.random.default_rng(5)
    hub_src = rng.integers(0, n, 3000).astype(np.uint32)
    s = np.concatenate([s, hub_src]).astype(np.uint32)
    t = np.concatenate([t, np.full(3000, 17, np.uint32)]).astype(np.uint32)
    w = np.exp(rng.standard_normal(len(s))).astype(np.float32)
    g = gt_hip.AdjList.with_vertices(n_all)
    g.add_edges(s, t)
    g.device().set_tuning(mm_tile_nnz=256)
    seeds = [17, 0, 5, n_all - 1] + list(range(100, 100 + 60))
    params = gt_hip.PowerIteration(1e-7, 1000, gt_hip.Plan(n_all, 64))
    wp = gt_hip.DenseProp.from_vec(g.graph_id, w) if weighted else gt_hip.Unity.NEW
    X, iters = gt_hip.pagerank_batched(g, wp, seeds, 0.85, params)
    o = oracle.Adj(n_all, s, t)
    for c in (0, 1, 2, 3, 40, 63):
        pers = np.zeros(n_all)
        pers[seeds[c]] = 1.0
        want, it_o = o.pagerank(weight=w.astype(np.float64) if weighted else None, pers=pers,
                                epsilon=1e-7, max_iter=1000)
        assert abs(int(iters[c]) - it_o) <= 1, (c, iters[c], it_o)
        assert np.abs(X[:, c].astype(np.float64) - want).sum() <= 2e-6, c
    g.device().set_tuning(mm_tile_nnz=0)
