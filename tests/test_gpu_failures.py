"""Failure paths beyond the reference's three panics (those are covered in
test_gpu_pagerank.py): NaN weights, a device allocation that fails, and a
zero-norm eigenvector on a filtered view -- each against the oracle, which
restates the reference's comparisons (centrality.rs:261-384, 405-497).

A NaN weight makes its source's weighted out-degree NaN; NaN != 0, so the
source is no sink and its out-neighbours' ranks turn NaN; the sweep's L1
change is NaN and `NaN >= epsilon` is false, so the loop stops after one
sweep (PageRank) -- the eigenvector's norm turns NaN likewise.
"""
import numpy as np
import pytest

import gt_hip
import oracle
from gt_hip import _lib
from graphs import thousand

pytestmark = pytest.mark.gpu


def _params(n, eps, max_iter=1000):
    return gt_hip.PowerIteration(eps, max_iter, gt_hip.Plan(n, 64))


def _same_nan_pattern(got, want, tol=1e-12):
    gn, wn = np.isnan(got), np.isnan(want)
    assert (gn == wn).all(), (int(gn.sum()), int(wn.sum()))
    assert np.abs(got[~gn] - want[~wn]).sum() <= tol


def test_nan_weight_pagerank_stops_after_one_sweep(gpu):
    n, s, t = thousand()
    g = gt_hip.AdjList.with_vertices(n)
    g.add_edges(s, t)
    w = np.random.default_rng(4).uniform(0.5, 1.5, len(s))
    w[17] = np.nan
    o = oracle.Adj(n, s, t)
    want, it_o = o.pagerank(weight=w, epsilon=1e-10, max_iter=1000)
    assert it_o == 1 and np.isnan(want).any()
    r = gt_hip.DenseProp.new(g.graph_id)
    it = gt_hip.pagerank(g, gt_hip.DenseProp.from_vec(g.graph_id, w), 0.85,
                         _params(n, 1e-10), r)
    assert it == it_o
    _same_nan_pattern(np.asarray(r.data), want)
    # The FP32 form stops after one sweep too, with no refinement sweep (a NaN
    # stop is no convergence), so the same NaNs.
    r32 = gt_hip.DenseProp.new(g.graph_id)
    it32 = gt_hip.pagerank_mixed(g, gt_hip.DenseProp.from_vec(g.graph_id, w), None, 0.85,
                                 _params(n, 1e-10), r32)
    assert it32 == 1
    x32 = np.asarray(r32.data)
    assert (np.isnan(x32) == np.isnan(want)).all()
    assert np.abs(x32 - want)[~np.isnan(want)].sum() <= 1e-6


def test_nan_weight_eigenvector_stops_after_one_sweep(gpu):
    n, s, t = thousand()
    g = gt_hip.AdjList.with_vertices(n)
    g.add_edges(s, t)
    w = np.ones(len(s))
    w[5] = np.nan
    xo, eo, ito = oracle.Adj(n, s, t).eigenvector(weight=w, epsilon=1e-10, max_iter=1000)
    x = gt_hip.DenseProp.new(g.graph_id)
    eig, it = gt_hip.eigenvector(g, gt_hip.DenseProp.from_vec(g.graph_id, w), _params(n, 1e-10),
                                 x)
    assert it == ito
    assert np.isnan(eig) == np.isnan(eo)
    _same_nan_pattern(np.asarray(x.data), xo)


def test_zero_norm_eigenvector_on_filtered_view(gpu):
    """Every edge filtered out: each sweep's sum is 0, the norm 0 (scale 1),
    the first sweep changes every kept slot by its start value, the second by
    nothing; dropped vertices keep theirs."""
    n = 1 << 12
    s, t = oracle.gen_rmat(12, 16 * n, 0.57, 0.19, 0.19, 8)
    g = gt_hip.AdjList.with_vertices(n)
    g.add_edges(s, t)
    vm = (np.random.default_rng(2).random(n) < 0.7).astype(np.uint8)
    em = np.zeros(len(s), np.uint8)
    view = g.masked(vm, em)
    xo, eo, ito = oracle.Adj(n, s, t).eigenvector(epsilon=1e-10, max_iter=1000, vmask=vm,
                                                  emask=em)
    x = gt_hip.DenseProp.new(g.graph_id)
    eig, it = gt_hip.eigenvector(view, gt_hip.Unity.NEW, _params(n, 1e-10), x)
    assert eo == 0.0 and eig == 0.0
    assert it == ito
    assert np.abs(np.asarray(x.data) - xo).sum() <= 1e-12


def test_device_allocation_failure_is_oom(gpu):
    n = 1 << 12
    s, t = oracle.gen_rmat(12, 16 * n, 0.57, 0.19, 0.19, 9)
    used = gt_hip.device_memory_in_use()
    try:
        gt_hip.set_device_memory_limit(used + 4096)
        with pytest.raises(gt_hip.GtError) as e:
            gt_hip.DeviceGraph.upload(77, n, len(s), *gt_hip.host_adjacency(n, s, t))
        assert e.value.status == _lib.GT_ERR_OOM
        assert "memory limit" in str(e.value)
        # A graph uploaded under no limit, then a solve whose schedule and
        # workspaces would pass a new one.
        gt_hip.set_device_memory_limit(0)
        g = gt_hip.AdjList.with_vertices(n)
        g.add_edges(s, t)
        g.device()
        gt_hip.set_device_memory_limit(gt_hip.device_memory_in_use() + 4096)
        r = gt_hip.DenseProp.new(g.graph_id)
        with pytest.raises(gt_hip.GtError) as e:
            gt_hip.pagerank(g, gt_hip.Unity.NEW, 0.85, _params(n, 1e-10), r)
        assert e.value.status == _lib.GT_ERR_OOM
        assert "exceeds" in _lib.lib().gt_last_error().decode()
    finally:
        gt_hip.set_device_memory_limit(0)
    # Without the limit the same graph solves as the oracle does.
    r = gt_hip.DenseProp.new(g.graph_id)
    it = gt_hip.pagerank(g, gt_hip.Unity.NEW, 0.85, _params(n, 1e-10), r)
    want, it_o = oracle.Adj(n, s, t).pagerank(epsilon=1e-10, max_iter=1000)
    assert abs(it - it_o) <= 1
    assert np.abs(np.asarray(r.data) - want).sum() <= 1e-10
