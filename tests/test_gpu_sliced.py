"""GPU parity of the sliced PageRank sweep (long rows' entries in the warm
slices gathered from LDS copies of those slices, the rest swept with the light
rows, the pieces added in slot order by the fold) against the oracle and
against the same solve with the split turned off.

Bars: the FP64 drop-in at L1 <= 1e-12 with sweeps +-1 (parity with the
oracle); the FP32 form at L1 <= 1e-6; a repeated solve bit for bit.  The
cases drive one slice per workgroup and many (1 KB slices), pieces longer
than a wave tile (chunked), personalisation, and the workgroup split.
"""
import numpy as np
import pytest

import gt_hip
import oracle

pytestmark = pytest.mark.gpu


def _params(n, eps, max_iter=1000):
    return gt_hip.PowerIteration(eps, max_iter, gt_hip.Plan(n, 64))


def _solve(g, fn, eps, pers=None):
    r = gt_hip.DenseProp.new(g.graph_id)
    if fn == "f64":
        it = (gt_hip.pagerank(g, gt_hip.Unity.NEW, 0.85, _params(g.vertex_bound(), eps), r)
              if pers is None else
              gt_hip.pagerank_with(g, gt_hip.Unity.NEW, pers, 0.85,
                                   _params(g.vertex_bound(), eps), r))
    else:
        it = gt_hip.pagerank_mixed(g, gt_hip.Unity.NEW, pers, 0.85,
                                   _params(g.vertex_bound(), eps), r)
    return np.asarray(r.data).copy(), it


def _parity(got, it, want, it_o, eps, tol=1e-12):
    l1 = float(np.abs(got - want).sum())
    if it == it_o:
        assert l1 <= tol, f"L1 {l1:.3e} at {it} sweeps"
    else:
        assert abs(it - it_o) == 1, f"sweeps {it} vs {it_o}"
        assert l1 <= eps, f"L1 {l1:.3e}, sweeps {it} vs {it_o}"


def _graph(n, s, t):
    g = gt_hip.AdjList.with_vertices(n)
    g.add_edges(s, t)
    return g


def _hub_graph():
    """R-MAT 12 plus two hubs.  Vertex 9 takes 6,000 in-edges from the five
    most gathered sources (multi-edges): a main piece longer than a wave tile.
    Vertex 11 takes 8 in-edges from each source ranked 256..511 by gather
    frequency: with 256-label slices, a warm piece longer than a wave tile
    (checked below on the final labelling)."""
    n = 1 << 12
    s, t = oracle.gen_rmat(12, 16 * n, 0.57, 0.19, 0.19, 21)
    deg = np.bincount(s, minlength=n)
    rank = np.argsort(-deg, kind="stable")
    top = rank[:5].astype(np.uint32)
    warm = rank[256:512].astype(np.uint32)
    s = np.concatenate([s, np.tile(top, 1200), np.repeat(warm, 8)]).astype(np.uint32)
    t = np.concatenate([t, np.full(6000, 9, np.uint32), np.full(256 * 8, 11, np.uint32)])
    t = t.astype(np.uint32)
    order = np.argsort(-np.bincount(s, minlength=n), kind="stable")
    label = np.empty(n, np.int64)
    label[order] = np.arange(n)
    into11 = label[s[t == 11]]
    assert ((into11 >= 256) & (into11 < 512)).sum() > 1024
    return n, s, t


CASES = {
    # (n, src, tgt, sweep_hot_kb, sliced_min_len)
    "rmat16_4kb": lambda: (1 << 16,) + oracle.gen_rmat(16, 16 << 16, 0.57, 0.19, 0.19, 6) + (4, 64),
    "rmat14_1kb": lambda: (1 << 14,) + oracle.gen_rmat(14, 16 << 14, 0.57, 0.19, 0.19, 7) + (1, 64),
    "hub_1kb": lambda: _hub_graph() + (1, 16),
}


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("personalised", [False, True])
def test_sliced_matches_oracle_and_unsliced(gpu, case, personalised):
    n, s, t, kb, min_len = CASES[case]()
    p = None
    if personalised:
        p = np.random.default_rng(n).uniform(0.0, 2.0, n)
        p /= p.sum()
    g = _graph(n, s, t)
    dev = g.device()
    # The FP32 solves on the wave engine (the sliced split is one of its sweeps).
    dev.set_tuning(sweep_hot_kb=kb, sliced_min_len=min_len, block_tile_nnz=-1)
    o = oracle.Adj(n, s, t)
    want, it_o = o.pagerank(pers=p, epsilon=1e-12, max_iter=1000, threads=8)
    pp = None if p is None else gt_hip.DenseProp.from_vec(g.graph_id, p)
    got, it = _solve(g, "f64", 1e-12, pp)
    lay = dev.layout()
    assert lay["slices"] >= 1 and lay["sliced_rows"] > 0, lay
    _parity(got, it, want, it_o, 1e-12)
    x32, _ = _solve(g, "mixed", 1e-9, pp)
    assert dev.layout()["slices"] >= 1
    assert np.abs(x32 - want).sum() <= 1e-6
    # The same solves with the split off: same answers to rounding.
    dev.set_tuning(sliced_min_len=-1)
    off64, it_off = _solve(g, "f64", 1e-12, pp)
    assert dev.layout()["slices"] == 0
    _parity(off64, it_off, want, it_o, 1e-12)
    off32, _ = _solve(g, "mixed", 1e-9, pp)
    assert np.abs(off32 - x32).sum() <= 1e-7
    dev.set_tuning(sweep_hot_kb=0, sliced_min_len=0, block_tile_nnz=0)


def test_sliced_repeats_bit_for_bit(gpu):
    n, s, t, kb, min_len = CASES["hub_1kb"]()
    g = _graph(n, s, t)
    g.device().set_tuning(sweep_hot_kb=kb, sliced_min_len=min_len, block_tile_nnz=-1)
    a, ia = _solve(g, "f64", 1e-12)
    b, ib = _solve(g, "f64", 1e-12)
    assert ia == ib and (a.view(np.uint64) == b.view(np.uint64)).all()
    c, ic = _solve(g, "mixed", 1e-9)
    d, id_ = _solve(g, "mixed", 1e-9)
    assert ic == id_ and (c.view(np.uint64) == d.view(np.uint64)).all()


@pytest.mark.parametrize("share", [1, 30, 70, 99])
def test_sliced_workgroup_split(gpu, share):
    """However the launch's workgroups are divided between pieces and light
    rows (down to one of either), the answer is the oracle's."""
    n = 1 << 15
    s, t = oracle.gen_rmat(15, 16 * n, 0.57, 0.19, 0.19, 13)
    g = _graph(n, s, t)
    dev = g.device()
    dev.set_tuning(sweep_hot_kb=4, sliced_min_len=64, sliced_share=share, block_tile_nnz=-1)
    want, it_o = oracle.Adj(n, s, t).pagerank(epsilon=1e-12, max_iter=1000, threads=8)
    got, it = _solve(g, "f64", 1e-12)
    assert dev.layout()["slices"] >= 1
    _parity(got, it, want, it_o, 1e-12)
    x32, _ = _solve(g, "mixed", 1e-9)
    assert np.abs(x32 - want).sum() <= 1e-6
