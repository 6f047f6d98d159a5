"""GPU parity of PageRank and eigenvector centrality against the oracle.

Bar (plain FP64 path): L1 <= 1e-12 against the oracle and the same sweep count
+-1 on the same inputs; the reference's known answers at their own tolerances;
reproducible bit for bit run to run.  Mixed path: L1 <= 1e-6.
"""
import math

import numpy as np
import pytest

import gt_hip
import oracle
from graphs import EDGES, Rng, bench, c1, sixty, thousand

pytestmark = pytest.mark.gpu


def adj(n, edges_or_src, tgt=None):
    g = gt_hip.AdjList.with_vertices(n)
    if tgt is None:
        for a, b in edges_or_src:
            g.add_edge(a, b)
    else:
        g.add_edges(edges_or_src, tgt)
    return g


def parity(got, it, want, it_o, eps, tol=1e-12):
    """Same sweep count and L1 <= tol; a count off by one is allowed only when
    the oracle's last delta sat within rounding of epsilon, and then the two
    answers are one sweep apart (L1 below epsilon)."""
    l1 = float(np.abs(np.asarray(got) - np.asarray(want)).sum())
    if it == it_o:
        assert l1 <= tol, f"L1 {l1:.3e} at {it} sweeps"
    else:
        assert abs(it - it_o) == 1, f"sweeps {it} vs {it_o}"
        assert l1 <= eps, f"L1 {l1:.3e}, sweeps {it} vs {it_o}"


def params(n, eps, max_iter, grain=64):
    return gt_hip.PowerIteration(eps, max_iter, gt_hip.Plan(n, grain))


def pr(g, eps, max_iter, weight=None, pers=None, rank=None, d=0.85, fn=None):
    r = gt_hip.DenseProp.new(g.graph_id) if rank is None else rank
    w = gt_hip.Unity.NEW if weight is None else weight
    if fn is not None:
        it = fn(g, w, pers, d, params(g.vertex_bound(), eps, max_iter), r)
    elif pers is None:
        it = gt_hip.pagerank(g, w, d, params(g.vertex_bound(), eps, max_iter), r)
    else:
        it = gt_hip.pagerank_with(g, w, pers, d, params(g.vertex_bound(), eps, max_iter), r)
    return np.asarray(r.data), it


# ------------------------------------------------------- known answers

def test_one_sweep_two_vertices(gpu):
    g = adj(2, [(0, 1)])
    r, it = pr(g, 1e-12, 1)
    assert it == 1
    assert abs(r[0] - 0.2875) < 1e-15 and abs(r[1] - 0.7125) < 1e-15
    r2, it2 = pr(g, 1e-12, 2)
    want = oracle.dense_pagerank(2, [(0, 1, 1.0)], [0.5, 0.5], 0.85, 2)
    assert it2 == 2 and np.abs(r2 - want).max() < 1e-15


def test_warm_start(gpu):
    g = adj(2, [(0, 1)])
    rank = gt_hip.DenseProp.from_vec(g.graph_id, np.array([0.2875, 0.7125]))
    r, _ = pr(g, 1e-12, 1, rank=rank)
    assert abs(r[0] - (0.15 * 0.5 + 0.85 * (0.7125 * 0.5))) < 1e-15
    assert abs(r[1] - (0.15 * 0.5 + 0.85 * (0.7125 * 0.5 + 0.2875))) < 1e-15


def test_partial_warm_start_seeds_only_new_slots(gpu):
    n, s, t = thousand()
    g = adj(n, s, t)
    start = np.linspace(0.0, 1e-3, 600)
    rank = gt_hip.DenseProp.from_vec(g.graph_id, start)
    r, it = pr(g, 1e-11, 1000, rank=rank)
    want, it_o = oracle.Adj(n, s, t).pagerank(epsilon=1e-11, max_iter=1000, rank=start)
    parity(r, it, want, it_o, 1e-11)


def test_zero_weight_predecessor(gpu):
    g = adj(3, [(0, 1), (1, 2)])
    w = gt_hip.DenseProp.from_vec(g.graph_id, np.array([0.0, 1.0]))
    r, _ = pr(g, 1e-13, 1000, weight=w)
    assert np.isfinite(r).all()
    h = adj(3, [(1, 2)])
    want, _ = pr(h, 1e-13, 1000)
    assert np.abs(r - want).max() < 1e-12


def test_personalisation_on_a_cycle(gpu):
    g = adj(4, [(0, 1), (1, 2), (2, 3), (3, 0)])
    flat, _ = pr(g, 1e-14, 10_000)
    assert np.abs(flat - 0.25).max() < 1e-12
    pers = gt_hip.DenseProp.from_vec(g.graph_id, np.array([0.7, 0.1, 0.1, 0.1]))
    got, _ = pr(g, 1e-14, 10_000, pers=pers)
    assert got[0] > got[1] > got[2] > got[3]
    want = oracle.dense_pagerank(4, [(0, 1, 1.0), (1, 2, 1.0), (2, 3, 1.0), (3, 0, 1.0)],
                                 [0.7, 0.1, 0.1, 0.1], 0.85, 400)
    assert np.abs(got - want).max() < 1e-12


def test_thousand_matches_dense_and_oracle(gpu):
    n, s, t = thousand()
    g = adj(n, s, t)
    got, it = pr(g, 1e-13, 10_000)
    want, it_o = oracle.Adj(n, s, t).pagerank(epsilon=1e-13, max_iter=10_000)
    dense = oracle.dense_pagerank(n, [(a, b, 1.0) for a, b in zip(s.tolist(), t.tolist())],
                                  np.full(n, 1.0 / n), 0.85, 400)
    parity(got, it, want, it_o, 1e-13)
    assert np.abs(got - dense).max() < 1e-12
    assert abs(got.sum() - 1.0) < 1e-9


def test_unity_is_explicit_ones_bitwise(gpu):
    n, s, t = thousand()
    g = adj(n, s, t)
    a, _ = pr(g, 1e-13, 10_000)
    ones = gt_hip.DenseProp.from_vec(g.graph_id, np.ones(len(s)))
    b, _ = pr(g, 1e-13, 10_000, weight=ones)
    assert (a.view(np.uint64) == b.view(np.uint64)).all()


def test_odd_and_even_truncation(gpu):
    n, s, t = thousand()
    g = adj(n, s, t)
    o = oracle.Adj(n, s, t)
    for k in (1, 2, 3, 7, 8):
        got, it = pr(g, 1e-30, k)
        want, it_o = o.pagerank(epsilon=1e-30, max_iter=k)
        assert it == it_o == k
        assert np.abs(got - want).sum() <= 1e-13


# -------------------------------------------------------- configurations

def test_plain_pagerank_workload_c1(gpu):
    """Directed G(10,000, 100,000), unweighted, d=0.85, eps 1e-6, max_iter 1000."""
    n, s, t = c1()
    g = adj(n, s, t)
    got, it = pr(g, 1e-6, 1000)
    want, it_o = oracle.Adj(n, s, t).pagerank(epsilon=1e-6, max_iter=1000)
    parity(got, it, want, it_o, 1e-6)


def test_bench_graph_and_rmat(gpu):
    for (n, s, t), eps in [(bench(), 1e-10), ((1 << 14,) + oracle.gen_rmat(
            14, 16 << 14, 0.57, 0.19, 0.19, 3), 1e-11)]:
        g = adj(n, s, t)
        got, it = pr(g, eps, 1000)
        want, it_o = oracle.Adj(n, s, t).pagerank(epsilon=eps, max_iter=1000, threads=8)
        parity(got, it, want, it_o, eps)


def test_long_rows_are_split_and_recombined(gpu):
    # A star: one hub with 20,000 in-entries (split into 10 chunks) plus noise.
    r = Rng(17)
    n = 3000
    s = [r.below(n) for _ in range(20_000)] + [r.below(n) for _ in range(5000)]
    t = [0] * 20_000 + [r.below(n) for _ in range(5000)]
    s, t = np.array(s, np.uint32), np.array(t, np.uint32)
    g = adj(n, s, t)
    w = np.array([0.5 + (i % 13) / 7.0 for i in range(len(s))])
    got, it = pr(g, 1e-12, 1000, weight=gt_hip.DenseProp.from_vec(g.graph_id, w))
    want, it_o = oracle.Adj(n, s, t).pagerank(weight=w, epsilon=1e-12, max_iter=1000)
    parity(got, it, want, it_o, 1e-12)


def test_weighted_and_personalised(gpu):
    r = Rng(23)
    n, m = 5000, 60_000
    s = np.array([r.below(n) for _ in range(m)], np.uint32)
    t = np.array([r.below(n) for _ in range(m)], np.uint32)
    w = np.array([(r.next() % 1000) / 100.0 for _ in range(m)])  # some zeros
    p = np.array([(r.next() % 100) / 50.0 for _ in range(n)])   # not normalised
    g = adj(n, s, t)
    got, it = pr(g, 1e-12, 1000, weight=gt_hip.DenseProp.from_vec(g.graph_id, w),
                 pers=gt_hip.DenseProp.from_vec(g.graph_id, p))
    want, it_o = oracle.Adj(n, s, t).pagerank(weight=w, pers=p, epsilon=1e-12, max_iter=1000)
    parity(got, it, want, it_o, 1e-12, tol=1e-12 * max(1.0, np.abs(want).sum()))


def test_reversed_view(gpu):
    n, s, t = thousand()
    g = adj(n, s, t)
    got, it = pr(g.reversed(), 1e-12, 1000)
    want, it_o = oracle.Adj(n, s, t).pagerank(reversed=True, epsilon=1e-12, max_iter=1000)
    parity(got, it, want, it_o, 1e-12)


def test_reproducible_bit_for_bit(gpu):
    n, s, t = bench()
    g = adj(n, s, t)
    a, ia = pr(g, 1e-10, 1000)
    b, ib = pr(g, 1e-10, 1000)
    assert ia == ib and (a.view(np.uint64) == b.view(np.uint64)).all()


def test_device_resident_maps(gpu):
    import torch
    n, s, t = bench()
    g = adj(n, s, t)
    rank = gt_hip.DenseProp(g.graph_id, torch.zeros(0, dtype=torch.float64, device="cuda"))
    it = gt_hip.pagerank(g, gt_hip.Unity.NEW, 0.85, params(n, 1e-10, 1000), rank)
    want, it_o = oracle.Adj(n, s, t).pagerank(epsilon=1e-10, max_iter=1000, threads=8)
    parity(rank.data.cpu().numpy(), it, want, it_o, 1e-10)


def test_mixed_precision_within_1e6(gpu):
    for n, s, t in [bench(), (1 << 14,) + oracle.gen_rmat(14, 16 << 14, 0.57, 0.19, 0.19, 3)]:
        g = adj(n, s, t)
        got, it = pr(g, 1e-9, 1000, fn=gt_hip.pagerank_mixed)
        want, _ = oracle.Adj(n, s, t).pagerank(epsilon=1e-12, max_iter=1000, threads=8)
        assert np.abs(got - want).sum() <= 1e-6
        assert it >= 1


@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("personalised", [False, True])
def test_mixed_hot_slice(gpu, weighted, personalised):
    """The wave engine's FP32 sweep that keeps the most gathered contributions
    in LDS (block tiles off) against the oracle, against the same solve with
    the slice off, and against the block-tile sweep (the default, weighted or
    not): graphs wholly inside the slice (scale 12), larger than it (scale
    16), with n not a multiple of 4 (1003), and a 1 KB slice."""
    cases = [oracle.gen_rmat(12, 16 << 12, 0.57, 0.19, 0.19, 5) + (1 << 12, 0),
             oracle.gen_rmat(16, 16 << 16, 0.57, 0.19, 0.19, 6) + (1 << 16, 0),
             oracle.gen_rmat(16, 16 << 16, 0.57, 0.19, 0.19, 6) + (1 << 16, 1)]
    r = Rng(31)
    n = 1003
    s = np.array([r.below(n) for _ in range(9000)], np.uint32)
    t = np.array([r.below(n) for _ in range(9000)], np.uint32)
    cases.append((s, t, n, 0))
    for s, t, n, kb in cases:
        rng = np.random.default_rng(n + kb)
        w = rng.uniform(0.1, 1.0, len(s)) if weighted else None
        # Normalised: an R-MAT graph has many sinks, whose mass the
        # personalisation hands back each sweep.
        p = rng.uniform(0.0, 2.0, n) if personalised else None
        if p is not None:
            p /= p.sum()
        g = adj(n, s, t)
        dev = g.device()
        want, _ = oracle.Adj(n, s, t).pagerank(weight=w, pers=p, epsilon=1e-12, max_iter=1000,
                                               threads=8)
        got = {}
        for hot, blk in ((kb, -1), (-1, -1), ("blk", 0)):
            dev.set_tuning(sweep_hot_kb=kb if hot == "blk" else hot, block_tile_nnz=blk)
            wp = None if w is None else gt_hip.DenseProp.from_vec(g.graph_id, w)
            pp = None if p is None else gt_hip.DenseProp.from_vec(g.graph_id, p)
            got[hot] = pr(g, 1e-9, 1000, weight=wp, pers=pp, fn=gt_hip.pagerank_mixed)
            assert (dev.layout()["block_tiles"] > 0) == (blk == 0)
        dev.set_tuning(sweep_hot_kb=0, block_tile_nnz=0)
        scale = max(1.0, np.abs(want).sum())
        for hot, (x, it) in got.items():
            assert np.abs(x - want).sum() <= 1e-6 * scale, (n, kb, hot)
        # Only the order of the FP64 partial folds differs between the two.
        assert abs(got[kb][1] - got[-1][1]) <= 1
        assert np.abs(got[kb][0] - got[-1][0]).sum() <= 1e-7 * scale
        # The block sweep's row sums are exact to 2^-61 of the mass.
        assert abs(got["blk"][1] - got[-1][1]) <= 1
        assert np.abs(got["blk"][0] - got[-1][0]).sum() <= 2e-7 * scale


@pytest.mark.parametrize("personalised", [False, True])
def test_fp64_hot_slice(gpu, personalised):
    """The FP64 drop-in on the wave engine (block_f64 = -1) with its hot slice
    (unweighted, unfiltered graphs) and without, and on block tiles (the
    default), against the oracle at the drop-in's bar: graphs inside the
    slice, larger than it, a 1 KB slice, n not a multiple of 4."""
    r = Rng(37)
    n = 1003
    cases = [oracle.gen_rmat(12, 16 << 12, 0.57, 0.19, 0.19, 8) + (1 << 12, 0),
             oracle.gen_rmat(15, 16 << 15, 0.57, 0.19, 0.19, 9) + (1 << 15, 0),
             oracle.gen_rmat(15, 16 << 15, 0.57, 0.19, 0.19, 9) + (1 << 15, 1),
             (np.array([r.below(n) for _ in range(9000)], np.uint32),
              np.array([r.below(n) for _ in range(9000)], np.uint32), n, 0)]
    for s, t, n, kb in cases:
        p = np.random.default_rng(n).uniform(0.0, 2.0, n) if personalised else None
        if p is not None:
            p /= p.sum()
        g = adj(n, s, t)
        dev = g.device()
        want, it_o = oracle.Adj(n, s, t).pagerank(pers=p, epsilon=1e-12, max_iter=1000,
                                                  threads=8)
        for hot, blk in ((kb, -1), (-1, -1), (kb, 1)):
            dev.set_tuning(sweep_hot_kb=hot, block_f64=blk)
            pp = None if p is None else gt_hip.DenseProp.from_vec(g.graph_id, p)
            got, it = pr(g, 1e-12, 1000, pers=pp)
            assert (dev.layout()["block_tiles"] > 0) == (blk == 1)
            parity(got, it, want, it_o, 1e-12)
        dev.set_tuning(sweep_hot_kb=0, block_f64=0)


# ---------------------------------------------------------------- errors

def test_errors(gpu):
    g = adj(4, [(0, 1)])
    with pytest.raises(ValueError, match="plan must cover"):
        gt_hip.pagerank(g, gt_hip.Unity.NEW, 0.85, params(3, 1e-6, 10),
                        gt_hip.DenseProp.new(g.graph_id))
    other = adj(4, [(0, 1)])
    with pytest.raises(ValueError, match="does not belong"):
        gt_hip.pagerank(g, gt_hip.Unity.NEW, 0.85, params(4, 1e-6, 10),
                        gt_hip.DenseProp.new(other.graph_id))
    with pytest.raises(ValueError, match="has no in-edges"):
        gt_hip.pagerank(g.undirect(), gt_hip.Unity.NEW, 0.85, params(4, 1e-6, 10),
                        gt_hip.DenseProp.new(g.graph_id))
    short = gt_hip.DenseProp.from_vec(g.graph_id, np.array([]))
    with pytest.raises(ValueError, match="shorter"):
        gt_hip.pagerank(g, short, 0.85, params(4, 1e-6, 10), gt_hip.DenseProp.new(g.graph_id))


def test_empty_and_edgeless_graphs(gpu):
    g0 = adj(0, [])
    r = gt_hip.DenseProp.new(g0.graph_id)
    assert gt_hip.pagerank(g0, gt_hip.Unity.NEW, 0.85, params(0, 1e-6, 0), r) == 1
    g = adj(5, [])
    got, it = pr(g, 1e-12, 100)
    want, it_o = oracle.Adj(5, np.zeros(0, np.uint32), np.zeros(0, np.uint32)).pagerank(
        epsilon=1e-12, max_iter=100)
    assert it == it_o and np.abs(got - want).max() < 1e-15


# ----------------------------------------------------------- eigenvector

def ev(g, eps, max_iter, weight=None, x=None):
    xm = gt_hip.DenseProp.new(g.graph_id) if x is None else x
    w = gt_hip.Unity.NEW if weight is None else weight
    eig, it = gt_hip.eigenvector(g, w, params(g.vertex_bound(), eps, max_iter), xm)
    return np.asarray(xm.data), eig, it


def test_eigenvector_known_answers(gpu):
    n, s, t, w = sixty()
    g = adj(n, s, t)
    c, eig, it = ev(g, 1e-14, 100_000, weight=gt_hip.DenseProp.from_vec(g.graph_id, w))
    co, eo, ito = oracle.Adj(n, s, t).eigenvector(weight=w, epsilon=1e-14, max_iter=100_000)
    parity(c, it, co, ito, 1e-14)
    assert abs(eig - eo) < 1e-12 * eo
    assert abs(np.sqrt((c * c).sum()) - 1.0) < 1e-12

    cyc = adj(7, [(i, (i + 1) % 7) for i in range(7)])
    c, eig, _ = ev(cyc, 1e-15, 1000)
    assert abs(eig - 1.0) < 1e-12 and np.abs(c - 1 / math.sqrt(7)).max() < 1e-12
    _, eig2, _ = ev(cyc, 1e-15, 1000, weight=gt_hip.Constant(2.5))
    assert abs(eig2 - 2.5) < 1e-12

    empty = adj(5, [])
    c, eig, it = ev(empty, 1e-12, 100)
    assert eig == 0.0 and it >= 1 and (c == 0.0).all()

    path = adj(3, [(0, 1), (1, 2)])
    c, eig, it = ev(path, 1e-12, 1)
    assert it == 1 and abs(eig - math.sqrt(2) / 3) < 1e-15
    h = 1 / math.sqrt(2)
    assert abs(c[0]) < 1e-15 and abs(c[1] - h) < 1e-15 and abs(c[2] - h) < 1e-15


def test_eigenvector_matches_oracle_on_symmetrised_rmat(gpu):
    s, t = oracle.gen_rmat(13, 16 << 13, 0.57, 0.19, 0.19, 5)
    S = np.concatenate([s, t])
    T = np.concatenate([t, s])
    n = 1 << 13
    g = adj(n, S, T)
    for k in (1, 2, 3):
        c, eig, it = ev(g, 1e-30, k)
        co, eo, ito = oracle.Adj(n, S, T).eigenvector(epsilon=1e-30, max_iter=k)
        assert it == ito == k and abs(eig - eo) <= 1e-12 * eo
        assert np.abs(c - co).sum() <= 1e-12
    c, eig, it = ev(g, 1e-9, 1000)
    co, eo, ito = oracle.Adj(n, S, T).eigenvector(epsilon=1e-9, max_iter=1000, threads=8)
    parity(c, it, co, ito, 1e-9)
    assert abs(eig - eo) <= 1e-9 * eo


# ------------------------------------------------------------ filtered views

def _masks(n, m, seed):
    r = Rng(seed)
    vmask = np.array([0 if r.below(7) == 3 else 1 for _ in range(n)], np.uint8)
    emask = np.array([0 if r.below(5) == 1 else 1 for _ in range(m)], np.uint8)
    return vmask, emask


def test_filtered_views(gpu):
    n, s, t = thousand()
    m = len(s)
    vmask, emask = _masks(n, m, 31)
    g = adj(n, s, t)
    o = oracle.Adj(n, s, t)
    w = np.array([0.25 + (i % 11) / 5.0 for i in range(m)])
    start = np.linspace(0.1, 0.9, n) / n
    for rev in (False, True):
        view = g.masked(vmask, emask, reversed=rev)
        got, it = pr(view, 1e-12, 1000)
        want, it_o = o.pagerank(vmask=vmask, emask=emask, reversed=rev, epsilon=1e-12,
                                max_iter=1000)
        parity(got, it, want, it_o, 1e-12)
        rank = gt_hip.DenseProp.from_vec(g.graph_id, start.copy())
        got, it = pr(view, 1e-12, 1000, weight=gt_hip.DenseProp.from_vec(g.graph_id, w),
                     rank=rank)
        want, it_o = o.pagerank(weight=w, vmask=vmask, emask=emask, reversed=rev,
                                epsilon=1e-12, max_iter=1000, rank=start)
        parity(got, it, want, it_o, 1e-12)
        dropped = np.flatnonzero(vmask == 0)
        assert (got[dropped] == start[dropped]).all()
        x = gt_hip.DenseProp.new(g.graph_id)
        eig, it = gt_hip.eigenvector(view, gt_hip.Unity.NEW, params(n, 1e-12, 1000), x)
        xo, eo, ito = o.eigenvector(vmask=vmask, emask=emask, reversed=rev, epsilon=1e-12,
                                    max_iter=1000)
        parity(x.data, it, xo, ito, 1e-12)
        assert abs(eig - eo) <= 1e-12 * max(1.0, abs(eo))


def test_short_masks_are_rejected(gpu):
    g = adj(4, [(0, 1), (1, 2)])
    with pytest.raises(ValueError, match="filter"):
        g.masked(np.ones(3, np.uint8), np.ones(2, np.uint8))
