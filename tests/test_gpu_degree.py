"""Degree kernels on the device against crates/gt-algo/tests/u15_degree.rs
(its six-vertex fixture and assertions) and against numpy on random graphs
and views."""
import numpy as np
import pytest

import gt_hip

pytestmark = pytest.mark.gpu

# u15_degree.rs:100-116: vertex 0 has out-degree ten (parallel edges, a
# self-loop) and one edge back into it.
EDGES = [(0, 1), (0, 1), (0, 2), (0, 2), (0, 3), (0, 3), (0, 4), (0, 4), (0, 5), (0, 0),
         (1, 2), (2, 0)]
N = 6


def build():
    g = gt_hip.AdjList.with_vertices(N)
    for s, t in EDGES:
        g.add_edge(s, t)
    return g


W = np.arange(1.0, len(EDGES) + 1.0)  # exact partial sums in any order


def ref_out(u):
    return sum(W[i] for i, (s, _) in enumerate(EDGES) if s == u)


def ref_in(u):
    return sum(W[i] for i, (_, t) in enumerate(EDGES) if t == u)


def test_unity_is_the_plain_degree(gpu):
    g = build()
    assert gt_hip.weighted_out_degrees(g).tolist() == [10, 1, 1, 0, 0, 0]
    assert gt_hip.weighted_in_degrees(g).tolist() == [2, 2, 3, 2, 2, 1]
    assert gt_hip.weighted_degrees(g).tolist() == [12, 3, 4, 2, 2, 1]


def test_constant_multiplies_rather_than_accumulates(gpu):
    g = build()
    acc = 0.0
    for _ in range(10):
        acc += 0.1
    assert acc != 1.0
    c = gt_hip.Constant(0.1)
    assert gt_hip.weighted_out_degree(g, 0, c) == 1.0 == 10 * 0.1
    assert gt_hip.weighted_degree(g, 0, gt_hip.Constant(2.0)) == 24.0
    od = gt_hip.weighted_out_degrees(g, c)
    assert od.tolist() == [d * 0.1 for d in [10, 1, 1, 0, 0, 0]]
    # no incident edge: 0, even for a constant whose product with 0 is NaN
    assert gt_hip.weighted_out_degree(g, 5, gt_hip.Constant(float("inf"))) == 0.0


def test_dense_map_sums_the_neighbourhood(gpu):
    g = build()
    wp = gt_hip.DenseProp.from_vec(g.graph_id, W)
    od, idg, td = (gt_hip.weighted_out_degrees(g, wp), gt_hip.weighted_in_degrees(g, wp),
                   gt_hip.weighted_degrees(g, wp))
    for u in range(N):
        assert od[u] == ref_out(u) and idg[u] == ref_in(u) and td[u] == ref_out(u) + ref_in(u)
    assert td.sum() == 2 * W.sum()
    assert od[5] == 0.0 and idg[5] == ref_in(5)


def test_views(gpu):
    g = build()
    wp = gt_hip.DenseProp.from_vec(g.graph_id, W)
    r = g.reversed()
    assert np.array_equal(gt_hip.weighted_out_degrees(r, wp), gt_hip.weighted_in_degrees(g, wp))
    assert np.array_equal(gt_hip.weighted_in_degrees(r, wp), gt_hip.weighted_out_degrees(g, wp))
    assert np.array_equal(gt_hip.weighted_degrees(r, wp), gt_hip.weighted_degrees(g, wp))
    u = g.undirect()
    assert np.array_equal(gt_hip.weighted_out_degrees(u, wp), gt_hip.weighted_degrees(g, wp))
    assert np.array_equal(gt_hip.weighted_degrees(u, wp), gt_hip.weighted_degrees(g, wp))
    assert gt_hip.weighted_out_degree(u, 0) == 12.0 and gt_hip.weighted_out_degree(g, 0) == 10.0
    with pytest.raises(ValueError, match="Undirected"):
        gt_hip.weighted_in_degrees(u, wp)


def test_out_degrees_and_histogram(gpu):
    g = build()
    p = gt_hip.DenseProp(g.graph_id, np.full(3, 99, np.int64), dtype=np.int64)
    assert gt_hip.out_degrees(g, p).data.tolist() == [10, 1, 1, 0, 0, 0]
    assert gt_hip.out_degrees(g.undirect()).data.tolist() == [12, 3, 4, 2, 2, 1]
    h = gt_hip.degree_histogram(g)
    want = [0] * 13
    for d in [12, 3, 4, 2, 2, 1]:
        want[d] += 1
    assert h == want and len(h) == 13 and sum(h) == N
    assert gt_hip.degree_histogram(g.undirect()) == h
    e = gt_hip.AdjList.with_vertices(4)
    assert gt_hip.degree_histogram(e) == [4]
    assert gt_hip.degree_histogram(gt_hip.AdjList.with_vertices(0)) == []


def test_random_graphs_and_filtered_views(gpu):
    rng = np.random.default_rng(7)
    n, m = 3000, 40000
    # a skewed source distribution gives rows past the wavefront and workgroup thresholds
    src = np.minimum((rng.pareto(1.1, m) * 3).astype(np.int64), n - 1).astype(np.uint32)
    tgt = rng.integers(0, n, m).astype(np.uint32)
    w = rng.random(m) + 0.5
    g = gt_hip.AdjList.with_vertices(n)
    g.add_edges(src, tgt)
    wp = gt_hip.DenseProp.from_vec(g.graph_id, w)
    out_w = np.bincount(src, weights=w, minlength=n)
    in_w = np.bincount(tgt, weights=w, minlength=n)
    assert np.allclose(gt_hip.weighted_out_degrees(g, wp), out_w, rtol=1e-13, atol=0)
    assert np.allclose(gt_hip.weighted_in_degrees(g, wp), in_w, rtol=1e-13, atol=0)
    assert np.allclose(gt_hip.weighted_degrees(g, wp), out_w + in_w, rtol=1e-13, atol=0)
    vm = (rng.random(n) < 0.7).astype(np.uint8)
    em = (rng.random(m) < 0.8).astype(np.uint8)
    f = g.masked(vm, em)
    keep = (em == 1) & (vm[src] == 1) & (vm[tgt] == 1)
    fo = np.bincount(src[keep], weights=w[keep], minlength=n) * vm
    fi = np.bincount(tgt[keep], weights=w[keep], minlength=n) * vm
    assert np.allclose(gt_hip.weighted_out_degrees(f, wp), fo, rtol=1e-13, atol=0)
    assert np.allclose(gt_hip.weighted_in_degrees(f, wp), fi, rtol=1e-13, atol=0)
    deg = (np.bincount(src[keep], minlength=n) + np.bincount(tgt[keep], minlength=n))
    hist = np.bincount(deg[vm == 1])
    assert gt_hip.degree_histogram(f) == hist.tolist()
    assert gt_hip.out_degrees(f).data.tolist() == (np.bincount(src[keep], minlength=n) * vm).tolist()


def test_weighted_out_degree_follows_storage_order_at_every_row_length(gpu):
    """Rows of 1 to 20,000 out-entries with weights whose sum depends on the
    order: up to 4,096 entries the device sum is the reference's sequential
    sum in storage order (centrality.rs:281-297) bit for bit; longer rows use
    a fixed tree order, within rounding of it."""
    import oracle
    lens = [1, 7, 31, 32, 33, 63, 64, 65, 1000, 4095, 4096, 4097, 9000, 20000]
    n = len(lens) + 50
    rng = np.random.default_rng(11)
    s = np.concatenate([np.full(L, v, np.uint32) for v, L in enumerate(lens)])
    t = rng.integers(0, n, len(s)).astype(np.uint32)
    w = rng.uniform(0.0, 1.0, len(s)) * 10.0 ** rng.integers(-8, 8, len(s))
    g = gt_hip.AdjList.with_vertices(n)
    g.add_edges(s, t)
    got = gt_hip.weighted_out_degrees(g, gt_hip.DenseProp.from_vec(g.graph_id, w))
    _, _, _, out_off, _, out_eid = oracle.Adj(n, s, t).export()
    for v, L in enumerate(lens):
        acc = 0.0
        for e in out_eid[out_off[v]:out_off[v + 1]]:
            acc += float(w[e])
        if L <= 4096:
            assert got[v] == acc, (L, got[v], acc)
        else:
            assert abs(got[v] - acc) <= 1e-12 * acc, (L, got[v], acc)
    assert (got[len(lens):] == 0.0).all()
