"""GPU device builder against the adjacency add_edge builds: every array word
for word (integer work, so the bar is bit-exact), on the reference's LAYOUT
fixture, random multigraphs with self-loops, hub-heavy R-MAT graphs, and the
empty cases."""
import numpy as np
import pytest

import gt_hip
import oracle
from graphs import EDGES, LAYOUT, Rng, thousand

pytestmark = pytest.mark.gpu


def check_build(n, s, t):
    s = np.ascontiguousarray(s, dtype=np.uint32)
    t = np.ascontiguousarray(t, dtype=np.uint32)
    g = gt_hip.DeviceGraph.build(n, s, t)
    got = g.download()
    want = oracle.Adj(n, s, t).export() if n else None
    if want is None:
        assert int(got[0][0]) == 0 and int(got[3][0]) == 0
        return got
    names = ["in_off", "in_other", "in_eid", "out_off", "out_other", "out_eid"]
    for name, x, y in zip(names, got, want):
        assert x.shape == y.shape, name
        bad = np.flatnonzero(x != y)
        assert bad.size == 0, f"{name}: first mismatch at {bad[:5]}"
    return got


def test_layout_fixture(gpu):
    s = [e[0] for e in EDGES]
    t = [e[1] for e in EDGES]
    in_off, in_o, in_e, out_off, out_o, out_e = check_build(4, s, t)
    for v, (entries, pos) in enumerate(LAYOUT):
        ins = list(zip(in_o[in_off[v]:in_off[v + 1]].tolist(), in_e[in_off[v]:in_off[v + 1]].tolist()))
        assert ins == entries[pos:]


def test_random_multigraphs_with_self_loops(gpu):
    r = Rng(41)
    for n, m in [(1, 1), (1, 50), (3, 100), (64, 5000), (1000, 40_000)]:
        s = [r.below(n) for _ in range(m)]
        t = [r.below(n) for _ in range(m)]
        check_build(n, s, t)


def test_thousand(gpu):
    check_build(*thousand())


def test_rmat_hubs(gpu):
    for scale in (12, 16):
        s, t = oracle.gen_rmat(scale, 16 << scale, 0.57, 0.19, 0.19, scale)
        check_build(1 << scale, s, t)


def test_many_vertices_few_edges_and_wide_ids(gpu):
    # n needs all radix passes: 2^24 + 5 vertices, endpoints spread over the range.
    n = (1 << 24) + 5
    r = Rng(8)
    s = [r.below(n) for _ in range(20_000)] + [n - 1] * 10
    t = [r.below(n) for _ in range(20_000)] + [0] * 10
    check_build(n, s, t)


def test_empty_graphs(gpu):
    check_build(0, [], [])
    g = gt_hip.DeviceGraph.build(7, np.zeros(0, np.uint32), np.zeros(0, np.uint32))
    in_off, *_ = g.download()
    assert (in_off == 0).all() and in_off.shape == (8,)


def test_no_such_vertex(gpu):
    with pytest.raises(ValueError, match="no such vertex"):
        gt_hip.DeviceGraph.build(3, np.array([0, 3], np.uint32), np.array([1, 1], np.uint32))


def test_device_resident_edge_list(gpu):
    import torch
    s, t = oracle.gen_rmat(14, 16 << 14, 0.57, 0.19, 0.19, 77)
    g = gt_hip.DeviceGraph.build(1 << 14, torch.from_numpy(s.astype(np.int32)).cuda(),
                                 torch.from_numpy(t.astype(np.int32)).cuda())
    got = g.download()
    want = oracle.Adj(1 << 14, s, t).export()
    for x, y in zip(got, want):
        assert (x == y).all()


def test_pagerank_on_device_built_graph(gpu):
    n, s, t = thousand()
    g = gt_hip.DeviceGraph.build(n, s, t)
    rank = gt_hip.DenseProp.new(g.graph_id)
    it = gt_hip.pagerank(g, gt_hip.Unity.NEW, 0.85,
                         gt_hip.PowerIteration(1e-12, 1000, gt_hip.Plan(n, 64)), rank)
    want, it_o = oracle.Adj(n, s, t).pagerank(epsilon=1e-12, max_iter=1000)
    assert it == it_o
    assert np.abs(rank.data - want).sum() <= 1e-12
