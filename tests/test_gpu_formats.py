"""Files into device graphs: the device build of a parsed file equals the
host adjacency, and lesmis (weighted, stored directed) is a real-graph
PageRank parity input against the oracle."""
from pathlib import Path

import numpy as np
import pytest

import gt_hip
import oracle

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"


@pytest.mark.parametrize("name", ["karate.gt", "lesmis.gt"])
def test_device_build_of_a_file_equals_the_host_adjacency(gpu, name):
    doc = gt_hip.read_gt(GOLD / name)
    dev = doc.device()
    s, t = doc.edges()
    want = gt_hip.host_adjacency(doc.n, s, t)
    got = dev.download()
    for a, b in zip(got, want):
        assert np.array_equal(np.asarray(a), np.asarray(b))


def test_lesmis_weighted_pagerank_matches_the_oracle(gpu):
    doc = gt_hip.read_gt(GOLD / "lesmis.gt")
    w = doc.properties()[4]["values"]
    g = doc.adjlist()
    s, t = doc.edges()
    o = oracle.Adj(doc.n, s, t)
    for d in (0.85, 0.5):
        want, it_o = o.pagerank(weight=w, damping=d, epsilon=1e-12, grain=8)
        r = gt_hip.DenseProp.new(g.graph_id)
        it = gt_hip.pagerank(g, gt_hip.DenseProp.from_vec(g.graph_id, w), d,
                             gt_hip.PowerIteration(1e-12, 0, gt_hip.Plan(doc.n, 8)), r)
        assert abs(it - it_o) <= 1
        assert np.abs(r.data - want).sum() <= 1e-12
        # the reversed view too
        want_r, it_r = o.pagerank(weight=w, damping=d, epsilon=1e-12, grain=8, reversed=True)
        r2 = gt_hip.DenseProp.new(g.graph_id)
        it2 = gt_hip.pagerank(g.reversed(), gt_hip.DenseProp.from_vec(g.graph_id, w), d,
                              gt_hip.PowerIteration(1e-12, 0, gt_hip.Plan(doc.n, 8)), r2)
        assert abs(it2 - it_r) <= 1 and np.abs(r2.data - want_r).sum() <= 1e-12


def test_edge_list_to_device_pagerank(gpu):
    rng = np.random.default_rng(11)
    pairs = rng.integers(0, 300, size=(3000, 2))
    text = "\n".join(f"v{a},v{b}" for a, b in pairs).encode()
    doc = gt_hip.read_edge_list(text)
    g = doc.adjlist()
    s, t = doc.edges()
    want, it_o = oracle.Adj(doc.n, s, t).pagerank(epsilon=1e-12, grain=16)
    r = gt_hip.DenseProp.new(g.graph_id)
    it = gt_hip.pagerank(g, gt_hip.Unity.NEW, 0.85,
                         gt_hip.PowerIteration(1e-12, 0, gt_hip.Plan(doc.n, 16)), r)
    assert abs(it - it_o) <= 1 and np.abs(r.data - want).sum() <= 1e-12
