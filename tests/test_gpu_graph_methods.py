"""The gt-py-shaped methods Graph.pagerank / Graph.eigenvector on a graph and
on each view kind, against the oracle: maps sized first, results in the given
map or the graph's own vp map, ValueError for a map of another graph or an
undirected view (the NotDirected message, crates/gt-py/src/dispatch.rs:176-179)."""
import numpy as np
import pytest

import gt_hip
import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def graph():
    n = 1 << 13
    s, t = oracle.gen_rmat(13, 16 * n, 0.57, 0.19, 0.19, 31)
    g = gt_hip.AdjList.with_vertices(n)
    g.add_edges(s, t)
    rng = np.random.default_rng(3)
    vm = (rng.random(n) < 0.8).astype(np.uint8)
    em = (rng.random(len(s)) < 0.9).astype(np.uint8)
    return n, s, t, g, oracle.Adj(n, s, t), vm, em


def _views(g, vm, em):
    return {"directed": (g, {}), "reversed": (g.reversed(), {"reversed": True}),
            "directed_filtered": (g.masked(vm, em), {"vmask": vm, "emask": em}),
            "reversed_filtered": (g.masked(vm, em, reversed=True),
                                  {"vmask": vm, "emask": em, "reversed": True})}


@pytest.mark.parametrize("kind", ["directed", "reversed", "directed_filtered",
                                  "reversed_filtered"])
def test_pagerank_and_eigenvector_methods(gpu, graph, kind):
    n, s, t, g, o, vm, em = graph
    view, kw = _views(g, vm, em)[kind]
    w = g.new_edge_property("double")
    w.data = np.random.default_rng(5).uniform(0.1, 1.0, len(s))
    rank = g.new_vertex_property("double")
    it = view.pagerank(weight=w, epsilon=1e-11, rank=rank)
    want, it_o = o.pagerank(weight=w.data, epsilon=1e-11, max_iter=0, **kw)
    assert abs(it - it_o) <= 1
    assert np.abs(rank.data - want).sum() <= (1e-12 if it == it_o else 1e-11)
    x = g.new_vertex_property("double")
    eig, ite = view.eigenvector(epsilon=1e-10, max_iter=500, x=x)
    xo, eo, ito = o.eigenvector(epsilon=1e-10, max_iter=500, **kw)
    assert abs(ite - ito) <= 1
    assert abs(eig - eo) <= 1e-9 * abs(eo)
    assert np.abs(x.data - xo).sum() <= (1e-12 if ite == ito else 1e-9) * max(
        1.0, np.abs(xo).sum())


def test_results_in_the_graphs_own_maps_and_warm_start(gpu, graph):
    n, s, t, g, o, vm, em = graph
    pers = g.new_vertex_property()
    pers.data = np.full(100, 2.0 / n)  # shorter than the bound: sized first
    it = g.pagerank(pers=pers, epsilon=1e-11)
    assert len(pers) == n and pers.data[100:].sum() == 0.0
    p = np.zeros(n)
    p[:100] = 2.0 / n
    want, it_o = o.pagerank(pers=p, epsilon=1e-11)
    got = g.vp["pagerank"].data
    assert abs(it - it_o) <= 1 and np.abs(got - want).sum() <= 1e-11
    # A second call starts from the stored ranks: one sweep suffices.
    assert g.pagerank(pers=pers, epsilon=1e-9) <= 2
    eig, _ = g.eigenvector(epsilon=1e-300, max_iter=10)
    assert g.vp["eigenvector"].data.shape == (n,) and eig > 0


def test_refusals(gpu, graph):
    n, s, t, g, o, vm, em = graph
    other = gt_hip.AdjList.with_vertices(4)
    with pytest.raises(ValueError, match="rank map does not belong to this graph"):
        g.pagerank(rank=other.new_vertex_property())
    with pytest.raises(ValueError, match="Undirected has no in-edges"):
        g.undirect().pagerank()
    with pytest.raises(ValueError, match="UndirectedFiltered has no in-edges"):
        g.masked(vm, em, undirected=True).eigenvector()
    with pytest.raises(ValueError, match="Invalid property type"):
        g.new_vertex_property("vector<string>")
