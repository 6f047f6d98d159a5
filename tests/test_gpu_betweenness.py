"""Betweenness on the device against the oracle's Brandes (oracle_betweenness).

Cases: the reference's 100 seeded multigraphs (u16_centrality.rs:707-770),
its path closed form (:772-799), relabelling (:801-831), map overwrite
(:833-844) and unreachable (:870-887) cases; every view kind; rows long enough
to be cut into pieces (hubs in both directions); several batches of 64
sources with a ragged last one; deep BFS trees; duplicate and restricted
pivots; device maps; the error paths.  The device sums dependencies in a
different order from the reference, so results agree to rounding: relative
1e-12 (absolute 1e-10 for values near zero), and exactly where every term is
an integer.
"""
import numpy as np
import pytest

import gt_hip
import oracle as O
from graphs import brandes_cases

pytestmark = pytest.mark.gpu

RTOL, ATOL = 1e-12, 1e-10


def adj(n, edges):
    g = gt_hip.AdjList.with_vertices(n)
    if edges:
        g.add_edges(np.array([e[0] for e in edges], np.uint32),
                    np.array([e[1] for e in edges], np.uint32))
    return g


def device_bc(view, g, pivots=None, edges=True):
    vb = gt_hip.DenseProp.new(g.graph_id)
    eb = gt_hip.DenseProp.new(g.graph_id) if edges else None
    st = gt_hip.betweenness(view, vb, eb, pivots=pivots)
    return vb.data, (eb.data if edges else None), st


def oracle_bc(n, edges, **kw):
    src = [e[0] for e in edges]
    tgt = [e[1] for e in edges]
    return O.betweenness(O.Adj(n, src, tgt), **kw)


def check(got, want):
    np.testing.assert_allclose(got[0], want[0], rtol=RTOL, atol=ATOL)
    if got[1] is not None:
        np.testing.assert_allclose(got[1], want[1], rtol=RTOL, atol=ATOL)


def test_reference_cases(gpu):
    for k, (n, edges) in enumerate(brandes_cases()):
        g = adj(n, edges)
        for undirected in (False, True):
            view = g.undirect() if undirected else g
            gv, ge, _ = device_bc(view, g)
            wv, we = oracle_bc(n, edges, undirected=undirected)
            np.testing.assert_allclose(gv, wv, rtol=RTOL, atol=ATOL, err_msg=f"case {k}")
            np.testing.assert_allclose(ge, we, rtol=RTOL, atol=ATOL, err_msg=f"case {k}")


@pytest.mark.parametrize("undirected", [False, True])
def test_path_closed_form(gpu, undirected):
    N = 9
    g = adj(N, [(i, i + 1) for i in range(N - 1)])
    gv, ge, _ = device_bc(g.undirect() if undirected else g, g)
    assert gv.tolist() == [float(i * (N - 1 - i)) for i in range(N)]
    assert ge.tolist() == [float((i + 1) * (N - 1 - i)) for i in range(N - 1)]


def test_relabelling_and_map_overwrite(gpu):
    edges = [(0, 1), (1, 2), (2, 3), (3, 0), (0, 2)]
    g = adj(4, edges)
    vb = gt_hip.DenseProp.new(g.graph_id)
    eb = gt_hip.DenseProp.new(g.graph_id)
    gt_hip.betweenness(g, vb, eb)
    first = (vb.data.copy(), eb.data.copy())
    gt_hip.betweenness(g, vb, eb)  # zero-filled again, not accumulated
    assert vb.data.tobytes() == first[0].tobytes()
    assert eb.data.tobytes() == first[1].tobytes()
    check(first, oracle_bc(4, edges))
    h = adj(4, [(3 - s, 3 - t) for s, t in edges])
    hv, _, _ = device_bc(h, h)
    np.testing.assert_allclose(hv[::-1], first[0], rtol=RTOL, atol=ATOL)


def test_unreachable_vertices_contribute_nothing(gpu):
    g = adj(6, [(0, 1), (1, 2), (3, 4)])
    gv, ge, _ = device_bc(g.undirect(), g)
    assert np.all(np.isfinite(gv))
    assert gv[1] == 1.0 and gv[0] == 0.0 and gv[5] == 0.0
    assert ge[2] == 1.0


def _rmat(scale, ef, seed):
    n = 1 << scale
    src, tgt = gt_hip.gen_rmat(scale, ef * n, 0.57, 0.19, 0.19, seed)
    return n, list(zip(src.tolist(), tgt.tolist()))


@pytest.mark.parametrize("kind", ["directed", "reversed", "undirected"])
def test_rmat_all_pivots(gpu, kind):
    """R-MAT scale 10 (hubs of a few hundred entries in both directions, cut
    into pieces on the undirected view), every vertex a pivot: 16 batches."""
    n, edges = _rmat(10, 8, 7)
    g = adj(n, edges)
    view = {"directed": g, "reversed": g.reversed(), "undirected": g.undirect()}[kind]
    gv, ge, st = device_bc(view, g)
    wv, we = oracle_bc(n, edges, reversed=kind == "reversed", undirected=kind == "undirected",
                       threads=8)
    check((gv, ge), (wv, we))
    assert st["batches"] == -(-st["sources"] // 64)
    # The 16 batches alternate between two streams; the lanes' sums are
    # added in a fixed order, so a second call is bit-identical.
    assert st["streams"] == 2
    gv2, ge2, _ = device_bc(view, g)
    assert np.array_equal(gv, gv2) and np.array_equal(ge, ge2)


def test_hubs_are_cut_into_pieces(gpu):
    """A vertex with 3000 out- and 3000 in-edges (rows of six pieces) among a
    sparse ring: both phases go through the piece fix-up kernels."""
    n = 3200
    rng = np.random.default_rng(4)
    edges = [(i, (i + 1) % n) for i in range(n)]
    edges += [(0, int(v)) for v in rng.integers(1, n, 3000)]
    edges += [(int(v), 0) for v in rng.integers(1, n, 3000)]
    g = adj(n, edges)
    piv = np.arange(0, n, 7, dtype=np.uint32)  # 458 sources: 8 batches, ragged
    for view, undirected in ((g, False), (g.undirect(), True)):
        gv, ge, _ = device_bc(view, g, pivots=piv)
        check((gv, ge), oracle_bc(n, edges, undirected=undirected, pivots=piv, threads=8))


@pytest.mark.parametrize("undirected", [False, True])
def test_deep_trees(gpu, undirected):
    """A 700-vertex path plus a few chords: hundreds of BFS levels per batch."""
    n = 700
    edges = [(i, i + 1) for i in range(n - 1)] + [(5, 300), (400, 20), (650, 651)]
    g = adj(n, edges)
    gv, ge, st = device_bc(g.undirect() if undirected else g, g)
    check((gv, ge), oracle_bc(n, edges, undirected=undirected))
    assert st["levels"] > 1000


@pytest.mark.parametrize("kind", ["directed", "reversed", "undirected"])
def test_filtered_views(gpu, kind):
    n, edges = _rmat(9, 6, 21)
    rng = np.random.default_rng(17)
    vmask = (rng.random(n) < 0.85).astype(np.uint8)
    emask = (rng.random(len(edges)) < 0.8).astype(np.uint8)
    g = adj(n, edges)
    view = g.masked(vmask, emask, reversed=kind == "reversed", undirected=kind == "undirected")
    gv, ge, _ = device_bc(view, g)
    want = oracle_bc(n, edges, reversed=kind == "reversed", undirected=kind == "undirected",
                     vmask=vmask, emask=emask, threads=8)
    check((gv, ge), want)
    assert np.all(gv[vmask == 0] == 0.0) and np.all(ge[emask == 0] == 0.0)


def test_pivots_duplicates_and_vertex_only(gpu):
    n, edges = _rmat(8, 8, 3)
    g = adj(n, edges)
    piv = np.array([3, 3, 200, 17, 3], np.uint32)
    gv, ge, st = device_bc(g, g, pivots=piv)
    check((gv, ge), oracle_bc(n, edges, pivots=piv))
    hv, he, _ = device_bc(g, g, pivots=piv, edges=False)
    assert he is None
    assert hv.tobytes() == gv.tobytes()


def test_device_maps(gpu):
    import torch
    n, edges = _rmat(8, 8, 5)
    g = adj(n, edges)
    vb = gt_hip.DenseProp.from_vec(g.graph_id, torch.zeros(n, dtype=torch.float64,
                                                           device="cuda"))
    eb = gt_hip.DenseProp.from_vec(g.graph_id, torch.zeros(len(edges), dtype=torch.float64,
                                                           device="cuda"))
    gt_hip.betweenness(g, vb, eb)
    torch.cuda.synchronize()
    check((vb.data.cpu().numpy(), eb.data.cpu().numpy()), oracle_bc(n, edges))


def test_errors(gpu):
    g = adj(5, [(0, 1), (1, 2)])
    other = adj(5, [(0, 1)])
    with pytest.raises(ValueError, match="vertex map does not belong to this graph"):
        gt_hip.betweenness(g, gt_hip.DenseProp.new(other.graph_id),
                           gt_hip.DenseProp.new(g.graph_id))
    with pytest.raises(ValueError, match="edge map does not belong to this graph"):
        gt_hip.betweenness(g, gt_hip.DenseProp.new(g.graph_id),
                           gt_hip.DenseProp.new(other.graph_id))
    with pytest.raises(ValueError, match="pivot out of range"):
        device_bc(g, g, pivots=[5])
    view = g.masked(np.array([1, 1, 0, 1, 1], np.uint8), np.ones(2, np.uint8))
    with pytest.raises(ValueError, match="not a vertex of the view"):
        device_bc(view, g, pivots=[2])


def test_empty_graph(gpu):
    g = adj(3, [])
    gv, ge, st = device_bc(g, g)
    assert gv.tolist() == [0.0, 0.0, 0.0] and len(ge) == 0 and st["sources"] == 0
