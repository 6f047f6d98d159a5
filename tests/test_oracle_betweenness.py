"""Pins the oracle's Brandes betweenness (oracle_betweenness) on the CPU.

Checks, each chosen so that a plausible slip (a dropped "1 +", sigma[u] for
sigma[v], the source counted in its own dependency, a self-loop entering the
DAG, parallel edges merged, the undirected halving missed) fails one of them:

* the reference's own 100 seeded graphs (u16_centrality.rs:707-770: SplitMix64
  seed 0x0123456789abcdef, n = 3 + below(10), m = below(3n + 1), self-loops
  and parallel edges included) against an independent restatement of its
  reference_brandes helper (:628-690), directed and undirected, at its 1e-9;
* the path closed form (:772-799): vertex i * (n-1-i), edge (i+1) * (n-1-i);
* the relabelling case (:801-831) and the unreachable case (:870-887);
* bit-identical results at 1, 4 and 16 threads (:846-868);
* brute force from the definition: all-pairs BFS distances and path counts,
  sigma_st(v) = sigma_sv * sigma_vt when d(s,v) + d(v,t) = d(s,t);
* views by construction: reversed = the graph with every edge turned round,
  filtered = the graph with the dropped vertices' and edges' incidences gone.
"""
from __future__ import annotations

from collections import deque

import numpy as np
import pytest

import oracle as O
from graphs import Rng, brandes_cases


def brandes_lists(n, edges, directed, plant=None):
    """Brandes over plain adjacency lists built in edge order (each undirected
    non-loop edge listed at both endpoints), the reference test's helper.
    ``plant`` injects a known mistake, to show the checks catch it."""
    adj = [[] for _ in range(n)]
    for ei, (s, t) in enumerate(edges):
        adj[s].append((t, ei))
        if not directed and s != t:
            adj[t].append((s, ei))
    vb = [0.0] * n
    eb = [0.0] * len(edges)
    for s in range(n):
        dist = [-1] * n
        sigma = [0.0] * n
        dep = [0.0] * n
        pred = [[] for _ in range(n)]
        order = [s]
        dist[s] = 0
        sigma[s] = 1.0
        head = 0
        while head < len(order):
            u = order[head]
            head += 1
            for w, ei in adj[u]:
                if w == u:
                    continue
                if dist[w] < 0:
                    dist[w] = dist[u] + 1
                    sigma[w] = sigma[u]
                    pred[w].append((u, ei))
                    order.append(w)
                elif dist[w] == dist[u] + 1:
                    sigma[w] += sigma[u]
                    pred[w].append((u, ei))
        for u in reversed(order):
            for p, ei in pred[u]:
                one = 0.0 if plant == "no_one" else 1.0
                f = (sigma[p] / sigma[u]) * (one + dep[u])
                dep[p] += f
                eb[ei] += f
            if u != s or plant == "count_source":
                vb[u] += dep[u]
    if not directed and plant != "no_halving":
        vb = [x / 2.0 for x in vb]
        eb = [x / 2.0 for x in eb]
    return np.array(vb), np.array(eb)


def definition(n, edges, directed):
    """Betweenness straight from its definition: for every ordered pair (s, t)
    and every v (or edge) on a shortest s-t path, add sigma_s(v) sigma_v(t) /
    sigma_s(t); undirected pairs are counted once."""
    arcs = []  # (a, b, edge id) usable a -> b
    for ei, (s, t) in enumerate(edges):
        if s == t:
            continue
        arcs.append((s, t, ei))
        if not directed:
            arcs.append((t, s, ei))
    out = [[] for _ in range(n)]
    for a, b, _ in arcs:
        out[a].append(b)
    D = np.full((n, n), -1, np.int64)
    S = np.zeros((n, n))
    for s in range(n):
        D[s, s], S[s, s] = 0, 1.0
        q = deque([s])
        while q:
            u = q.popleft()
            for w in out[u]:
                if D[s, w] < 0:
                    D[s, w] = D[s, u] + 1
                    q.append(w)
                if D[s, w] == D[s, u] + 1:
                    S[s, w] += S[s, u]
    vb = np.zeros(n)
    eb = np.zeros(len(edges))
    for s in range(n):
        for t in range(n):
            if t == s or D[s, t] < 0:
                continue
            for v in range(n):
                if v in (s, t) or D[s, v] < 0 or D[v, t] < 0:
                    continue
                if D[s, v] + D[v, t] == D[s, t]:
                    vb[v] += S[s, v] * S[v, t] / S[s, t]
            for a, b, ei in arcs:
                if D[s, a] >= 0 and D[b, t] >= 0 and D[s, a] + 1 + D[b, t] == D[s, t]:
                    eb[ei] += S[s, a] * S[b, t] / S[s, t]
    if not directed:
        vb /= 2.0
        eb /= 2.0
    return vb, eb


def run(n, edges, **kw):
    src = [e[0] for e in edges]
    tgt = [e[1] for e in edges]
    return O.betweenness(O.Adj(n, src, tgt), **kw)


def reference_cases():
    return brandes_cases()


def test_reference_cases_match_brandes_over_lists():
    cases = reference_cases()
    assert len(cases) == 100
    assert any(s == t for _, es in cases for s, t in es), "no self-loop drawn"
    assert any(len(set(es)) < len(es) for _, es in cases), "no parallel edge drawn"
    for k, (n, edges) in enumerate(cases):
        for directed in (True, False):
            gv, ge = run(n, edges, undirected=not directed)
            wv, we = brandes_lists(n, edges, directed)
            np.testing.assert_allclose(gv, wv, rtol=0, atol=1e-9, err_msg=f"case {k} {directed}")
            np.testing.assert_allclose(ge, we, rtol=0, atol=1e-9, err_msg=f"case {k} {directed}")


def test_reference_cases_match_the_definition():
    for k, (n, edges) in enumerate(reference_cases()[:40]):
        for directed in (True, False):
            gv, ge = run(n, edges, undirected=not directed)
            wv, we = definition(n, edges, directed)
            np.testing.assert_allclose(gv, wv, rtol=1e-12, atol=1e-12, err_msg=f"case {k}")
            np.testing.assert_allclose(ge, we, rtol=1e-12, atol=1e-12, err_msg=f"case {k}")


@pytest.mark.parametrize("plant", ["no_one", "count_source", "no_halving"])
def test_planted_mistakes_are_caught_by_the_definition(plant):
    """The definition check is sharp: Brandes with a dropped '1 +', with the
    source counted in its own dependency, or without the undirected halving
    disagrees with it on the first reference cases."""
    caught = False
    for n, edges in reference_cases()[:10]:
        for directed in (True, False):
            wv, we = definition(n, edges, directed)
            bv, be = brandes_lists(n, edges, directed, plant=plant)
            if not (np.allclose(bv, wv, atol=1e-9) and np.allclose(be, we, atol=1e-9)):
                caught = True
    assert caught


@pytest.mark.parametrize("undirected", [False, True])
def test_path_closed_form(undirected):
    N = 9
    edges = [(i, i + 1) for i in range(N - 1)]
    gv, ge = run(N, edges, undirected=undirected)
    assert gv.tolist() == [float(i * (N - 1 - i)) for i in range(N)]
    assert ge.tolist() == [float((i + 1) * (N - 1 - i)) for i in range(N - 1)]


def test_edge_back_into_an_earlier_source_and_relabelling():
    edges = [(0, 1), (1, 2), (2, 3), (3, 0), (0, 2)]
    gv, ge = run(4, edges)
    wv, we = brandes_lists(4, edges, True)
    np.testing.assert_allclose(gv, wv, atol=1e-12)
    np.testing.assert_allclose(ge, we, atol=1e-12)
    hv, _ = run(4, [(3 - s, 3 - t) for s, t in edges])
    np.testing.assert_allclose(hv[::-1], gv, atol=1e-12)


def test_unreachable_vertices_contribute_nothing():
    gv, ge = run(6, [(0, 1), (1, 2), (3, 4)], undirected=True)
    assert np.all(np.isfinite(gv))
    assert gv[1] == 1.0 and gv[0] == 0.0 and gv[5] == 0.0
    assert ge[2] == 1.0


def test_bit_identical_across_thread_counts():
    r = Rng(0xDEADBEEF12345678)
    n = 220
    edges = [(i, (i + 1) % n) for i in range(n)]
    for _ in range(600):
        s = r.below(n)
        t = r.below(n)
        edges.append((s, t))
    base = run(n, edges, threads=1)
    for t in (4, 16):
        got = run(n, edges, threads=t)
        assert got[0].tobytes() == base[0].tobytes()
        assert got[1].tobytes() == base[1].tobytes()


def _random_graph(seed, n, m):
    rng = np.random.default_rng(seed)
    return [(int(a), int(b)) for a, b in zip(rng.integers(0, n, m), rng.integers(0, n, m))]


def test_reversed_view_is_the_turned_graph():
    n, edges = 40, _random_graph(3, 40, 120)
    gv, ge = run(n, edges, reversed=True)
    wv, we = run(n, [(t, s) for s, t in edges])
    np.testing.assert_allclose(gv, wv, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(ge, we, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("undirected", [False, True])
def test_filtered_view_is_the_induced_graph(undirected):
    n, edges = 50, _random_graph(5, 50, 160)
    rng = np.random.default_rng(9)
    vmask = (rng.random(n) < 0.8).astype(np.uint8)
    emask = (rng.random(len(edges)) < 0.8).astype(np.uint8)
    gv, ge = run(n, edges, undirected=undirected, vmask=vmask, emask=emask)
    # The same graph with the dropped edges and every edge touching a dropped
    # vertex turned into a self-loop on vertex 0 (a self-loop is never on a
    # shortest path, and keeps the edge ids in place); dropped vertices are
    # excluded as pivots.
    kept = [(s, t) if emask[i] and vmask[s] and vmask[t] else (0, 0)
            for i, (s, t) in enumerate(edges)]
    pivots = np.flatnonzero(vmask).astype(np.uint32)
    wv, we = run(n, kept, undirected=undirected, pivots=pivots)
    np.testing.assert_allclose(gv, wv, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(ge, we, rtol=1e-12, atol=1e-12)
    assert np.all(gv[vmask == 0] == 0.0)
    assert np.all(ge[emask == 0] == 0.0)


def test_pivots_partition_the_sum():
    """A pivot list restricts the sources; the pivot sets of a partition add
    up to all pivots (the accumulation is linear in the source set)."""
    n, edges = 30, _random_graph(11, 30, 90)
    full_v, full_e = run(n, edges)
    a_v, a_e = run(n, edges, pivots=np.arange(0, 13))
    b_v, b_e = run(n, edges, pivots=np.arange(13, 30))
    np.testing.assert_allclose(a_v + b_v, full_v, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(a_e + b_e, full_e, rtol=1e-12, atol=1e-12)
    with pytest.raises(ValueError):
        run(n, edges, pivots=[30])
