"""Pinning the oracle (CPU): the reference's own known answers, closed forms,
library solves and brute force.  Each check is chosen so a plausible slip in
the restatement (a dropped sink term, a wrong half, a transposed edge, an
inverted parity fix-up) fails it.
"""
import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle
from graphs import EDGES, LAYOUT, Rng, sixty, thousand


def build(n, edges):
    src = np.array([e[0] for e in edges], np.uint32)
    tgt = np.array([e[1] for e in edges], np.uint32)
    return oracle.Adj(n, src, tgt)


# ------------------------------------------------------------ generators

def test_splitmix_known_vector():
    # SplitMix64 seeded with 0: the published first outputs.
    s = oracle.splitmix_stream(0, 3)
    assert int(s[0]) == 0xE220A8397B1DCDAF
    assert int(s[1]) == 0x6E789E6AA1B965F4
    assert int(s[2]) == 0x06C45D188009454F


def test_er_generator_is_the_sequential_stream():
    seed = 0x243F6A8885A308D3
    s, t = oracle.gen_er(97, 50, seed)
    stream = oracle.splitmix_stream(seed, 100)
    assert (s == (stream[0::2] % 97)).all()
    assert (t == (stream[1::2] % 97)).all()


def test_rmat_generator_quadrants():
    seed = 12345
    scale, m = 5, 40
    s, t = oracle.gen_rmat(scale, m, 0.57, 0.19, 0.19, seed)
    stream = oracle.splitmix_stream(seed, scale * m)
    u = (stream >> np.uint64(11)).astype(np.float64) * 2.0**-53
    for i in range(m):
        a = b = 0
        for l in range(scale):
            r = u[i * scale + l]
            bit = 1 << (scale - 1 - l)
            if r < 0.57:
                pass
            elif r < 0.76:
                b |= bit
            elif r < 0.95:
                a |= bit
            else:
                a |= bit
                b |= bit
        assert (s[i], t[i]) == (a, b)


# ---------------------------------------------------------------- layout

def test_layout_fixture_push_swap():
    g = build(4, EDGES)
    for v, (entries, pos) in enumerate(LAYOUT):
        blk, out_len = g.block(v)
        assert blk == entries, f"v{v}"
        assert out_len == pos


def test_export_halves_match_blocks():
    g = build(4, EDGES)
    in_off, in_o, in_e, out_off, out_o, out_e = g.export()
    for v, (entries, pos) in enumerate(LAYOUT):
        ins = list(zip(in_o[in_off[v]:in_off[v + 1]].tolist(), in_e[in_off[v]:in_off[v + 1]].tolist()))
        outs = list(zip(out_o[out_off[v]:out_off[v + 1]].tolist(),
                        out_e[out_off[v]:out_off[v + 1]].tolist()))
        assert outs == entries[:pos]
        assert ins == entries[pos:]


def test_in_half_is_a_queue_under_push_swap():
    # Brute-force model: per vertex, a deque with push_back and rotate.
    from collections import deque
    r = Rng(99)
    n, m = 12, 200
    edges = [(r.below(n), r.below(n)) for _ in range(m)]
    g = build(n, edges)
    q = [deque() for _ in range(n)]
    outs = [[] for _ in range(n)]
    for i, (s, t) in enumerate(edges):
        outs[s].append((t, i))
        if q[s]:
            q[s].rotate(-1)
        q[t].append((s, i))
    for v in range(n):
        blk, ol = g.block(v)
        assert blk[:ol] == outs[v]
        assert blk[ol:] == list(q[v])


# -------------------------------------------------------------- pagerank

def test_one_sweep_two_vertices():
    g = build(2, [(0, 1)])
    r, it = g.pagerank(epsilon=1e-12, max_iter=1)
    assert it == 1
    assert r[0] != 0.5
    assert abs(r[0] - 0.2875) < 1e-15 and abs(r[1] - 0.7125) < 1e-15
    r2, it2 = g.pagerank(epsilon=1e-12, max_iter=2)
    assert it2 == 2
    want = oracle.dense_pagerank(2, [(0, 1, 1.0)], [0.5, 0.5], 0.85, 2)
    assert np.abs(r2 - want).max() < 1e-15


def test_warm_start():
    g = build(2, [(0, 1)])
    r, _ = g.pagerank(epsilon=1e-12, max_iter=1, rank=np.array([0.2875, 0.7125]))
    assert abs(r[0] - (0.15 * 0.5 + 0.85 * (0.7125 * 0.5))) < 1e-15
    assert abs(r[1] - (0.15 * 0.5 + 0.85 * (0.7125 * 0.5 + 0.2875))) < 1e-15


def test_zero_weight_predecessor_is_a_sink():
    g = build(3, [(0, 1), (1, 2)])
    r, _ = g.pagerank(weight=np.array([0.0, 1.0]), epsilon=1e-13, max_iter=1000)
    assert np.isfinite(r).all()
    alone = build(3, [(1, 2)])
    want, _ = alone.pagerank(epsilon=1e-13, max_iter=1000)
    assert np.abs(r - want).max() < 1e-12


def test_personalisation_on_a_cycle():
    g = build(4, [(0, 1), (1, 2), (2, 3), (3, 0)])
    flat, _ = g.pagerank(epsilon=1e-14, max_iter=10_000)
    assert np.abs(flat - 0.25).max() < 1e-12
    pers = [0.7, 0.1, 0.1, 0.1]
    got, _ = g.pagerank(pers=pers, epsilon=1e-14, max_iter=10_000)
    assert got[0] > 0.25 and got[0] > got[1] > got[2] > got[3]
    assert abs(got.sum() - 1.0) < 1e-9
    want = oracle.dense_pagerank(4, [(0, 1, 1.0), (1, 2, 1.0), (2, 3, 1.0), (3, 0, 1.0)],
                                 pers, 0.85, 400)
    assert np.abs(got - want).max() < 1e-12


def test_thousand_matches_dense_power_iteration():
    n, s, t = thousand()
    g = oracle.Adj(n, s, t)
    want = oracle.dense_pagerank(n, [(a, b, 1.0) for a, b in zip(s.tolist(), t.tolist())],
                                 np.full(n, 1.0 / n), 0.85, 400)
    got, it = g.pagerank(epsilon=1e-13, max_iter=10_000)
    assert 5 <= it < 10_000
    assert np.abs(got - want).max() < 1e-12
    assert abs(got.sum() - 1.0) < 1e-9


def test_bit_identical_across_thread_counts():
    n, s, t = thousand()
    g = oracle.Adj(n, s, t)
    base, it = g.pagerank(epsilon=1e-13, max_iter=10_000, threads=1)
    for th in (4, 16):
        got, it2 = g.pagerank(epsilon=1e-13, max_iter=10_000, threads=th)
        assert it2 == it
        assert (got.view(np.uint64) == base.view(np.uint64)).all()


def test_the_plan_fixes_the_fold_order():
    n, s, t = thousand()
    g = oracle.Adj(n, s, t)
    fine, _ = g.pagerank(epsilon=1e-13, max_iter=10_000, grain=16)
    coarse, _ = g.pagerank(epsilon=1e-13, max_iter=10_000, grain=1000)
    assert np.abs(fine - coarse).max() < 1e-14
    assert (fine.view(np.uint64) != coarse.view(np.uint64)).any()


def test_unity_is_explicit_ones_bitwise():
    n, s, t = thousand()
    g = oracle.Adj(n, s, t)
    a, _ = g.pagerank(epsilon=1e-13, max_iter=10_000)
    b, _ = g.pagerank(weight=np.ones(len(s)), epsilon=1e-13, max_iter=10_000)
    assert (a.view(np.uint64) == b.view(np.uint64)).all()


def test_plan_mismatch_is_rejected():
    g = build(4, [(0, 1)])
    with pytest.raises(ValueError, match="plan must cover"):
        g.pagerank(epsilon=1e-6, max_iter=10, plan_len=3)


def _scipy_fixed_point(n, s, t, w, pers, d):
    deg = np.bincount(s, weights=w, minlength=n)
    sink = deg == 0.0
    vals = np.where(deg[s] != 0, w / np.where(deg[s] != 0, deg[s], 1.0), 0.0)
    M = sp.csr_matrix((vals, (t, s)), shape=(n, n))
    # x = (1-d) p + d (M x + (sink . x) p)  =>  (I - d M - d p sink^T) x = (1-d) p
    A = sp.identity(n) - d * M - d * sp.csr_matrix(np.outer(pers, sink.astype(float)))
    return spla.spsolve(A.tocsc(), (1 - d) * pers)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_fixed_point_is_the_linear_solve(seed):
    r = Rng(seed)
    n, m = 300, 1500
    s = np.array([r.below(n) for _ in range(m)], np.int64)
    t = np.array([r.below(n) for _ in range(m)], np.int64)
    w = np.array([1.0 + (r.next() % 1000) / 250.0 for _ in range(m)])
    pers = np.array([1.0 + (r.next() % 7) for _ in range(n)], np.float64)
    pers /= pers.sum()
    g = oracle.Adj(n, s, t)
    got, _ = g.pagerank(weight=w, pers=pers, epsilon=1e-15, max_iter=10_000)
    want = _scipy_fixed_point(n, s, t, w, pers, 0.85)
    assert np.abs(got - want).max() < 1e-12


def test_reversed_view_is_the_reversed_graph():
    r = Rng(7)
    n, m = 200, 1200
    s = np.array([r.below(n) for _ in range(m)], np.uint32)
    t = np.array([r.below(n) for _ in range(m)], np.uint32)
    g = oracle.Adj(n, s, t)
    h = oracle.Adj(n, t, s)
    a, ia = g.pagerank(reversed=True, epsilon=1e-13, max_iter=10_000)
    b, ib = h.pagerank(epsilon=1e-13, max_iter=10_000)
    assert abs(ia - ib) <= 1
    assert np.abs(a - b).max() < 1e-13


def test_filtered_view_is_the_subgraph_with_carried_slots():
    r = Rng(11)
    n, m = 200, 1200
    s = np.array([r.below(n) for _ in range(m)], np.uint32)
    t = np.array([r.below(n) for _ in range(m)], np.uint32)
    vmask = np.array([0 if (i % 7 == 3) else 1 for i in range(n)], np.uint8)
    emask = np.array([0 if (j % 5 == 1) else 1 for j in range(m)], np.uint8)
    g = oracle.Adj(n, s, t)
    start = np.linspace(0.1, 0.9, n)
    got, it = g.pagerank(vmask=vmask, emask=emask, epsilon=1e-13, max_iter=10_000,
                         rank=start)
    # Compact: kept vertices relabelled, kept edges between kept vertices.
    keep = np.flatnonzero(vmask)
    lab = -np.ones(n, np.int64)
    lab[keep] = np.arange(len(keep))
    ke = (emask == 1) & (vmask[s] == 1) & (vmask[t] == 1)
    h = oracle.Adj(len(keep), lab[s[ke]], lab[t[ke]])
    want, iw = h.pagerank(epsilon=1e-13, max_iter=10_000, rank=start[keep])
    assert abs(it - iw) <= 1
    assert np.abs(got[keep] - want).max() < 1e-13
    dropped = np.flatnonzero(vmask == 0)
    assert (got[dropped] == start[dropped]).all()


# ----------------------------------------------------------- eigenvector

def _eigen_residual(n, s, t, w, eig, c):
    ac = np.zeros(n)
    np.add.at(ac, t, w * c[s])
    return np.abs(ac - eig * c).max()


def test_eigenvector_satisfies_its_equation():
    n, s, t, w = sixty()
    g = oracle.Adj(n, s, t)
    c, eig, it = g.eigenvector(weight=w, epsilon=1e-14, max_iter=100_000)
    assert it > 1 and eig > 1.0
    assert _eigen_residual(n, s.astype(np.int64), t.astype(np.int64), w, eig, c) < 1e-10
    assert abs(np.sqrt((c * c).sum()) - 1.0) < 1e-12


def test_directed_cycle_has_unit_eigenvalue():
    n = 7
    g = build(n, [(i, (i + 1) % n) for i in range(n)])
    c, eig, _ = g.eigenvector(epsilon=1e-15, max_iter=1000)
    assert abs(eig - 1.0) < 1e-12
    assert np.abs(c - 1.0 / math.sqrt(n)).max() < 1e-12
    _, eig2, _ = g.eigenvector(weight=2.5, epsilon=1e-15, max_iter=1000)
    assert abs(eig2 - 2.5) < 1e-12


def test_edgeless_graph_has_zero_eigenvalue():
    g = build(5, [])
    c, eig, it = g.eigenvector(epsilon=1e-12, max_iter=100)
    assert eig == 0.0 and it >= 1
    assert (c == 0.0).all()


def test_one_eigenvector_sweep_on_a_path():
    g = build(3, [(0, 1), (1, 2)])
    c, eig, it = g.eigenvector(epsilon=1e-12, max_iter=1)
    assert it == 1
    assert abs(eig - math.sqrt(2) / 3) < 1e-15
    h = 1 / math.sqrt(2)
    assert abs(c[0]) < 1e-15 and abs(c[1] - h) < 1e-15 and abs(c[2] - h) < 1e-15


def test_eigenvector_matches_dense_eigenpair_on_symmetric_graph():
    r = Rng(5)
    n = 80
    pairs = [(r.below(n), r.below(n)) for _ in range(300)]
    pairs += [(i, (i + 1) % n) for i in range(n)]
    s = [a for a, b in pairs] + [b for a, b in pairs]
    t = [b for a, b in pairs] + [a for a, b in pairs]
    g = oracle.Adj(n, np.array(s, np.uint32), np.array(t, np.uint32))
    c, eig, _ = g.eigenvector(epsilon=1e-15, max_iter=100_000)
    A = np.zeros((n, n))
    np.add.at(A, (np.array(t), np.array(s)), 1.0)
    vals, vecs = np.linalg.eigh(A)
    v = vecs[:, -1] * np.sign(vecs[:, -1].sum())
    assert abs(eig - vals[-1]) < 1e-9
    assert np.abs(c - v).max() < 1e-7


def test_nan_weight_stops_after_one_sweep():
    """IEEE comparisons fix the reference's behaviour on a NaN weight: the
    source's degree is NaN (not 0: no sink), its out-neighbours' ranks NaN,
    delta NaN, and `while delta >= epsilon` ends after one sweep
    (centrality.rs:284-291, 333-372).  Vertices no NaN reaches keep finite
    ranks."""
    n, s, t = thousand()
    w = np.ones(len(s))
    w[17] = np.nan
    x, it = oracle.Adj(n, s, t).pagerank(weight=w, epsilon=1e-10, max_iter=1000)
    assert it == 1
    bad = np.isnan(x)
    assert bad[t[s == s[17]]].all()
    assert bad.sum() == len(set(t[s == s[17]].tolist()))
    xe, eig, ite = oracle.Adj(n, s, t).eigenvector(weight=w, epsilon=1e-10, max_iter=1000)
    assert ite == 1 and np.isnan(eig) and np.isnan(xe).all()
