"""GPU parity above 2^22 vertices, the width at which a source label and a
10-bit slot stop fitting one 32-bit word (the reference's ids are u32 up to
MAX_INDEX = u32::MAX - 1, crates/gt-core/src/ids.rs:24-32).

Two layouts of the wave-tile stream are reachable there, and each is driven
here and identified through gt_last_layout:
  * packed -- ((source - tile base) << 10) | slot, when every tile's sources
    span fewer than 2^22 labels (local edges, block-diagonal batches);
  * plain  -- 32-bit sources in row order, when some tile's sources span more
    (a random functional graph on 4.5 M vertices: every vertex is a source, so
    the relabelled labels cover the whole range, and so does almost every
    tile).
Each layout is checked against the oracle: FP64 PageRank (unity, and weighted
with a dense personalisation) at L1 <= 1e-12 with sweeps +-1, the FP32 form at
L1 <= 1e-6, eigenvector iterates at L1 <= 1e-12 ||x||_1, FP64 / FP32 adjoints, filtered
views, and an uneven batch whose union graph exceeds 2^22 vertices.
"""
import numpy as np
import pytest

import gt_hip
import oracle
from test_gpu_uneven import check_forward

pytestmark = pytest.mark.gpu

WIDE = 1 << 22
THREADS = 8


def _params(n, eps, max_iter=0, grain=4096):
    return gt_hip.PowerIteration(eps, max_iter, gt_hip.Plan(n, grain))


def _parity(got, it, want, it_o, eps, tol=1e-12):
    l1 = float(np.abs(np.asarray(got) - np.asarray(want)).sum())
    if it == it_o:
        assert l1 <= tol, f"L1 {l1:.3e} at {it} sweeps"
    else:
        assert abs(it - it_o) == 1, f"sweeps {it} vs {it_o}"
        assert l1 <= eps, f"L1 {l1:.3e}, sweeps {it} vs {it_o}"


def _functional(n, seed, hub_in=3000):
    """Every vertex one out-edge to a uniform target, plus `hub_in` extra
    in-edges of vertex 7 (a split row) from uniform sources."""
    rng = np.random.default_rng(seed)
    s = np.arange(n, dtype=np.uint32)
    t = rng.integers(0, n, n, dtype=np.uint32)
    hs = rng.integers(0, n, hub_in, dtype=np.uint32)
    return np.concatenate([s, hs]), np.concatenate([t, np.full(hub_in, 7, np.uint32)])


def _local(n):
    """v -> v+1 and v -> v+2 (those below n) for v not a multiple of 3;
    multiples of 3 are sinks.  The relabelling puts the sources first in
    about their old order, so every tile spans a few labels; the sinks'
    dangling mass mixes the iteration as an R-MAT graph's does."""
    v = np.arange(n, dtype=np.int64)
    v = v[v % 3 != 0]
    s = np.concatenate([v, v])
    t = np.concatenate([v + 1, v + 2])
    keep = t < n
    return s[keep].astype(np.uint32), t[keep].astype(np.uint32)


@pytest.fixture(scope="module")
def plain_graph():
    n = 4_500_000
    s, t = _functional(n, 11)
    g = gt_hip.AdjList.with_vertices(n)
    g.add_edges(s, t)
    return n, s, t, g, oracle.Adj(n, s, t)


@pytest.fixture(scope="module")
def local_graph():
    n = WIDE + 300_001
    s, t = _local(n)
    g = gt_hip.AdjList.with_vertices(n)
    g.add_edges(s, t)
    return n, s, t, g, oracle.Adj(n, s, t)


def _pr(g, n, fn, weight=None, pers=None, eps=1e-10, max_iter=0):
    r = gt_hip.DenseProp.new(g.graph_id)
    w = gt_hip.Unity.NEW if weight is None else weight
    it = fn(g, w, pers, 0.85, _params(n, eps, max_iter), r)
    return np.asarray(r.data), it


def _f64(g, w, pers, d, params, r):
    if pers is None:
        return gt_hip.pagerank(g, w, d, params, r)
    return gt_hip.pagerank_with(g, w, pers, d, params, r)


def _engines(g):
    """Runs the body on block tiles (the default) and on the wave engine,
    checking which one ran."""
    dev = g.device()
    for blk in (0, -1):
        dev.set_tuning(block_tile_nnz=blk)
        yield blk
        assert (dev.layout()["block_tiles"] > 0) == (blk == 0), dev.layout()
    dev.set_tuning(block_tile_nnz=0)


def test_plain_stream_pagerank_f64(gpu, plain_graph):
    n, s, t, g, o = plain_graph
    want, it_o = o.pagerank(epsilon=1e-10, grain=4096, threads=THREADS)
    for _ in _engines(g):
        got, it = _pr(g, n, _f64)
        assert g.device().layout()["slot_bits"] == 0
        _parity(got, it, want, it_o, 1e-10)


def test_plain_stream_weighted_personalised(gpu, plain_graph):
    n, s, t, g, o = plain_graph
    rng = np.random.default_rng(3)
    w = 0.1 + 0.9 * rng.random(len(s))
    p = rng.random(n) + 0.05
    want, it_o = o.pagerank(weight=w, pers=p, epsilon=1e-10, grain=4096, threads=THREADS)
    for _ in _engines(g):
        got, it = _pr(g, n, _f64, weight=gt_hip.DenseProp.from_vec(g.graph_id, w),
                      pers=gt_hip.DenseProp.from_vec(g.graph_id, p))
        assert g.device().layout()["slot_bits"] == 0
        _parity(got, it, want, it_o, 1e-10, tol=1e-12 * max(1.0, float(p.sum())))


def test_plain_stream_mixed(gpu, plain_graph):
    """The FP32 form on the block tiles (the default) and on the wave engine's
    plain stream."""
    n, s, t, g, o = plain_graph
    want, _ = o.pagerank(epsilon=1e-12, grain=4096, threads=THREADS)
    dev = g.device()
    for blk in (0, -1):
        dev.set_tuning(block_tile_nnz=blk)
        got, it = _pr(g, n, gt_hip.pagerank_mixed, eps=1e-9)
        lay = dev.layout()
        assert lay["slot_bits"] == 0 and (lay["block_tiles"] > 0) == (blk == 0), lay
        assert np.abs(got - want).sum() <= 1e-6
        assert it >= 1
    dev.set_tuning(block_tile_nnz=0)


def test_plain_stream_eigenvector(gpu, plain_graph):
    # A random functional graph need not converge: both sides run 25 sweeps.
    n, s, t, g, o = plain_graph
    xo, eo, ito = o.eigenvector(epsilon=1e-300, max_iter=25, grain=4096, threads=THREADS)
    for _ in _engines(g):
        x = gt_hip.DenseProp.new(g.graph_id)
        eig, it = gt_hip.eigenvector(g, gt_hip.Unity.NEW, _params(n, 1e-300, 25), x)
        assert g.device().layout()["slot_bits"] == 0
        assert it == ito == 25
        assert np.abs(np.asarray(x.data) - xo).sum() <= 1e-12 * max(1.0, np.abs(xo).sum())
        assert abs(eig - eo) <= 1e-12 * abs(eo)


def test_plain_stream_filtered_view(gpu, plain_graph):
    # Filtered views sweep in the caller's labels (no relabelling): every
    # tile of a random graph spans the whole range.
    n, s, t, g, o = plain_graph
    rng = np.random.default_rng(8)
    vm = (rng.random(n) < 0.9).astype(np.uint8)
    em = (rng.random(len(s)) < 0.8).astype(np.uint8)
    view = g.masked(vm, em)
    r = gt_hip.DenseProp.new(g.graph_id)
    it = gt_hip.pagerank(view, gt_hip.Unity.NEW, 0.85, _params(n, 1e-10), r)
    assert g.device().layout()["slot_bits"] == 0
    want, it_o = o.pagerank(epsilon=1e-10, grain=4096, vmask=vm, emask=em, threads=THREADS)
    _parity(np.asarray(r.data), it, want, it_o, 1e-10)


@pytest.mark.parametrize("precision,tol", [("f64", 1e-9), ("f32", 2e-4)])
def test_plain_stream_adjoint(gpu, plain_graph, precision, tol):
    n, s, t, g, o = plain_graph
    rng = np.random.default_rng(5)
    w = (1.0 - 0.9 * rng.random(len(s))).astype(np.float32).astype(np.float64)
    c = rng.standard_normal(n)
    x, _ = o.pagerank(weight=w, epsilon=1e-14, max_iter=400, grain=4096, threads=THREADS)
    # Both sides stop at the same cap, so lambda is the same truncated series.
    gw_o, gp_o, gd_o, _, it_o = o.pagerank_grad(x, c, weight=w, epsilon=1e-300, max_iter=40)
    wm = gt_hip.DenseProp.from_vec(g.graph_id, w)
    gw, gp, gd, it = gt_hip.pagerank_grad(g, wm, None, 0.85, _params(n, 1e-300, 40), x, c,
                                          precision=precision)
    assert it == it_o == 40
    assert np.abs(gw - gw_o).max() <= tol * np.abs(gw_o).max()
    assert np.abs(gp - gp_o).max() <= tol * np.abs(gp_o).max()
    assert abs(gd - gd_o) <= tol * max(1.0, abs(gd_o))


def test_packed_stream_above_2_22(gpu, local_graph):
    n, s, t, g, o = local_graph
    want, it_o = o.pagerank(epsilon=1e-10, grain=4096, threads=THREADS)
    for _ in _engines(g):
        got, it = _pr(g, n, _f64, eps=1e-10)
        lay = g.device().layout()
        assert lay["slot_bits"] == 10, lay
        _parity(got, it, want, it_o, 1e-10)
    dev = g.device()
    for blk in (0, -1):
        dev.set_tuning(block_tile_nnz=blk)
        got32, _ = _pr(g, n, gt_hip.pagerank_mixed, eps=1e-9)
        lay = dev.layout()
        assert lay["slot_bits"] == 10 and (lay["block_tiles"] > 0) == (blk == 0), lay
        assert np.abs(got32 - want).sum() <= 1e-6
    dev.set_tuning(block_tile_nnz=0)
    xo, eo, ito = o.eigenvector(epsilon=1e-300, max_iter=20, grain=4096, threads=THREADS)
    for _ in _engines(g):
        x = gt_hip.DenseProp.new(g.graph_id)
        eig, it = gt_hip.eigenvector(g, gt_hip.Unity.NEW, _params(n, 1e-300, 20), x)
        assert it == ito == 20
        # Unit L2 norm over millions of entries: the bar is relative to ||x||_1.
        assert np.abs(np.asarray(x.data) - xo).sum() <= 1e-12 * max(1.0, np.abs(xo).sum())
        assert abs(eig - eo) <= 1e-12 * abs(eo)


def test_uneven_batch_union_above_2_22(gpu):
    G = 600
    voff, eoff, d, s, t, w = gt_hip.gen_chung_lu_batch(G, 0x5EED, n_min=64, n_max=65536)
    n = int(voff[-1])
    assert n > WIDE
    dev = gt_hip.DeviceGraph.build(n, s, t)
    x, it = gt_hip.pagerank_uneven(dev, voff, d, weight=w, epsilon=1e-6)
    assert dev.layout()["slot_bits"] == 10
    rng = np.random.default_rng(2)
    sample = sorted(set(rng.choice(G, 24, replace=False).tolist()) | {G - 1})
    for k in sample:
        lo, hi = int(eoff[k]), int(eoff[k + 1])
        nk = int(voff[k + 1] - voff[k])
        ok = oracle.Adj(nk, s[lo:hi] - voff[k], t[lo:hi] - voff[k])
        wk = w[lo:hi].astype(np.float64)
        check_forward(ok, x[voff[k]:voff[k + 1]], int(it[k]), wk, float(d[k]))
    # The adjoint from the exact fixed point on the graphs it is checked on.
    c = rng.standard_normal(n)
    xs = np.array(x, dtype=np.float64)
    checked = sample[:6] + [G - 1]
    grads = {}
    for k in checked:
        lo, hi = int(eoff[k]), int(eoff[k + 1])
        vk = slice(int(voff[k]), int(voff[k + 1]))
        ok = oracle.Adj(vk.stop - vk.start, s[lo:hi] - voff[k], t[lo:hi] - voff[k])
        wk = w[lo:hi].astype(np.float64)
        xk, _ = ok.pagerank(weight=wk, damping=float(d[k]), epsilon=1e-15, max_iter=100_000)
        xs[vk] = xk
        grads[k] = ok.pagerank_grad(xk, c[vk], weight=wk, damping=float(d[k]), epsilon=1e-12)
    gw, gp, gd, _ = gt_hip.pagerank_uneven_grad(dev, voff, d, xs, c, weight=w, epsilon=1e-9)
    assert dev.layout()["slot_bits"] == 10
    for k in checked:
        gw_o, gp_o, gd_o, _, _ = grads[k]
        lo, hi = int(eoff[k]), int(eoff[k + 1])
        vk = slice(int(voff[k]), int(voff[k + 1]))
        assert np.abs(gw[lo:hi] - gw_o).max() <= 2e-4 * np.abs(gw_o).max(), k
        assert np.abs(gp[vk] - gp_o).max() <= 2e-4 * np.abs(gp_o).max(), k
        assert abs(gd[k] - gd_o) <= 2e-4 * max(1.0, abs(gd_o)), k
