"""The block-tile sweeps against the oracle.

The FP32 sweep is the default FP32 engine of pagerank_mixed on relabelled,
unfiltered, unweighted graphs. Each result is checked against the oracle at
the FP32 form's bar (L1 <= 1e-6 of the FP64 fixed point), against the wave
engine (the same solve with block_tile_nnz = -1), and bit for bit run to run.
The two-word FP64 sweeps -- the FP64 PageRank drop-in's and the eigenvector's,
both on by default -- are held to the drop-in's bar (L1 <= 1e-12 and the same
sweep count +-1) and to the wave engine's result (block_f64 = -1, or
block_tile_nnz = -1) to rounding.  Weighted PageRank runs on the block tiles
too when every weight is finite and non-negative.

The cases cover the layout's edge cases:
- long rows summed in lane slots;
- rows longer than a tile, cut into chunk tiles;
- tiles capped by their row range (empty rows);
- groups cut by their source span;
- warm starts, a sweep cap, and a personalisation with negative entries;
- an iterate whose mass the fixed point cannot hold, where the solve must
  come from the wave engine.
"""
import numpy as np
import pytest

import gt_hip
import oracle
from graphs import Rng

pytestmark = pytest.mark.gpu


def adj(n, s, t):
    g = gt_hip.AdjList.with_vertices(n)
    g.add_edges(np.asarray(s, np.uint32), np.asarray(t, np.uint32))
    return g


def params(n, eps, max_iter):
    return gt_hip.PowerIteration(eps, max_iter, gt_hip.Plan(n, 64))


def mixed(g, eps=1e-9, max_iter=1000, pers=None, rank=None, reversed_=False, weight=None):
    r = gt_hip.DenseProp.new(g.graph_id) if rank is None else rank
    target = g.reversed() if reversed_ else g
    w = gt_hip.Unity.NEW if weight is None else weight
    it = gt_hip.pagerank_mixed(target, w, pers, 0.85, params(g.vertex_bound(), eps, max_iter), r)
    return np.asarray(r.data).copy(), it


def run_both(g, **kw):
    """(block result, its layout), (wave-engine result) for one solve."""
    dev = g.device()
    dev.set_tuning(block_tile_nnz=kw.pop("tile", 0), block_long_min=kw.pop("long_min", 0))
    blk = mixed(g, **kw)
    lay = dev.layout()
    dev.set_tuning(block_tile_nnz=-1)
    wave = mixed(g, **kw)
    assert dev.layout()["block_tiles"] == 0
    dev.set_tuning(block_tile_nnz=0, block_long_min=0)
    return blk, lay, wave


def check(n, s, t, pers=None, tile=0, long_min=0, reversed_=False):
    g = adj(n, s, t)
    pp = None if pers is None else gt_hip.DenseProp.from_vec(g.graph_id, pers)
    (x, it), lay, (xw, itw) = run_both(g, tile=tile, long_min=long_min, pers=pp,
                                       reversed_=reversed_)
    assert lay["block_tiles"] > 0, lay
    want, _ = oracle.Adj(n, s, t).pagerank(pers=pers, epsilon=1e-12, max_iter=1000,
                                           reversed=reversed_, threads=8)
    scale = max(1.0, float(np.abs(want).sum()))
    assert np.abs(x - want).sum() <= 1e-6 * scale, (np.abs(x - want).sum(), it)
    # The two engines differ only in the FP32 row sums' rounding.
    assert abs(it - itw) <= 1
    assert np.abs(x - xw).sum() <= 2e-7 * scale
    return g, lay, x, it


def test_rmat_default_tiles(gpu):
    """R-MAT 16 (1 M entries, 16 default tiles, hub rows longer than the
    long-row threshold) and R-MAT 12 (one tile)."""
    for scale, seed in ((16, 41), (12, 42)):
        n = 1 << scale
        s, t = oracle.gen_rmat(scale, 16 * n, 0.57, 0.19, 0.19, seed)
        g, lay, _, _ = check(n, s, t)
        if scale == 16:
            assert lay["block_tiles"] >= 16 and lay["block_long_rows"] > 0, lay


def test_small_tiles_long_rows_and_chunks(gpu):
    """1,024-entry tiles and a long-row threshold of 16: many long rows per
    tile, and every hub longer than a tile cut into chunk tiles."""
    n = 1 << 13
    s, t = oracle.gen_rmat(13, 16 * n, 0.57, 0.19, 0.19, 43)
    indeg = np.bincount(t, minlength=n)
    assert indeg.max() > 1024
    _, lay, _, _ = check(n, s, t, tile=1024, long_min=16)
    assert lay["block_long_rows"] >= int((indeg > 16).sum())


def test_reversed_view_and_personalisation_with_negative_entries(gpu):
    """Out-halves as the pull orientation; personalisation with negative
    entries (normalised to sum 1), so the iterate's L1 mass exceeds its sum
    and the fixed-point scale follows the mass."""
    n = 1 << 14
    s, t = oracle.gen_rmat(14, 16 * n, 0.57, 0.19, 0.19, 44)
    p = np.random.default_rng(7).uniform(-0.5, 1.5, n)
    p /= p.sum()
    check(n, s, t, pers=p, reversed_=True)
    check(n, s, t, pers=p, tile=2048, long_min=64)


def test_empty_rows_cap_the_tile_ranges(gpu):
    """100,003 vertices and 20,000 edges: most rows are empty, so tiles are
    cut by their row range (8,064 rows) rather than their entries."""
    r = Rng(45)
    n, m = 100_003, 20_000
    s = np.array([r.below(n) for _ in range(m)], np.uint32)
    t = np.array([r.below(n) for _ in range(m)], np.uint32)
    _, lay, _, _ = check(n, s, t)
    assert lay["block_tiles"] >= n // 8064


def test_groups_cut_by_source_span(gpu):
    """2^22 + 5 vertices, vertex v's one out-edge to v mod 600,000: every
    vertex is gathered once (so relabelling keeps the ids), and row r's
    sources are r + 600,000 k.  A 1,024-entry tile sorts into runs of about
    146 consecutive sources 600,000 labels apart; a group spans fewer than
    2^19 labels, so every run ends a group early.  Without such cuts a tile
    needs at most entries / 64 + 4 groups (whole blocks of four)."""
    n = (1 << 22) + 5
    s = np.arange(n, dtype=np.uint32)
    t = (s % 600_000).astype(np.uint32)
    _, lay, _, _ = check(n, s, t, tile=1024)
    assert lay["block_groups"] > n // 64 + 4 * lay["block_tiles"] + 6_000, lay


def test_warm_start_and_cap(gpu):
    """A full and a partial warm start, and a cap of three sweeps (no FP64
    refinement) against the oracle's three sweeps.

    The FP32 phase stops once delta < 2^-22 * mass.  From a random start the
    slowest mode decays at the damping d, so that iterate lies within
    delta * d / (1 - d) of the fixed point, and the FP64 sweep multiplies
    that by d: about 1.2e-6 in L1 for mass 1 at d = 0.85.  That is the bar
    here (the cold starts above stay far inside 1e-6); the wave engine must
    agree with the block sweep to rounding."""
    n = 1 << 13
    s, t = oracle.gen_rmat(13, 16 * n, 0.57, 0.19, 0.19, 47)
    g = adj(n, s, t)
    warm = np.random.default_rng(8).uniform(0.0, 2.0 / n, n)
    bar = 2.0**-22 * 0.85 / 0.15 * 0.85 * 1.05
    for had in (n, n // 3):
        r = gt_hip.DenseProp.from_vec(g.graph_id, warm[:had].copy())
        x, it = mixed(g, rank=r)
        assert g.device().layout()["block_tiles"] > 0
        g.device().set_tuning(block_tile_nnz=-1)
        r = gt_hip.DenseProp.from_vec(g.graph_id, warm[:had].copy())
        xw, itw = mixed(g, rank=r)
        g.device().set_tuning(block_tile_nnz=0)
        want, _ = oracle.Adj(n, s, t).pagerank(epsilon=1e-12, max_iter=1000,
                                               rank=warm[:had].copy(), threads=8)
        assert np.abs(x - want).sum() <= bar, (np.abs(x - want).sum(), had)
        assert np.abs(xw - want).sum() <= bar, (np.abs(xw - want).sum(), had)
        assert abs(it - itw) <= 1 and np.abs(x - xw).sum() <= 2e-7
    x, it = mixed(g, max_iter=3)
    want, it_o = oracle.Adj(n, s, t).pagerank(epsilon=1e-9, max_iter=3, threads=8)
    assert it == it_o == 3
    assert np.abs(x - want).sum() <= 1e-6


def test_bit_for_bit_run_to_run(gpu):
    """Integer row sums: the same bits whatever order the workgroups add in."""
    n = 1 << 15
    s, t = oracle.gen_rmat(15, 16 * n, 0.57, 0.19, 0.19, 48)
    g = adj(n, s, t)
    a, ia = mixed(g)
    b, ib = mixed(g)
    assert ia == ib and np.array_equal(a, b)


def test_unrepresentable_mass_runs_on_the_wave_engine(gpu):
    """A personalisation of 1e300 overflows FP32, so the first iterate's mass is
    infinite: the block sweep gives up and the solve is the wave engine's, bit
    for bit (NaN where it has NaN), with its sweep count."""
    n = 1 << 10
    s, t = oracle.gen_rmat(10, 16 * n, 0.57, 0.19, 0.19, 49)
    g = adj(n, s, t)
    p = np.full(n, 1e300)
    pp = gt_hip.DenseProp.from_vec(g.graph_id, p)
    (x, it), lay, (xw, itw) = run_both(g, pers=pp)
    assert lay["block_tiles"] == 0
    assert it == itw and np.array_equal(x, xw, equal_nan=True)


# ------------------------------------------------ FP64 PageRank

def parity(got, it, want, it_o, eps, tol=1e-12):
    """The drop-in's bar: same sweep count and L1 <= tol, or one sweep apart
    when the oracle's last delta sat within rounding of epsilon (then the
    answers are within epsilon)."""
    l1 = float(np.abs(np.asarray(got) - np.asarray(want)).sum())
    if it == it_o:
        assert l1 <= tol, f"L1 {l1:.3e} at {it} sweeps"
    else:
        assert abs(it - it_o) == 1, f"sweeps {it} vs {it_o}"
        assert l1 <= eps, f"L1 {l1:.3e}, sweeps {it} vs {it_o}"


def pr64(g, eps=1e-12, max_iter=1000, pers=None, rank=None, reversed_=False, weight=None):
    r = gt_hip.DenseProp.new(g.graph_id) if rank is None else rank
    target = g.reversed() if reversed_ else g
    p = params(g.vertex_bound(), eps, max_iter)
    w = gt_hip.Unity.NEW if weight is None else weight
    if pers is None:
        it = gt_hip.pagerank(target, w, 0.85, p, r)
    else:
        it = gt_hip.pagerank_with(target, w, pers, 0.85, p, r)
    return np.asarray(r.data).copy(), it


def check64(n, s, t, pers=None, tile=0, long_min=0, reversed_=False):
    """The FP64 drop-in on block tiles (the default), against the oracle at
    the drop-in's bar and against the wave engine (block_f64 = -1) to
    rounding."""
    g = adj(n, s, t)
    dev = g.device()
    pp = None if pers is None else gt_hip.DenseProp.from_vec(g.graph_id, pers)
    dev.set_tuning(block_f64=1, block_tile_nnz=tile, block_long_min=long_min)
    x, it = pr64(g, pers=pp, reversed_=reversed_)
    lay = dev.layout()
    dev.set_tuning(block_f64=-1, block_tile_nnz=0, block_long_min=0)
    pp = None if pers is None else gt_hip.DenseProp.from_vec(g.graph_id, pers)
    xw, itw = pr64(g, pers=pp, reversed_=reversed_)
    assert dev.layout()["block_tiles"] == 0
    dev.set_tuning(block_f64=0)
    assert lay["block_tiles"] > 0, lay
    want, it_o = oracle.Adj(n, s, t).pagerank(pers=pers, epsilon=1e-12, max_iter=1000,
                                              reversed=reversed_, threads=8)
    parity(x, it, want, it_o, 1e-12)
    assert abs(it - itw) <= 1
    if it == itw:
        assert np.abs(x - xw).sum() <= 1e-12
    return g, lay, x, it


def test_f64_rmat_and_small_tiles(gpu):
    """R-MAT 16 on default tiles; R-MAT 13 on 1,024-entry tiles with a long-row
    threshold of 16 (long rows in lane slots, hubs cut into chunk tiles)."""
    n = 1 << 16
    s, t = oracle.gen_rmat(16, 16 * n, 0.57, 0.19, 0.19, 51)
    _, lay, _, _ = check64(n, s, t)
    assert lay["block_long_rows"] > 0
    n = 1 << 13
    s, t = oracle.gen_rmat(13, 16 * n, 0.57, 0.19, 0.19, 52)
    check64(n, s, t, tile=1024, long_min=16)


def test_f64_reversed_and_negative_personalisation(gpu):
    n = 1 << 14
    s, t = oracle.gen_rmat(14, 16 * n, 0.57, 0.19, 0.19, 53)
    p = np.random.default_rng(9).uniform(-0.5, 1.5, n)
    p /= p.sum()
    check64(n, s, t, pers=p, reversed_=True)
    check64(n, s, t, pers=p, tile=2048, long_min=64)


def test_f64_empty_rows_warm_start_and_truncation(gpu):
    """Mostly empty rows; a partial warm start; caps of 1, 2 and 3 sweeps
    against the oracle's."""
    r = Rng(54)
    n, m = 100_003, 20_000
    s = np.array([r.below(n) for _ in range(m)], np.uint32)
    t = np.array([r.below(n) for _ in range(m)], np.uint32)
    check64(n, s, t)
    n = 1 << 13
    s, t = oracle.gen_rmat(13, 16 * n, 0.57, 0.19, 0.19, 55)
    g = adj(n, s, t)
    g.device().set_tuning(block_f64=1)
    warm = np.random.default_rng(10).uniform(0.0, 2.0 / n, n // 3)
    x, it = pr64(g, rank=gt_hip.DenseProp.from_vec(g.graph_id, warm.copy()))
    assert g.device().layout()["block_tiles"] > 0
    want, it_o = oracle.Adj(n, s, t).pagerank(epsilon=1e-12, max_iter=1000, rank=warm.copy(),
                                              threads=8)
    parity(x, it, want, it_o, 1e-12)
    for k in (1, 2, 3):
        x, it = pr64(g, eps=1e-30, max_iter=k)
        want, it_o = oracle.Adj(n, s, t).pagerank(epsilon=1e-30, max_iter=k, threads=8)
        assert it == it_o == k
        assert np.abs(x - want).sum() <= 1e-14
    g.device().set_tuning(block_f64=0)


def test_f64_bit_for_bit_and_unrepresentable_mass(gpu):
    """Integer row sums: the same bits run to run.  A personalisation of
    1e300 puts the iterate's mass above what the fixed point holds: the solve
    is the wave engine's, bit for bit."""
    n = 1 << 15
    s, t = oracle.gen_rmat(15, 16 * n, 0.57, 0.19, 0.19, 56)
    g = adj(n, s, t)
    g.device().set_tuning(block_f64=1)
    a, ia = pr64(g)
    b, ib = pr64(g)
    assert g.device().layout()["block_tiles"] > 0
    assert ia == ib and np.array_equal(a.view(np.uint64), b.view(np.uint64))
    p = np.full(n, 1e300)
    x, it = pr64(g, eps=1e-6, pers=gt_hip.DenseProp.from_vec(g.graph_id, p))
    assert g.device().layout()["block_tiles"] == 0
    g.device().set_tuning(block_f64=-1)
    xw, itw = pr64(g, eps=1e-6, pers=gt_hip.DenseProp.from_vec(g.graph_id, p))
    g.device().set_tuning(block_f64=0)
    assert it == itw and np.array_equal(x, xw, equal_nan=True)


# ---------------------------------------------------------- eigenvector

def ev(g, eps, max_iter, x=None, reversed_=False):
    xm = gt_hip.DenseProp.new(g.graph_id) if x is None else x
    target = g.reversed() if reversed_ else g
    eig, it = gt_hip.eigenvector(target, gt_hip.Unity.NEW, params(g.vertex_bound(), eps, max_iter),
                                 xm)
    return np.asarray(xm.data).copy(), eig, it


def sym_rmat(scale, seed):
    s, t = oracle.gen_rmat(scale, 16 << scale, 0.57, 0.19, 0.19, seed)
    return 1 << scale, np.concatenate([s, t]), np.concatenate([t, s])


def check_ev(n, s, t, eps, max_iter, tile=0, long_min=0, x0=None, reversed_=False):
    """The block eigenvector (the default) against the oracle at the drop-in's
    bar, and against the wave engine to rounding."""
    g = adj(n, s, t)
    dev = g.device()
    dev.set_tuning(block_tile_nnz=tile, block_long_min=long_min)
    xm = None if x0 is None else gt_hip.DenseProp.from_vec(g.graph_id, x0.copy())
    c, eig, it = ev(g, eps, max_iter, x=xm, reversed_=reversed_)
    lay = dev.layout()
    dev.set_tuning(block_tile_nnz=-1)
    xm = None if x0 is None else gt_hip.DenseProp.from_vec(g.graph_id, x0.copy())
    cw, eigw, itw = ev(g, eps, max_iter, x=xm, reversed_=reversed_)
    assert dev.layout()["block_tiles"] == 0
    dev.set_tuning(block_tile_nnz=0, block_long_min=0)
    assert lay["block_tiles"] > 0, lay
    co, eo, ito = oracle.Adj(n, s, t).eigenvector(epsilon=eps, max_iter=max_iter,
                                                  x=None if x0 is None else x0.copy(),
                                                  reversed=reversed_, threads=8)
    parity(c, it, co, ito, eps)
    assert abs(eig - eo) <= max(1e-12, eps) * eo
    assert abs(it - itw) <= 1
    if it == itw:
        assert np.abs(c - cw).sum() <= 1e-12 and abs(eig - eigw) <= 1e-13 * eigw
    return g, lay, c, eig, it


def test_ev_symmetrised_rmat_and_truncation(gpu):
    """Symmetrised R-MAT 13 at caps of 1, 2 and 3 sweeps and to epsilon 1e-9;
    R-MAT 16 (directed, hub rows in lane slots) to 1e-10."""
    n, S, T = sym_rmat(13, 57)
    for k in (1, 2, 3):
        check_ev(n, S, T, 1e-30, k)
    check_ev(n, S, T, 1e-9, 1000)
    n = 1 << 16
    s, t = oracle.gen_rmat(16, 16 * n, 0.57, 0.19, 0.19, 58)
    S, T = np.concatenate([s, t]), np.concatenate([t, s])
    _, lay, _, _, _ = check_ev(n, S, T, 1e-10, 1000)
    assert lay["block_long_rows"] > 0


def test_ev_small_tiles_reversed_and_warm_start(gpu):
    n, S, T = sym_rmat(13, 59)
    check_ev(n, S, T, 1e-10, 1000, tile=1024, long_min=16)
    s, t = oracle.gen_rmat(12, 16 << 12, 0.45, 0.15, 0.15, 60)
    n = 1 << 12
    # A strongly connected core keeps the directed iteration convergent.
    S = np.concatenate([s, np.arange(n, dtype=np.uint32)])
    T = np.concatenate([t, (np.arange(n, dtype=np.uint32) + 1) % n]).astype(np.uint32)
    check_ev(n, S, T, 1e-10, 2000, reversed_=True)
    warm = np.random.default_rng(11).uniform(0.5, 1.5, n // 2)
    check_ev(n, S, T, 1e-10, 2000, x0=warm)


def test_ev_bit_for_bit_and_unrepresentable_start(gpu):
    """Integer row sums: the same bits run to run.  A start of 1e300 per
    vertex puts the mass (sum of |x| * out-degree) above what the fixed point
    holds: the solve is the wave engine's, bit for bit."""
    n, S, T = sym_rmat(14, 61)
    g = adj(n, S, T)
    a, ea, ia = ev(g, 1e-10, 1000)
    assert g.device().layout()["block_tiles"] > 0
    b, eb, ib = ev(g, 1e-10, 1000)
    assert ia == ib and ea == eb and np.array_equal(a.view(np.uint64), b.view(np.uint64))
    big = np.full(n, 1e300)
    x, e, it = ev(g, 1e-10, 1000, x=gt_hip.DenseProp.from_vec(g.graph_id, big.copy()))
    assert g.device().layout()["block_tiles"] == 0
    g.device().set_tuning(block_tile_nnz=-1)
    xw, ew, itw = ev(g, 1e-10, 1000, x=gt_hip.DenseProp.from_vec(g.graph_id, big.copy()))
    g.device().set_tuning(block_tile_nnz=0)
    assert it == itw and np.array_equal(x, xw, equal_nan=True)
    assert e == ew or (np.isnan(e) and np.isnan(ew))


# ------------------------------------------------------- weighted solves

def _weight_maps(g, w):
    """(map, the values the oracle is given) for an f64, an f32 and a
    Constant weight map."""
    w32 = w.astype(np.float32)
    return [(gt_hip.DenseProp.from_vec(g.graph_id, w.copy()), w),
            (gt_hip.DenseProp.from_vec(g.graph_id, w32.copy()), w32.astype(np.float64)),
            (gt_hip.Constant(2.5), np.full(len(w), 2.5))]


def test_weighted_fp32_form_on_block_tiles(gpu):
    """The weighted FP32 form on block tiles (weights in block order), with
    and without a personalisation, against the oracle at the FP32 form's bar
    and against the wave engine."""
    n = 1 << 14
    s, t = oracle.gen_rmat(14, 16 * n, 0.57, 0.19, 0.19, 62)
    rng = np.random.default_rng(12)
    w = 1.0 - 0.9 * rng.random(len(s))
    p = rng.random(n) + 0.05
    p /= p.sum()  # a larger sum makes the iteration diverge (dangling mass times sum p)
    g = adj(n, s, t)
    dev = g.device()
    for wm, wv in _weight_maps(g, w):
        for pers in (None, p):
            pp = None if pers is None else gt_hip.DenseProp.from_vec(g.graph_id, pers.copy())
            x, it = mixed(g, weight=wm, pers=pp)
            assert dev.layout()["block_tiles"] > 0
            dev.set_tuning(block_tile_nnz=-1)
            pp = None if pers is None else gt_hip.DenseProp.from_vec(g.graph_id, pers.copy())
            xw, itw = mixed(g, weight=wm, pers=pp)
            dev.set_tuning(block_tile_nnz=0)
            want, _ = oracle.Adj(n, s, t).pagerank(weight=wv, pers=pers, epsilon=1e-12,
                                                   max_iter=1000, threads=8)
            scale = max(1.0, float(np.abs(want).sum()))
            assert np.abs(x - want).sum() <= 1e-6 * scale
            assert abs(it - itw) <= 1 and np.abs(x - xw).sum() <= 2e-7 * scale


def test_weighted_fp64_on_block_tiles(gpu):
    """The weighted FP64 drop-in on block tiles against the oracle at the
    drop-in's bar, on default and small tiles, and against the wave engine
    (block_f64 = -1)."""
    n = 1 << 14
    s, t = oracle.gen_rmat(14, 16 * n, 0.57, 0.19, 0.19, 63)
    rng = np.random.default_rng(13)
    w = 0.1 + 0.9 * rng.random(len(s))
    g = adj(n, s, t)
    dev = g.device()
    for tile in (0, 1024):
        for wm, wv in _weight_maps(g, w):
            dev.set_tuning(block_f64=1, block_tile_nnz=tile, block_long_min=16 if tile else 0)
            x, it = pr64(g, weight=wm)
            assert dev.layout()["block_tiles"] > 0
            dev.set_tuning(block_f64=-1, block_tile_nnz=0, block_long_min=0)
            xw, itw = pr64(g, weight=wm)
            dev.set_tuning(block_f64=0)
            want, it_o = oracle.Adj(n, s, t).pagerank(weight=wv, epsilon=1e-12, max_iter=1000,
                                                      threads=8)
            parity(x, it, want, it_o, 1e-12)
            assert abs(it - itw) <= 1
            if it == itw:
                assert np.abs(x - xw).sum() <= 1e-12


def test_unit_weights_give_the_unweighted_result_bitwise(gpu):
    """centrality.rs's unity fast path must agree with the general sum: on
    block tiles, unit weights (an f64, an f32 and a Constant map) give the
    Unity solve bit for bit, for the FP64 drop-in and the FP32 form."""
    n = 1 << 13
    s, t = oracle.gen_rmat(13, 16 * n, 0.57, 0.19, 0.19, 64)
    g = adj(n, s, t)
    dev = g.device()
    ones = [gt_hip.DenseProp.from_vec(g.graph_id, np.ones(len(s))),
            gt_hip.DenseProp.from_vec(g.graph_id, np.ones(len(s), np.float32)),
            gt_hip.Constant(1.0)]
    dev.set_tuning(block_f64=1)
    for solve in (pr64, mixed):
        a, ia = solve(g)
        assert dev.layout()["block_tiles"] > 0
        for wm in ones:
            b, ib = solve(g, weight=wm)
            assert dev.layout()["block_tiles"] > 0
            assert ia == ib and np.array_equal(a.view(np.uint64), b.view(np.uint64))
    dev.set_tuning(block_f64=0)


def test_negative_weight_runs_on_the_wave_engine(gpu):
    """The block sweeps bound their sums by the iterate's mass, which a
    negative weight breaks: such a solve runs on the wave engine, bit for bit
    as with the block sweeps turned off."""
    n = 1 << 12
    s, t = oracle.gen_rmat(12, 16 * n, 0.57, 0.19, 0.19, 65)
    w = np.random.default_rng(14).uniform(0.5, 1.5, len(s))
    w[7] = -0.25
    g = adj(n, s, t)
    dev = g.device()
    dev.set_tuning(block_f64=1)
    for solve in (pr64, mixed):
        x, it = solve(g, weight=gt_hip.DenseProp.from_vec(g.graph_id, w.copy()))
        assert dev.layout()["block_tiles"] == 0
        dev.set_tuning(block_tile_nnz=-1)
        xw, itw = solve(g, weight=gt_hip.DenseProp.from_vec(g.graph_id, w.copy()))
        dev.set_tuning(block_tile_nnz=0)
        assert it == itw and np.array_equal(x.view(np.uint64), xw.view(np.uint64))
    want, it_o = oracle.Adj(n, s, t).pagerank(weight=w, epsilon=1e-12, max_iter=1000, threads=8)
    x, it = pr64(g, weight=gt_hip.DenseProp.from_vec(g.graph_id, w.copy()))
    parity(x, it, want, it_o, 1e-12)
    dev.set_tuning(block_f64=0)


# --------------------------------------------------------------- adjoint

@pytest.mark.parametrize("precision,tol", [("f64", 1e-9), ("f32", 2e-4)])
def test_adjoint_on_block_tiles(gpu, precision, tol):
    """The adjoint's transposed sweeps and its mu pass on block tiles (the
    default), weighted and unweighted, directed and reversed, on default and
    1,024-entry tiles: against the oracle's FP64 adjoint and against the wave
    engine (block_tile_nnz = -1)."""
    n = 1 << 13
    s, t = oracle.gen_rmat(13, 16 * n, 0.57, 0.19, 0.19, 66)
    rng = np.random.default_rng(15)
    w = (1.0 - 0.9 * rng.random(len(s))).astype(np.float32).astype(np.float64)
    c = rng.standard_normal(n)
    p = rng.random(n) + 0.1
    p /= p.sum()
    g = adj(n, s, t)
    dev = g.device()
    o = oracle.Adj(n, s, t)
    for weighted, reversed_, tile, pers in ((True, False, 0, None), (False, False, 0, p),
                                            (True, True, 1024, p), (False, True, 1024, None)):
        wv = w if weighted else None
        x, _ = o.pagerank(weight=wv, pers=pers, epsilon=1e-15, max_iter=10_000,
                          reversed=reversed_, threads=8)
        gw_o, gp_o, gd_o, _, _ = o.pagerank_grad(x, c, weight=wv, pers=pers, epsilon=1e-13,
                                                 reversed=reversed_)
        view = g.reversed() if reversed_ else g
        wm = gt_hip.DenseProp.from_vec(g.graph_id, w) if weighted else gt_hip.Unity.NEW
        pm = None if pers is None else gt_hip.DenseProp.from_vec(g.graph_id, pers)
        got = {}
        for blk in (tile, -1):
            dev.set_tuning(block_tile_nnz=blk, block_long_min=16 if blk > 0 else 0)
            got[blk] = gt_hip.pagerank_grad(view, wm, pm, 0.85, params(n, 1e-13, 0), x, c,
                                            precision=precision)
            assert (dev.layout()["block_tiles"] > 0) == (blk != -1), (blk, dev.layout())
        dev.set_tuning(block_tile_nnz=0, block_long_min=0)
        for gw, gp, gd, _ in got.values():
            assert np.abs(gw - gw_o).max() <= tol * np.abs(gw_o).max()
            assert np.abs(gp - gp_o).max() <= tol * np.abs(gp_o).max()
            assert abs(gd - gd_o) <= tol * max(1.0, abs(gd_o))
        if precision == "f64":
            assert abs(got[tile][3] - got[-1][3]) <= 1


def test_adjoint_negative_weight_runs_on_the_wave_engine(gpu):
    """A negative weight breaks the fixed point's bound: the adjoint runs on
    the wave engine, bit for bit as with the block tiles turned off."""
    n = 1 << 11
    s, t = oracle.gen_rmat(11, 16 * n, 0.57, 0.19, 0.19, 67)
    rng = np.random.default_rng(16)
    w = rng.uniform(0.5, 1.5, len(s))
    w[3] = -0.125
    c = rng.standard_normal(n)
    x, _ = oracle.Adj(n, s, t).pagerank(weight=w, epsilon=1e-15, max_iter=10_000)
    g = adj(n, s, t)
    dev = g.device()
    wm = gt_hip.DenseProp.from_vec(g.graph_id, w)
    a = gt_hip.pagerank_grad(g, wm, None, 0.85, params(n, 1e-12, 0), x, c, precision="f64")
    assert dev.layout()["block_tiles"] == 0
    dev.set_tuning(block_tile_nnz=-1)
    b = gt_hip.pagerank_grad(g, wm, None, 0.85, params(n, 1e-12, 0), x, c, precision="f64")
    dev.set_tuning(block_tile_nnz=0)
    assert a[3] == b[3] and np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert a[2] == b[2]


def ev_w(g, eps, max_iter, weight):
    xm = gt_hip.DenseProp.new(g.graph_id)
    eig, it = gt_hip.eigenvector(g, weight, params(g.vertex_bound(), eps, max_iter), xm)
    return np.asarray(xm.data).copy(), eig, it


def test_weighted_eigenvector_on_block_tiles(gpu):
    """The weighted eigenvector on block tiles against the oracle at the
    drop-in's bar and against the wave engine; unit weights give the Unity
    solve bit for bit; a negative weight runs on the wave engine."""
    n, S, T = sym_rmat(13, 68)
    rng = np.random.default_rng(17)
    w = 0.1 + 0.9 * rng.random(len(S))
    g = adj(n, S, T)
    dev = g.device()
    o = oracle.Adj(n, S, T)
    for eps, mi in ((1e-30, 3), (1e-10, 1000)):
        want, eo, ito = o.eigenvector(weight=w, epsilon=eps, max_iter=mi, threads=8)
        x, eig, it = ev_w(g, eps, mi, gt_hip.DenseProp.from_vec(g.graph_id, w.copy()))
        assert dev.layout()["block_tiles"] > 0
        dev.set_tuning(block_tile_nnz=-1)
        xw, eigw, itw = ev_w(g, eps, mi, gt_hip.DenseProp.from_vec(g.graph_id, w.copy()))
        dev.set_tuning(block_tile_nnz=0)
        parity(x, it, want, ito, eps)
        assert abs(eig - eo) <= max(1e-12, eps) * eo
        assert abs(it - itw) <= 1
        if it == itw:
            assert np.abs(x - xw).sum() <= 1e-12
    a, ea, ia = ev(g, 1e-10, 1000)
    b, eb, ib = ev_w(g, 1e-10, 1000, gt_hip.DenseProp.from_vec(g.graph_id, np.ones(len(S))))
    assert ia == ib and ea == eb and np.array_equal(a.view(np.uint64), b.view(np.uint64))
    wn = w.copy()
    wn[11] = -0.5
    x, eig, it = ev_w(g, 1e-10, 1000, gt_hip.DenseProp.from_vec(g.graph_id, wn.copy()))
    assert dev.layout()["block_tiles"] == 0
    dev.set_tuning(block_tile_nnz=-1)
    xw, eigw, itw = ev_w(g, 1e-10, 1000, gt_hip.DenseProp.from_vec(g.graph_id, wn.copy()))
    dev.set_tuning(block_tile_nnz=0)
    assert it == itw and eig == eigw and np.array_equal(x.view(np.uint64), xw.view(np.uint64))
