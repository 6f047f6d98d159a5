"""CPU-side checks of the product: the C ABI library loads and exports every
declared symbol; host-side generators and the host adjacency equal the
oracle's bit for bit; the Python mirror of Plan / maps behaves like the
reference.  No device compute is called here.
"""
import re
from pathlib import Path

import numpy as np
import pytest

import gt_hip
import oracle
from graphs import EDGES, LAYOUT, Rng, thousand
from gt_hip import _lib

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    text = (ROOT / "include" / "gt_hip.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:gt_status|const char \*)\s*(gt_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared_symbols()
    assert "gt_pagerank_f64" in names and "gt_eigenvector_f64" in names
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the binding declares exactly what the header declares
    assert sorted(_lib.EXPORTED) == names


def test_probe_library_exports_its_header():
    """libgt_probe.so (the measurement library bench.py times the random-gather
    rate with) exports what include/gt_probe.h declares."""
    import ctypes
    text = (ROOT / "include" / "gt_probe.h").read_text()
    names = sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(gt_\w+)\s*\(", text, re.M)))
    assert names == ["gt_probe_error", "gt_probe_gather"]
    L = ctypes.CDLL(str(ROOT / "graph-tool-rs_amd" / "gt_hip" / "libgt_probe.so"))
    assert all(hasattr(L, n) for n in names)


def test_er_generator_matches_oracle():
    for n, m, seed in [(10_000, 100_000, 0x243F6A8885A308D3), (7, 5, 1), (100, 200_000, 42)]:
        a = gt_hip.gen_er(n, m, seed)
        b = oracle.gen_er(n, m, seed)
        assert (a[0] == b[0]).all() and (a[1] == b[1]).all()


def test_rmat_generator_matches_oracle():
    a = gt_hip.gen_rmat(16, 300_000, 0.57, 0.19, 0.19, 7)
    b = oracle.gen_rmat(16, 300_000, 0.57, 0.19, 0.19, 7)
    assert (a[0] == b[0]).all() and (a[1] == b[1]).all()


def _adj_equal(n, s, t):
    got = gt_hip.host_adjacency(n, s, t)
    want = oracle.Adj(n, s, t).export()
    for x, y in zip(got, want):
        assert x.shape == y.shape and (x == y).all()


def test_host_adjacency_layout_fixture():
    s = np.array([e[0] for e in EDGES], np.uint32)
    t = np.array([e[1] for e in EDGES], np.uint32)
    _adj_equal(4, s, t)
    g = gt_hip.AdjList.with_vertices(4)
    for a, b in EDGES:
        g.add_edge(a, b)
    for v, (entries, pos) in enumerate(LAYOUT):
        assert g.out_edges(v) == entries[:pos]
        assert g.in_edges(v) == entries[pos:]


def test_host_adjacency_matches_oracle_on_random_graphs():
    r = Rng(3)
    for n, m in [(1, 10), (5, 100), (50, 3000), (1000, 20_000)]:
        s = np.array([r.below(n) for _ in range(m)], np.uint32)
        t = np.array([r.below(n) for _ in range(m)], np.uint32)
        _adj_equal(n, s, t)
    _adj_equal(*thousand())


def test_host_adjacency_hub_heavy():
    # Many out-inserts on a hub with a non-empty in-half: long rotation runs.
    s, t = oracle.gen_rmat(12, 60_000, 0.57, 0.19, 0.19, 9)
    _adj_equal(4096, s, t)


def test_no_such_vertex():
    with pytest.raises(gt_hip.GtValueError):
        gt_hip.host_adjacency(3, np.array([0, 3], np.uint32), np.array([1, 1], np.uint32))
    g = gt_hip.AdjList.with_vertices(2)
    with pytest.raises(ValueError):
        g.add_edge(0, 2)


def test_plan_matches_reference_partition():
    p = gt_hip.Plan(1_000_001, 1000)
    assert p.chunks() == 1001
    cover = [x for k in range(p.chunks()) for x in p.range(k)]
    assert cover == list(range(1_000_001))
    lens = {len(p.range(k)) for k in range(p.chunks())}
    assert lens <= {999, 1000}
    assert gt_hip.Plan(0, 64).chunks() == 1
    with pytest.raises(ValueError):
        gt_hip.Plan(10, 0)


def test_dense_prop_grows_and_never_shrinks():
    m = gt_hip.DenseProp.from_vec(5, np.array([1.0, 2.0]))
    assert m.sized_for(5, 4) == 2
    assert m.data.tolist() == [1.0, 2.0, 0.0, 0.0]
    assert m.sized_for(5, 3) == 4
    assert len(m) == 4
    with pytest.raises(ValueError):
        m.sized_for(6, 4)


def test_device_count_without_gpu_is_not_an_error():
    assert gt_hip.device_count() >= 0


def test_chung_lu_batch_generator():
    voff, eoff, d, s, t, w = gt_hip.gen_chung_lu_batch(300, 0x5EED, n_min=64, n_max=4096)
    G = 300
    ng = np.diff(voff.astype(np.int64))
    mg = np.diff(eoff)
    assert voff[0] == 0 and len(voff) == G + 1
    assert ng.min() >= 64 and ng.max() <= 4096
    # log-uniform sizes: about half the graphs below the geometric midpoint 512
    assert 0.35 < np.mean(ng < 512) < 0.65
    assert np.array_equal(mg, np.rint(8.0 * ng).astype(np.int64))
    assert np.all((d >= 0.5) & (d <= 0.95))
    assert np.all((w > 0) & (w <= 1))
    # every edge stays inside its graph's vertex range
    gid = np.repeat(np.arange(G), mg)
    assert np.all(s >= voff[gid]) and np.all(s < voff[gid + 1])
    assert np.all(t >= voff[gid]) and np.all(t < voff[gid + 1])
    # power-law sources: the lowest vertex of a graph is drawn far more often
    # than the median one (theta_0 / theta_{n/2} = (n/2)^(1/1.1))
    big = int(np.argmax(ng))
    ls = s[eoff[big]:eoff[big + 1]] - voff[big]
    cnt = np.bincount(ls, minlength=ng[big])
    assert cnt[0] > 20 * max(1, cnt[ng[big] // 2])
    # deterministic, and graph k depends only on (seed, k)
    again = gt_hip.gen_chung_lu_batch(300, 0x5EED, n_min=64, n_max=4096)
    assert np.array_equal(again[3], s) and np.array_equal(again[4], t)
    assert np.array_equal(again[5], w)
    few = gt_hip.gen_chung_lu_batch(10, 0x5EED, n_min=64, n_max=4096)
    assert np.array_equal(few[0], voff[:11])
    assert np.array_equal(few[3], s[:eoff[10]])


def test_device_memory_limit_api():
    """The limit round-trips without a device; nothing is held before a graph
    is built (allocation under the limit is exercised on the GPU)."""
    before = gt_hip.device_memory_in_use()
    assert before >= 0
    gt_hip.set_device_memory_limit(1 << 20)
    gt_hip.set_device_memory_limit(0)
    with pytest.raises(gt_hip.GtValueError):
        _lib.call("gt_device_memory_in_use", None)
