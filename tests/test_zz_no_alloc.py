"""The no-allocation PageRank entry point (gt_pagerank_f64_noalloc), called
through the C ABI, against gt_pagerank_f64 on the same graph, on two calls in
a row, the second on a smaller graph (the shared scratch keeps the first
call's contents): bit for bit against gt_pagerank_f64 on the wave engine,
whose kernels it shares (block_f64 = -1); at the drop-in's bar against its
default block-tile sweeps; and against the oracle at the drop-in's bar.
Also its refusals: an unprepared graph, a graph over the scratch size, a
host rank map, weights."""
import ctypes as C

import numpy as np
import pytest

import gt_hip
import oracle
from gt_hip import _lib

pytestmark = pytest.mark.gpu


def _device_graph(n, s, t, gid):
    return gt_hip.DeviceGraph.upload(gid, n, len(s), *gt_hip.host_adjacency(n, s, t))


def _call(fn, dev, n, gid, eps, pers=None, kind=_lib.GT_VIEW_DIRECTED, loc=_lib.GT_DEVICE,
          weight=None):
    import torch
    view = _lib.GtView(kind, 0, None, 0, None, 0)
    unity = weight if weight is not None else _lib.GtProp(_lib.GT_MAP_UNITY, 0, 0, 0, 1.0, None,
                                                          0, 0)
    params = _lib.GtPowerIteration(eps, 1000, _lib.GtPlan(n, 64))
    if loc == _lib.GT_DEVICE:
        rank = torch.zeros(n, dtype=torch.float64, device="cuda")
        out = _lib.GtVertexOut(C.c_void_p(rank.data_ptr()), n, 0, _lib.GT_DEVICE, 0, gid)
    else:
        rank = np.zeros(n)
        out = _lib.GtVertexOut(rank.ctypes.data_as(C.c_void_p), n, 0, _lib.GT_HOST, 0, gid)
    it = C.c_uint64()
    dev.set_stream(None)
    _lib.call(fn, dev.handle, C.byref(view), C.byref(unity), C.byref(pers) if pers else None,
              0.85, C.byref(params), C.byref(out), C.byref(it))
    x = rank.cpu().numpy() if loc == _lib.GT_DEVICE else rank
    return x, int(it.value)


def _case(scale, seed, gid):
    n = 1 << scale
    s, t = oracle.gen_rmat(scale, 16 * n, 0.57, 0.19, 0.19, seed)
    return n, s, t, _device_graph(n, s, t, gid)


def test_noalloc_matches_f64_twice_shrinking(gpu):
    import torch
    cases = [_case(17, 3, 901), _case(12, 4, 902)]  # second call on a smaller graph
    for kind in (_lib.GT_VIEW_DIRECTED, _lib.GT_VIEW_REVERSED):
        for n, s, t, dev in cases:
            p = torch.from_numpy(np.random.default_rng(n).uniform(0.1, 1.0, n)).cuda()
            pers = _lib.GtProp(_lib.GT_MAP_DENSE, _lib.GT_F64, _lib.GT_DEVICE, 0, 0.0,
                               C.c_void_p(p.data_ptr()), n, 0)
            for pm in (None, pers):
                dflt, it_d = _call("gt_pagerank_f64", dev, n, dev.graph_id, 1e-12, pm, kind)
                dev.set_tuning(block_f64=-1)
                want, it_w = _call("gt_pagerank_f64", dev, n, dev.graph_id, 1e-12, pm, kind)
                dev.set_tuning(block_f64=0)
                got, it_g = _call("gt_pagerank_f64_noalloc", dev, n, dev.graph_id, 1e-12, pm,
                                  kind)
                assert it_g == it_w
                assert (got.view(np.uint64) == want.view(np.uint64)).all()
                # An unnormalised personalisation makes the iteration grow
                # without bound: past the fixed point's range the default
                # solve is the wave engine's, bit for bit (NaN where it has
                # NaN).
                if not np.array_equal(got, dflt, equal_nan=True):
                    scale = max(1.0, float(np.abs(dflt).sum()))
                    assert abs(it_g - it_d) <= 1
                    assert np.abs(got - dflt).sum() <= 1e-12 * scale or abs(it_g - it_d) == 1
            ref, it_o = oracle.Adj(n, s, t).pagerank(epsilon=1e-12, max_iter=1000, threads=8,
                                                     reversed=kind == _lib.GT_VIEW_REVERSED)
            got, it_g = _call("gt_pagerank_f64_noalloc", dev, n, dev.graph_id, 1e-12, None, kind)
            assert abs(it_g - it_o) <= 1
            assert np.abs(got - ref).sum() <= 1e-12 or abs(it_g - it_o) == 1


def test_noalloc_refusals(gpu):
    n, s, t, dev = _case(10, 5, 903)
    with pytest.raises(gt_hip.GtError, match="schedule is not built"):
        _call("gt_pagerank_f64_noalloc", dev, n, dev.graph_id, 1e-10)
    _call("gt_pagerank_f64", dev, n, dev.graph_id, 1e-10)
    with pytest.raises(gt_hip.GtError, match="device memory"):
        _call("gt_pagerank_f64_noalloc", dev, n, dev.graph_id, 1e-10, loc=_lib.GT_HOST)
    w = _lib.GtProp(_lib.GT_MAP_CONSTANT, 0, 0, 0, 2.0, None, 0, 0)
    with pytest.raises(gt_hip.GtError, match="Unity"):
        _call("gt_pagerank_f64_noalloc", dev, n, dev.graph_id, 1e-10, weight=w)
    # Over the scratch: 3 M vertices (edgeless, prepared by one ordinary call).
    big = 3_000_000
    e = np.zeros(0, np.uint32)
    bdev = _device_graph(big, e, e, 904)
    _call("gt_pagerank_f64", bdev, big, 904, 1e-10)
    with pytest.raises(gt_hip.GtError, match="bytes of scratch"):
        _call("gt_pagerank_f64_noalloc", bdev, big, 904, 1e-10)
