"""GPU adjoint against the FP64 CPU adjoint of the oracle (itself pinned by
finite differences), against central finite differences of the GPU forward
solve, the row-scale identity, and through torch.autograd."""
import numpy as np
import pytest

import gt_hip
import oracle
from graphs import Rng

pytestmark = pytest.mark.gpu


def _problem(scale, seed):
    n = 1 << scale
    s, t = oracle.gen_rmat(scale, 16 * n, 0.57, 0.19, 0.19, seed)
    rng = np.random.default_rng(seed)
    w = (1.0 - 0.9 * rng.random(len(s))).astype(np.float32).astype(np.float64)
    c = rng.standard_normal(n)
    return n, s, t, w, c


def _params(n, eps, max_iter=0):
    return gt_hip.PowerIteration(eps, max_iter, gt_hip.Plan(n, 64))


@pytest.mark.parametrize("precision,tol", [("f64", 1e-9), ("f32", 2e-4)])
def test_backward_matches_oracle_adjoint(gpu, precision, tol):
    n, s, t, w, c = _problem(12, 5)
    g = gt_hip.AdjList.with_vertices(n)
    g.add_edges(s, t)
    o = oracle.Adj(n, s, t)
    x, _ = o.pagerank(weight=w, epsilon=1e-15, max_iter=10_000)
    gw_o, gp_o, gd_o, _, it_o = o.pagerank_grad(x, c, weight=w, epsilon=1e-13)
    wm = gt_hip.DenseProp.from_vec(g.graph_id, w)
    gw, gp, gd, it = gt_hip.pagerank_grad(g, wm, None, 0.85, _params(n, 1e-13), x, c,
                                          precision=precision)
    scale = np.abs(gw_o).max()
    assert np.abs(gw - gw_o).max() <= tol * scale
    assert np.abs(gp - gp_o).max() <= tol * np.abs(gp_o).max()
    assert abs(gd - gd_o) <= tol * max(1.0, abs(gd_o))
    if precision == "f64":
        # At epsilon 1e-13 the last deltas are rounding noise; compare sweep
        # counts where the L1 change is still well above it.  (An FP32 adjoint
        # reaches its own rounding floor first, so its count is not compared.)
        _, _, _, it = gt_hip.pagerank_grad(g, wm, None, 0.85, _params(n, 1e-8), x, c)
        _, _, _, _, it_o = o.pagerank_grad(x, c, weight=w, epsilon=1e-8)
        assert abs(it - it_o) <= 1


def test_fp64_backward_matches_oracle_adjoint_scale16(gpu):
    """The FP64 GPU adjoint against the oracle's FP64 adjoint on a scale-16
    R-MAT instance (65,536 vertices, 1,048,576 weighted edges): every edge
    weight's, the personalisation's and the damping's gradient."""
    n, s, t, w, c = _problem(16, 23)
    g = gt_hip.AdjList.with_vertices(n)
    g.add_edges(s, t)
    o = oracle.Adj(n, s, t)
    x, _ = o.pagerank(weight=w, epsilon=1e-15, max_iter=10_000, threads=8)
    gw_o, gp_o, gd_o, _, it_o = o.pagerank_grad(x, c, weight=w, epsilon=1e-13)
    wm = gt_hip.DenseProp.from_vec(g.graph_id, w)
    gw, gp, gd, it = gt_hip.pagerank_grad(g, wm, None, 0.85, _params(n, 1e-13), x, c,
                                          precision="f64")
    assert np.abs(gw - gw_o).max() <= 1e-9 * np.abs(gw_o).max()
    assert np.abs(gp - gp_o).max() <= 1e-9 * np.abs(gp_o).max()
    assert abs(gd - gd_o) <= 1e-9 * max(1.0, abs(gd_o))
    # Sweep counts where the L1 change is still well above rounding noise.
    _, _, _, it = gt_hip.pagerank_grad(g, wm, None, 0.85, _params(n, 1e-8), x, c)
    _, _, _, _, it_o = o.pagerank_grad(x, c, weight=w, epsilon=1e-8)
    assert abs(it - it_o) <= 1


def test_backward_on_reversed_view_matches_oracle(gpu):
    # Reversed view: the transposed sweep pulls over the in-halves and every
    # edge's gradient row is its target.
    n, s, t, w, c = _problem(11, 9)
    g = gt_hip.AdjList.with_vertices(n)
    g.add_edges(s, t)
    o = oracle.Adj(n, s, t)
    x, _ = o.pagerank(weight=w, epsilon=1e-15, max_iter=10_000, reversed=True)
    gw_o, gp_o, gd_o, _, _ = o.pagerank_grad(x, c, weight=w, epsilon=1e-13, reversed=True)
    wm = gt_hip.DenseProp.from_vec(g.graph_id, w)
    gw, gp, gd, _ = gt_hip.pagerank_grad(g.reversed(), wm, None, 0.85, _params(n, 1e-13), x, c)
    assert np.abs(gw - gw_o).max() <= 1e-9 * np.abs(gw_o).max()
    assert np.abs(gp - gp_o).max() <= 1e-9 * np.abs(gp_o).max()
    assert abs(gd - gd_o) <= 1e-9 * max(1.0, abs(gd_o))


@pytest.mark.parametrize("reversed_", [False, True])
def test_f32_hot_slice_on_and_off(gpu, reversed_):
    """The FP32 transposed sweep runs over the relabelled orientation and keeps
    the most gathered entries of lambda in LDS (the default).  With the slice
    off the gradients agree with it, and both with the oracle adjoint at the
    FP32 bar: weighted and personalised, on a graph inside the slice (scale
    10) and one larger than it (scale 16)."""
    for scale, seed in ((10, 21), (16, 22)):
        n, s, t, w, c = _problem(scale, seed)
        p = np.random.default_rng(seed).uniform(0.0, 2.0, n)
        p /= p.sum()
        g = gt_hip.AdjList.with_vertices(n)
        g.add_edges(s, t)
        o = oracle.Adj(n, s, t)
        x, _ = o.pagerank(weight=w, pers=p, epsilon=1e-15, max_iter=10_000, reversed=reversed_)
        gw_o, gp_o, gd_o, _, _ = o.pagerank_grad(x, c, weight=w, pers=p, epsilon=1e-13,
                                                 reversed=reversed_)
        wm = gt_hip.DenseProp.from_vec(g.graph_id, w)
        pm = gt_hip.DenseProp.from_vec(g.graph_id, p)
        view = g.reversed() if reversed_ else g
        got = {}
        for hot in (0, -1):
            g.device().set_tuning(sweep_hot_kb=hot)
            got[hot] = gt_hip.pagerank_grad(view, wm, pm, 0.85, _params(n, 1e-13), x, c,
                                            precision="f32")
        g.device().set_tuning(sweep_hot_kb=0)
        sw, sp = np.abs(gw_o).max(), np.abs(gp_o).max()
        for gw, gp, gd, _ in got.values():
            assert np.abs(gw - gw_o).max() <= 2e-4 * sw, (scale, reversed_)
            assert np.abs(gp - gp_o).max() <= 2e-4 * sp
            assert abs(gd - gd_o) <= 2e-4 * max(1.0, abs(gd_o))
        # Only the order of the FP64 partial folds differs between the two.
        assert np.abs(got[0][0] - got[-1][0]).max() <= 1e-5 * sw


def test_backward_matches_finite_differences_on_gpu(gpu):
    n, s, t, w, c = _problem(10, 11)
    g = gt_hip.AdjList.with_vertices(n)
    g.add_edges(s, t)
    p = np.random.default_rng(3).random(n) + 0.5
    p /= p.sum()

    def loss(wv, pv, d):
        r = gt_hip.DenseProp.new(g.graph_id)
        gt_hip.pagerank_with(g, gt_hip.DenseProp.from_vec(g.graph_id, wv),
                             gt_hip.DenseProp.from_vec(g.graph_id, pv), d, _params(n, 1e-15, 100_000), r)
        return float(c @ r.data), r.data

    L0, x = loss(w, p, 0.85)
    gw, gp, gd, _ = gt_hip.pagerank_grad(g, gt_hip.DenseProp.from_vec(g.graph_id, w),
                                         gt_hip.DenseProp.from_vec(g.graph_id, p), 0.85,
                                         _params(n, 1e-14), x, c)
    h = 1e-4
    rng = np.random.default_rng(7)
    scale = np.abs(gw).max()
    for e in rng.choice(len(w), 40, replace=False):
        wp, wmn = w.copy(), w.copy()
        wp[e] += h
        wmn[e] -= h
        fd = (loss(wp, p, 0.85)[0] - loss(wmn, p, 0.85)[0]) / (2 * h)
        assert abs(fd - gw[e]) <= 1e-6 * scale + 1e-5 * abs(fd), (e, fd, gw[e])
    fd = (loss(w, p, 0.85 + h)[0] - loss(w, p, 0.85 - h)[0]) / (2 * h)
    assert abs(fd - gd) <= 1e-5 * max(1.0, abs(fd))


def test_row_scale_identity_on_gpu(gpu):
    n, s, t, w, c = _problem(12, 2)
    g = gt_hip.AdjList.with_vertices(n)
    g.add_edges(s, t)
    wm = gt_hip.DenseProp.from_vec(g.graph_id, w)
    r = gt_hip.DenseProp.new(g.graph_id)
    gt_hip.pagerank(g, wm, 0.85, _params(n, 1e-13), r)
    gw, _, _, _ = gt_hip.pagerank_grad(g, wm, None, 0.85, _params(n, 1e-13), r.data, c)
    per_src = np.zeros(n)
    np.add.at(per_src, s.astype(np.int64), w * gw)
    assert np.abs(per_src).max() <= 1e-10 * np.abs(w * gw).max()


def test_torch_autograd(gpu):
    import torch
    from gt_hip import autograd
    n, s, t, w, c = _problem(11, 4)
    g = gt_hip.AdjList.with_vertices(n)
    g.add_edges(s, t)
    wt = torch.tensor(w, dtype=torch.float64, device="cuda", requires_grad=True)
    d = torch.tensor(0.85, dtype=torch.float64, device="cuda", requires_grad=True)
    pt = torch.full((n,), 1.0 / n, dtype=torch.float64, device="cuda", requires_grad=True)
    x = autograd.pagerank(g, wt, pt, d, epsilon=1e-13)
    L = (torch.tensor(c, device="cuda") * x).sum()
    L.backward()
    o = oracle.Adj(n, s, t)
    xo, _ = o.pagerank(weight=w, pers=np.full(n, 1.0 / n), epsilon=1e-15, max_iter=10_000)
    gw_o, gp_o, gd_o, _, _ = o.pagerank_grad(xo, c, weight=w, pers=np.full(n, 1.0 / n),
                                             epsilon=1e-13)
    assert np.abs(x.detach().cpu().numpy() - xo).sum() <= 1e-11
    assert np.abs(wt.grad.cpu().numpy() - gw_o).max() <= 1e-8 * np.abs(gw_o).max()
    assert np.abs(pt.grad.cpu().numpy() - gp_o).max() <= 1e-8 * np.abs(gp_o).max()
    assert abs(d.grad.item() - gd_o) <= 1e-8 * max(1.0, abs(gd_o))
