"""PageRank as a differentiable torch operation (extension).

``pagerank(graph, weight, pers=None, damping=0.85, ...)`` returns the rank
vector as a CUDA float64 tensor whose backward solves the transposed fixed
point on the GPU (``gt_pagerank_backward``): gradients flow to the edge
weights, the personalisation vector and the damping factor.  No forward
iterate is kept -- only the solution itself -- so memory is O(n + m)
whatever the sweep count.  Tensors cross into the C ABI by device pointer, the
zero-copy path DLPack capsules (``torch.from_dlpack``) produce as well.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import torch

from . import (DeviceGraph, PowerIteration, Plan, _bind_stream, _lib, _prop, _resolve,
               Unity)
from ._lib import GT_DEVICE, GtVertexOut, call


class _PageRank(torch.autograd.Function):
    @staticmethod
    def forward(ctx, weight, pers, damping, graph, epsilon, max_iter, mixed, adj_precision):
        dev, view, gid, n, nv = _resolve(graph)
        d = float(damping)
        params = PowerIteration(epsilon, max_iter, Plan(n, 4096))
        wp = _prop(weight if weight is not None else Unity.NEW, gid)
        pp = _prop(pers, gid) if pers is not None else None
        x = torch.empty(n, dtype=torch.float64, device="cuda")
        _bind_stream(dev, x)
        out = GtVertexOut(C.c_void_p(x.data_ptr()), n, 0, GT_DEVICE, 0, gid)
        it = C.c_uint64()
        call("gt_pagerank_mixed" if mixed else "gt_pagerank_f64", dev.handle, C.byref(view),
             C.byref(wp), C.byref(pp) if pp else None, d, C.byref(params._c()), C.byref(out),
             C.byref(it))
        ctx.graph = graph
        ctx.damping = d
        ctx.params = params
        ctx.prec = adj_precision
        ctx.damping_is_tensor = torch.is_tensor(damping)
        ctx.save_for_backward(weight if weight is not None else torch.empty(0),
                              pers if pers is not None else torch.empty(0), x)
        ctx.has_w = weight is not None
        ctx.has_p = pers is not None
        ctx.sweeps = int(it.value)
        return x

    @staticmethod
    def backward(ctx, gx):
        weight, pers, x = ctx.saved_tensors
        dev, view, gid, n, nv = _resolve(ctx.graph)
        _, _, eb, _ = dev.info()
        wp = _prop(weight if ctx.has_w else Unity.NEW, gid)
        pp = _prop(pers, gid) if ctx.has_p else None
        gx = gx.contiguous().to(torch.float64)
        xp = _prop(x, gid)
        cp = _prop(gx, gid)
        need_w = ctx.has_w and ctx.needs_input_grad[0]
        need_p = ctx.has_p and ctx.needs_input_grad[1]
        gw = torch.empty(max(eb, 1), dtype=torch.float64, device="cuda") if need_w else None
        gp = torch.empty(max(n, 1), dtype=torch.float64, device="cuda") if need_p else None
        gd = C.c_double()
        it = C.c_uint64()
        _bind_stream(dev, x)
        call("gt_pagerank_backward", dev.handle, C.byref(view), C.byref(wp),
             C.byref(pp) if pp else None, ctx.damping, C.byref(ctx.params._c()), C.byref(xp),
             C.byref(cp), _lib.GT_F64 if ctx.prec == "f64" else _lib.GT_F32,
             C.c_void_p(gw.data_ptr()) if gw is not None else None, _lib.GT_F64, GT_DEVICE,
             C.c_void_p(gp.data_ptr()) if gp is not None else None, C.byref(gd), C.byref(it))
        grad_w = gw[:eb].to(weight.dtype) if gw is not None else None
        grad_p = gp[:n].to(pers.dtype) if gp is not None else None
        grad_d = (torch.tensor(gd.value, dtype=torch.float64, device="cuda")
                  if ctx.damping_is_tensor and ctx.needs_input_grad[2] else None)
        return grad_w, grad_p, grad_d, None, None, None, None, None


def pagerank(graph, weight: Optional[torch.Tensor] = None, pers: Optional[torch.Tensor] = None,
             damping=0.85, epsilon: float = 1e-10, max_iter: int = 0, mixed: bool = False,
             adjoint_precision: str = "f64") -> torch.Tensor:
    """Differentiable PageRank on a DeviceGraph / AdjList / View.

    weight: CUDA tensor [edge_bound] (float32 or float64) or None (unity);
    pers: CUDA float64 [n] or None (uniform); damping: float or 0-dim tensor."""
    return _PageRank.apply(weight, pers, damping, graph, epsilon, max_iter, mixed,
                           adjoint_precision)
