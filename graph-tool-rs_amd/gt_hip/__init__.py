"""gt_hip -- the PageRank / eigenvector-centrality path of graph-tool-rs on MI355X.

The call surface mirrors gt-algo's centrality module
(crates/gt-algo/src/centrality.rs:79-99, 234-497) and the gt-core types it is
written against:

* :class:`Plan` -- ``Plan::new(n, grain)`` (crates/gt-core/src/par/plan.rs:55-101);
* :class:`PowerIteration` -- ``{epsilon, max_iter, plan}`` (centrality.rs:79-99);
* :class:`Unity`, :class:`Constant`, :class:`DenseProp` -- the property maps the
  path reads (crates/gt-core/src/prop/dense.rs:88-127, 373-460);
* :class:`AdjList` -- ``with_vertices`` / ``add_edge`` (crates/gt-core/src/adj/list.rs:59-97,
  258-280) with the four directed views ``dispatch_bidi`` accepts
  (crates/gt-py/src/dispatch.rs:55-60, 315-336);
* :func:`pagerank`, :func:`pagerank_with`, :func:`eigenvector` with the
  reference's signatures and return values.

Every computation runs in libgt_hip.so on the GPU; this module only sizes maps,
resolves views and translates errors (ValueError for caller errors, as
crates/gt-py/src/module.rs:78-101 maps them; RuntimeError for device failures).
"""
from __future__ import annotations

import ctypes as C
import itertools
import math
from dataclasses import dataclass
from typing import Optional, Tuple

import numpy as np

from . import _lib
from ._lib import (GT_DEVICE, GT_F32, GT_F64, GT_HOST, GT_MAP_CONSTANT, GT_MAP_DENSE,
                   GT_MAP_UNITY, GtError, GtProp, GtPowerIteration, GtValueError,
                   GtVertexOut, GtView, call, lib)

__all__ = [
    "Plan", "PowerIteration", "Unity", "Constant", "DenseProp", "AdjList", "View",
    "DeviceGraph", "pagerank", "pagerank_with", "pagerank_mixed", "eigenvector",
    "gen_er", "gen_rmat", "host_adjacency", "GtError", "GtValueError", "device_count",
    "set_device_memory_limit", "device_memory_in_use",
    "pagerank_uneven", "pagerank_uneven_grad", "gen_chung_lu_batch",
    "Document", "read_gt", "read_edge_list",
    "weighted_out_degree", "weighted_in_degree", "weighted_degree", "weighted_out_degrees",
    "weighted_in_degrees", "weighted_degrees", "out_degrees", "degree_histogram",
]

_u32p = C.POINTER(C.c_uint32)


def _ptr(a: np.ndarray, t=_u32p):
    return a.ctypes.data_as(t)


def set_device_memory_limit(nbytes: int) -> None:
    """Cap the device memory the library's buffers may hold (0: no cap); an
    allocation past it raises GtError with status GT_ERR_OOM."""
    call("gt_set_device_memory_limit", int(nbytes))


def device_memory_in_use() -> int:
    """Device bytes the library's buffers hold now."""
    v = C.c_uint64()
    call("gt_device_memory_in_use", C.byref(v))
    return int(v.value)


def device_count() -> int:
    n = C.c_int32(0)
    call("gt_device_count", C.byref(n))
    return int(n.value)


# --------------------------------------------------------------------- Plan

class Plan:
    """``Plan::new(n, grain)``: ``max(1, ceil(n / grain))`` balanced chunks."""

    def __init__(self, n: int, grain: int):
        if grain <= 0:
            raise ValueError("Plan::new: grain must be non-zero")
        self._n = int(n)
        self._chunks = max(1, -(-self._n // int(grain)))

    @classmethod
    def from_parts(cls, n: int, chunks: int) -> "Plan":
        p = cls.__new__(cls)
        p._n, p._chunks = int(n), int(chunks)
        return p

    def len(self) -> int:
        return self._n

    def chunks(self) -> int:
        return self._chunks

    def range(self, k: int) -> range:
        if not 0 <= k < self._chunks:
            raise IndexError(f"Plan::range: chunk {k} out of range (chunks = {self._chunks})")
        base, rem = divmod(self._n, self._chunks)
        start = k * base + min(k, rem)
        return range(start, start + base + (1 if k < rem else 0))

    def __eq__(self, o) -> bool:
        return isinstance(o, Plan) and (self._n, self._chunks) == (o._n, o._chunks)

    def __repr__(self) -> str:
        return f"Plan(n={self._n}, chunks={self._chunks})"


@dataclass
class PowerIteration:
    """Convergence threshold on the per-sweep L1 change, sweep cap (0 = none)
    and the plan that must cover the vertex index space."""
    epsilon: float
    max_iter: int
    plan: Plan

    def _c(self) -> GtPowerIteration:
        return GtPowerIteration(float(self.epsilon), int(self.max_iter),
                                _lib.GtPlan(self.plan.len(), self.plan.chunks()))


# ---------------------------------------------------------- property maps

class Unity:
    """The all-ones edge map (``Unity``; ``IS_UNITY`` selects the integer
    out-degree)."""
    NEW: "Unity"


Unity.NEW = Unity()


class Constant:
    """A map with one value everywhere (``Constant::new``)."""

    def __init__(self, value: float):
        self.value = float(value)


def _is_torch_cuda(x) -> bool:
    return hasattr(x, "is_cuda") and bool(getattr(x, "is_cuda"))


class DenseProp:
    """A dense f64 (or f32) map owned by one graph.

    ``data`` is a numpy array (host memory) or a CUDA torch tensor (device
    memory).  ``sized_for(bound)`` grows it to exactly ``bound`` and never
    shrinks it (crates/gt-core/src/prop/dense.rs:88-127).
    """

    def __init__(self, graph_id: int, data=None, dtype=np.float64):
        self.graph_id = int(graph_id)
        self.data = np.zeros(0, dtype=dtype) if data is None else data

    @classmethod
    def new(cls, graph_id: int) -> "DenseProp":
        return cls(graph_id)

    @classmethod
    def from_vec(cls, graph_id: int, data) -> "DenseProp":
        if not _is_torch_cuda(data):
            data = np.ascontiguousarray(data)
            if data.dtype not in (np.float64, np.float32):
                data = data.astype(np.float64)
        return cls(graph_id, data)

    def __len__(self) -> int:
        return int(self.data.shape[0])

    def as_slice(self):
        return self.data

    def sized_for(self, graph_id: int, bound: int) -> int:
        """Grow to ``bound`` (zero-filled); returns the length it had."""
        if graph_id != self.graph_id:
            raise GtValueError(_lib.GT_ERR_WRONG_GRAPH,
                               f"map belongs to graph {self.graph_id}, not {graph_id}")
        had = len(self)
        if had < bound:
            if _is_torch_cuda(self.data):
                import torch
                grown = torch.zeros(bound, dtype=self.data.dtype, device=self.data.device)
                grown[:had] = self.data
            else:
                grown = np.zeros(bound, dtype=self.data.dtype)
                grown[:had] = self.data
            self.data = grown
        return had


def _prop(m, graph_id: int, keep=None) -> Optional[GtProp]:
    """A read-only map as a gt_prop (``keep`` collects arrays to keep alive)."""
    if m is None:
        return None
    if isinstance(m, Unity):
        return GtProp(GT_MAP_UNITY, GT_F64, GT_HOST, 0, 1.0, None, 0, 0)
    if isinstance(m, Constant):
        return GtProp(GT_MAP_CONSTANT, GT_F64, GT_HOST, 0, m.value, None, 0, 0)
    gid = 0
    if isinstance(m, DenseProp):
        gid, m = m.graph_id, m.data
    if _is_torch_cuda(m):
        import torch
        if not m.is_contiguous():
            raise GtValueError(_lib.GT_ERR_INVALID_ARG, "device map must be contiguous")
        dt = {torch.float64: GT_F64, torch.float32: GT_F32}.get(m.dtype)
        if dt is None:
            raise GtValueError(_lib.GT_ERR_INVALID_ARG, "map dtype must be float64 or float32")
        return GtProp(GT_MAP_DENSE, dt, GT_DEVICE, 0, 0.0, C.c_void_p(m.data_ptr()),
                      int(m.numel()), gid)
    a = np.ascontiguousarray(m)
    if a.dtype not in (np.float64, np.float32):
        a = a.astype(np.float64)
    if keep is not None:
        keep.append(a)
    dt = GT_F64 if a.dtype == np.float64 else GT_F32
    return GtProp(GT_MAP_DENSE, dt, GT_HOST, 0, 0.0, a.ctypes.data_as(C.c_void_p),
                  int(a.shape[0]), gid)


# ------------------------------------------------------------------ graphs

_graph_ids = itertools.count(1)


def host_adjacency(n: int, src: np.ndarray, tgt: np.ndarray):
    """The flattened in-halves / out-halves ``add_edge(src[i], tgt[i])`` builds."""
    src = np.ascontiguousarray(src, dtype=np.uint32)
    tgt = np.ascontiguousarray(tgt, dtype=np.uint32)
    m = int(src.shape[0])
    in_off = np.zeros(n + 1, np.uint32)
    out_off = np.zeros(n + 1, np.uint32)
    arrs = [np.zeros(max(m, 1), np.uint32) for _ in range(4)]
    call("gt_host_adjacency", n, m, _ptr(src), _ptr(tgt), _ptr(in_off), _ptr(arrs[0]),
         _ptr(arrs[1]), _ptr(out_off), _ptr(arrs[2]), _ptr(arrs[3]))
    return in_off, arrs[0][:m], arrs[1][:m], out_off, arrs[2][:m], arrs[3][:m]


class DeviceGraph:
    """A device-resident graph handle (``gt_graph``)."""

    def __init__(self, handle: int, graph_id: int):
        self._h = C.c_void_p(handle)
        self.graph_id = graph_id

    @classmethod
    def upload(cls, graph_id: int, n: int, edge_bound: int, in_off, in_other, in_eid, out_off,
               out_other, out_eid) -> "DeviceGraph":
        arrs = [np.ascontiguousarray(a, dtype=np.uint32) for a in
                (in_off, in_other, in_eid, out_off, out_other, out_eid)]
        h = _lib.GraphP()
        call("gt_graph_upload", graph_id, n, edge_bound, *[_ptr(a) for a in arrs], C.byref(h))
        return cls(h.value, graph_id)

    @classmethod
    def from_edges(cls, n: int, src, tgt, graph_id: Optional[int] = None) -> "DeviceGraph":
        """Host adjacency (push-swap simulated on the CPU), then uploaded."""
        gid = next(_graph_ids) if graph_id is None else graph_id
        adj = host_adjacency(n, src, tgt)
        return cls.upload(gid, n, int(np.asarray(src).shape[0]), *adj)

    @classmethod
    def build(cls, n: int, src, tgt, graph_id: Optional[int] = None) -> "DeviceGraph":
        """Device builder: histogram, scan, stable scatter and push-swap replay
        on the GPU.  ``src`` / ``tgt`` are uint32 numpy arrays or CUDA tensors."""
        gid = next(_graph_ids) if graph_id is None else graph_id
        h = _lib.GraphP()
        if _is_torch_cuda(src):
            import torch
            s = src.to(torch.int32).contiguous()
            t = tgt.to(torch.int32).contiguous()
            torch.cuda.current_stream(s.device).synchronize()
            call("gt_graph_build_edges", gid, n, int(s.numel()), C.c_void_p(s.data_ptr()),
                 C.c_void_p(t.data_ptr()), GT_DEVICE, C.byref(h))
        else:
            s = np.ascontiguousarray(src, dtype=np.uint32)
            t = np.ascontiguousarray(tgt, dtype=np.uint32)
            if s.shape != t.shape:
                raise ValueError("source and target arrays differ in length")
            call("gt_graph_build_edges", gid, n, int(s.shape[0]), s.ctypes.data_as(C.c_void_p),
                 t.ctypes.data_as(C.c_void_p), GT_HOST, C.byref(h))
        return cls(h.value, gid)

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    def info(self) -> Tuple[int, int, int, int]:
        n, m, eb, gid = C.c_uint32(), C.c_uint64(), C.c_uint64(), C.c_uint64()
        call("gt_graph_info", self._h, C.byref(n), C.byref(m), C.byref(eb), C.byref(gid))
        return n.value, m.value, eb.value, gid.value

    def download(self):
        n, m, _, _ = self.info()
        out = [np.zeros(n + 1, np.uint32), np.zeros(max(m, 1), np.uint32),
               np.zeros(max(m, 1), np.uint32), np.zeros(n + 1, np.uint32),
               np.zeros(max(m, 1), np.uint32), np.zeros(max(m, 1), np.uint32)]
        call("gt_graph_download", self._h, *[_ptr(a) for a in out])
        return out[0], out[1][:m], out[2][:m], out[3], out[4][:m], out[5][:m]

    def set_stream(self, stream_ptr: Optional[int]) -> None:
        call("gt_graph_set_stream", self._h, C.c_void_p(stream_ptr or 0))

    TUNING_FIELDS = ("sweep_blocks_per_cu", "tile_rows", "short_row", "mm_tile_nnz",
                     "mm_blocks_per_cu", "sweep_hot_kb", "sliced_min_len", "sliced_share",
                     "block_tile_nnz", "block_long_min", "block_per_cu",
                     "block_f64")

    def tuning(self) -> dict:
        """Launch / schedule parameters in effect (gt_graph_get_tuning)."""
        t = _lib.GtTuning()
        call("gt_graph_get_tuning", self.handle, C.byref(t))
        return {f: int(getattr(t, f)) for f in self.TUNING_FIELDS}

    def set_tuning(self, **kw) -> dict:
        """Set some parameters (others keep their values); 0 restores a default.
        Returns the values in effect."""
        cur = self.tuning()
        for k, v in kw.items():
            if k not in self.TUNING_FIELDS:
                raise GtValueError(_lib.GT_ERR_INVALID_ARG, f"unknown tuning field {k}")
            cur[k] = int(v)
        t = _lib.GtTuning(*[cur[f] for f in self.TUNING_FIELDS])
        call("gt_graph_set_tuning", self.handle, C.byref(t))
        return self.tuning()

    def set_profiling(self, enable: bool) -> None:
        call("gt_graph_set_profiling", self._h, 1 if enable else 0)

    def sweep_stats(self, kind: int = _lib.GT_SWEEP_PAGERANK_F64) -> Tuple[float, int, int]:
        """(summed event ms, launches) of one sweep kernel kind, and the
        sweep count of the last call."""
        ms, la, sw = C.c_double(), C.c_uint64(), C.c_uint64()
        call("gt_sweep_stats", self._h, kind, C.byref(ms), C.byref(la), C.byref(sw))
        return ms.value, la.value, sw.value

    def layout(self) -> dict:
        """How the last pull-sweep solve laid out its orientation
        (gt_last_layout): slot_bits 10 for the packed tile stream, 0 for
        plain sources; the tile and split-row counts; the sliced rows."""
        lay = _lib.GtSweepLayout()
        call("gt_last_layout", self._h, C.byref(lay))
        return {k: int(getattr(lay, k)) for k, _ in _lib.GtSweepLayout._fields_}

    def close(self) -> None:
        if self._h is not None and self._h.value:
            lib().gt_graph_free(self._h)
        self._h = C.c_void_p(0)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _GraphMethods:
    """gt-py's shape of the centrality calls on a graph or one of its views
    (crates/gt-py/src/module.rs:3-18): the read-only maps are sized first (a
    graph-tool map grows on read); the output map is only checked for its
    owner -- the solver sizes it, and the length it had decides which slots
    are seeded and which warm-start (centrality.rs:198-205, 299).  The call
    runs outside the GIL (ctypes releases it around every library call);
    errors arrive as
    ValueError (a map or argument) or RuntimeError (the device).  The four
    directed view kinds are served; the undirected ones are refused with the
    NotDirected message (crates/gt-py/src/dispatch.rs:315-336).  Without a
    map argument the result goes to the graph's own vertex map
    ``vp["pagerank"]`` / ``vp["eigenvector"]``."""

    def _graph(self) -> "AdjList":
        raise NotImplementedError

    def new_vertex_property(self, value_type: str = "double") -> "DenseProp":
        """A new double vertex map of this graph."""
        return self._graph()._new_map(value_type)

    def new_edge_property(self, value_type: str = "double") -> "DenseProp":
        """A new double edge map of this graph."""
        return self._graph()._new_map(value_type)

    def _owned(self, m, what: str) -> None:
        if isinstance(m, DenseProp) and m.graph_id != self._graph().graph_id:
            raise GtValueError(_lib.GT_ERR_WRONG_GRAPH, f"{what} map does not belong to this graph")

    def _sized(self, m, bound: int, what: str):
        if isinstance(m, DenseProp):
            self._owned(m, what)
            m.sized_for(m.graph_id, bound)
        return m

    def pagerank(self, weight=None, pers=None, damping: float = 0.85, epsilon: float = 1e-6,
                 max_iter: int = 0, rank=None) -> int:
        """``pagerank`` / ``pagerank_with`` (centrality.rs:234-384) into `rank`
        (warm start from its existing slots); returns the sweep count."""
        base = self._graph()
        if rank is None:
            rank = base.vp.setdefault("pagerank", DenseProp.new(base.graph_id))
        self._owned(rank, "pagerank: rank")
        self._sized(weight, base.edge_bound(), "pagerank: weight")
        self._sized(pers, base.vertex_bound(), "pagerank: personalisation")
        params = PowerIteration(epsilon, max_iter, Plan(base.vertex_bound(), 64))
        w = Unity.NEW if weight is None else weight
        if pers is None:
            return pagerank(self, w, damping, params, rank)
        return pagerank_with(self, w, pers, damping, params, rank)

    def eigenvector(self, weight=None, epsilon: float = 1e-6, max_iter: int = 0,
                    x=None) -> Tuple[float, int]:
        """``eigenvector`` (centrality.rs:405-497) into `x`; returns
        ``(eigenvalue, sweeps)``."""
        base = self._graph()
        if x is None:
            x = base.vp.setdefault("eigenvector", DenseProp.new(base.graph_id))
        self._owned(x, "eigenvector: centrality")
        self._sized(weight, base.edge_bound(), "eigenvector: weight")
        params = PowerIteration(epsilon, max_iter, Plan(base.vertex_bound(), 64))
        return eigenvector(self, Unity.NEW if weight is None else weight, params, x)


class AdjList(_GraphMethods):
    """An adjacency built by ``add_edge`` calls; edge ``i`` is the ``i``-th added.

    The device copy is built on first use and rebuilt after any mutation.
    """

    def __init__(self, n: int = 0):
        self.graph_id = next(_graph_ids)
        self._n = int(n)
        self._src = np.zeros(16, np.uint32)
        self._tgt = np.zeros(16, np.uint32)
        self._m = 0
        self._dev: Optional[DeviceGraph] = None
        self._dev_epoch = -1
        self._epoch = 0
        self._host = None
        self.vp = {}  # the graph's own vertex maps (name -> DenseProp)
        self.ep = {}  # and edge maps

    def _graph(self) -> "AdjList":
        return self

    def _new_map(self, value_type: str) -> "DenseProp":
        if value_type not in ("double", "float64"):
            raise GtValueError(_lib.GT_ERR_INVALID_ARG,
                               f"Invalid property type: {value_type}")
        return DenseProp.new(self.graph_id)

    @classmethod
    def with_vertices(cls, n: int) -> "AdjList":
        return cls(n)

    def add_vertex(self) -> int:
        self._n += 1
        self._epoch += 1
        return self._n - 1

    def _grow(self, need: int) -> None:
        cap = self._src.shape[0]
        if need <= cap:
            return
        cap = max(need, 2 * cap)
        for name in ("_src", "_tgt"):
            a = np.zeros(cap, np.uint32)
            a[: self._m] = getattr(self, name)[: self._m]
            setattr(self, name, a)

    def add_edge(self, s: int, t: int) -> int:
        if not 0 <= s < self._n:
            raise GtValueError(_lib.GT_ERR_NO_SUCH_VERTEX, f"no such vertex: {s}")
        if not 0 <= t < self._n:
            raise GtValueError(_lib.GT_ERR_NO_SUCH_VERTEX, f"no such vertex: {t}")
        self._grow(self._m + 1)
        self._src[self._m] = s
        self._tgt[self._m] = t
        self._m += 1
        self._epoch += 1
        return self._m - 1

    def add_edges(self, src, tgt) -> None:
        """``add_edge`` for every pair, in order (ParBuilder's result)."""
        src = np.asarray(src, dtype=np.int64)
        tgt = np.asarray(tgt, dtype=np.int64)
        if src.shape != tgt.shape:
            raise ValueError("source and target arrays differ in length")
        if src.size and (src.min() < 0 or tgt.min() < 0 or src.max() >= self._n
                         or tgt.max() >= self._n):
            raise GtValueError(_lib.GT_ERR_NO_SUCH_VERTEX, "no such vertex")
        k = int(src.shape[0])
        self._grow(self._m + k)
        self._src[self._m:self._m + k] = src
        self._tgt[self._m:self._m + k] = tgt
        self._m += k
        self._epoch += 1

    def num_vertices(self) -> int:
        return self._n

    def num_edges(self) -> int:
        return self._m

    def vertex_bound(self) -> int:
        return self._n

    def edge_bound(self) -> int:
        return self._m

    def edges(self) -> Tuple[np.ndarray, np.ndarray]:
        return self._src[: self._m].copy(), self._tgt[: self._m].copy()

    def _adjacency(self):
        if self._host is None or self._host[0] != self._epoch:
            self._host = (self._epoch, host_adjacency(self._n, self._src[: self._m],
                                                      self._tgt[: self._m]))
        return self._host[1]

    def in_edges(self, v: int):
        """``(other, edge id)`` pairs in storage order."""
        in_off, in_o, in_e, _, _, _ = self._adjacency()
        lo, hi = int(in_off[v]), int(in_off[v + 1])
        return list(zip(in_o[lo:hi].tolist(), in_e[lo:hi].tolist()))

    def out_edges(self, v: int):
        _, _, _, out_off, out_o, out_e = self._adjacency()
        lo, hi = int(out_off[v]), int(out_off[v + 1])
        return list(zip(out_o[lo:hi].tolist(), out_e[lo:hi].tolist()))

    def device(self) -> DeviceGraph:
        if self._dev is None or self._dev_epoch != self._epoch:
            if self._dev is not None:
                self._dev.close()
            self._dev = DeviceGraph.upload(self.graph_id, self._n, self._m, *self._adjacency())
            self._dev_epoch = self._epoch
        return self._dev

    # views -------------------------------------------------------------
    def reversed(self) -> "View":
        return View(self, _lib.GT_VIEW_REVERSED)

    def undirect(self) -> "View":
        return View(self, _lib.GT_VIEW_UNDIRECTED)

    def masked(self, vmask, emask, reversed: bool = False, undirected: bool = False) -> "View":
        kind = (_lib.GT_VIEW_UNDIRECTED_FILTERED if undirected else
                _lib.GT_VIEW_REVERSED_FILTERED if reversed else _lib.GT_VIEW_DIRECTED_FILTERED)
        vmask = np.ascontiguousarray(vmask, dtype=np.uint8)
        emask = np.ascontiguousarray(emask, dtype=np.uint8)
        if vmask.shape[0] < self._n or emask.shape[0] < self._m:
            raise GtValueError(_lib.GT_ERR_INVALID_ARG, "filter: mask shorter than the bound")
        return View(self, kind, vmask, emask)


class View(_GraphMethods):
    """One of the six view kinds over an :class:`AdjList`."""

    def __init__(self, base: AdjList, kind: int, vmask=None, emask=None):
        self.base, self.kind, self.vmask, self.emask = base, kind, vmask, emask

    def _graph(self) -> AdjList:
        return self.base

    @property
    def graph_id(self) -> int:
        return self.base.graph_id

    def vertex_bound(self) -> int:
        return self.base.vertex_bound()

    def num_vertices(self) -> int:
        if self.vmask is None:
            return self.base.num_vertices()
        return int(np.count_nonzero(self.vmask[: self.base.vertex_bound()]))

    def _c(self) -> GtView:
        v = GtView(self.kind, 0, None, 0, None, 0)
        if self.vmask is not None:
            v.vmask = self.vmask.ctypes.data_as(C.c_void_p)
            v.vmask_len = self.vmask.shape[0]
            v.emask = self.emask.ctypes.data_as(C.c_void_p)
            v.emask_len = self.emask.shape[0]
        return v


def _resolve(g):
    """(device graph, gt_view, graph_id, vertex_bound, num_vertices)."""
    if isinstance(g, View):
        return g.base.device(), g._c(), g.graph_id, g.vertex_bound(), g.num_vertices()
    if isinstance(g, AdjList):
        return (g.device(), GtView(_lib.GT_VIEW_DIRECTED, 0, None, 0, None, 0), g.graph_id,
                g.vertex_bound(), g.num_vertices())
    if isinstance(g, DeviceGraph):
        n, _, _, gid = g.info()
        return g, GtView(_lib.GT_VIEW_DIRECTED, 0, None, 0, None, 0), gid, n, n
    raise TypeError("expected an AdjList, a View or a DeviceGraph")


def _vertex_out(m: DenseProp, graph_id: int, bound: int) -> GtVertexOut:
    if m.graph_id != graph_id:
        raise GtValueError(_lib.GT_ERR_WRONG_GRAPH, "map does not belong to this graph")
    had = m.sized_for(graph_id, bound)
    d = m.data
    if _is_torch_cuda(d):
        import torch
        if d.dtype != torch.float64 or not d.is_contiguous():
            raise GtValueError(_lib.GT_ERR_INVALID_ARG, "rank map must be contiguous float64")
        return GtVertexOut(C.c_void_p(d.data_ptr()), int(d.numel()), had, GT_DEVICE, 0, graph_id)
    if d.dtype != np.float64 or not d.flags.c_contiguous:
        m.data = d = np.ascontiguousarray(d, dtype=np.float64)
    return GtVertexOut(d.ctypes.data_as(C.c_void_p), int(d.shape[0]), had, GT_HOST, 0, graph_id)


# -------------------------------------------------------------- centrality

def _check_map_owner(m: DenseProp, gid: int, what: str) -> None:
    if isinstance(m, DenseProp) and m.graph_id != gid:
        raise GtValueError(_lib.GT_ERR_WRONG_GRAPH,
                           f"{what}: rank map does not belong to this graph")


def pagerank(g, weight, damping: float, params: PowerIteration, rank: DenseProp) -> int:
    """PageRank with the uniform personalisation ``1 / num_vertices``
    (centrality.rs:234-248).  Returns the number of sweeps."""
    return _pagerank("gt_pagerank_f64", g, weight, None, damping, params, rank)


def pagerank_with(g, weight, pers, damping: float, params: PowerIteration,
                  rank: DenseProp) -> int:
    """PageRank with a personalisation map, not normalised (centrality.rs:261-384)."""
    if pers is None:
        raise ValueError("pagerank_with needs a personalisation map")
    return _pagerank("gt_pagerank_f64", g, weight, pers, damping, params, rank)


def pagerank_mixed(g, weight, pers, damping: float, params: PowerIteration,
                   rank: DenseProp) -> int:
    """FP32-value / FP64-residual PageRank (extension); ``pers=None`` is uniform."""
    return _pagerank("gt_pagerank_mixed", g, weight, pers, damping, params, rank)


def _bind_stream(dev: DeviceGraph, *maps) -> None:
    """Run on torch's current stream when any map is a CUDA tensor, so the
    call is ordered after the work that produced it."""
    for m in maps:
        d = m.data if isinstance(m, DenseProp) else m
        if _is_torch_cuda(d):
            import torch
            dev.set_stream(torch.cuda.current_stream(d.device).cuda_stream)
            return
    dev.set_stream(None)


def _pagerank(fn, g, weight, pers, damping, params, rank) -> int:
    dev, view, gid, bound, nv = _resolve(g)
    _bind_stream(dev, rank, weight, pers)
    if rank.graph_id != gid:
        raise GtValueError(_lib.GT_ERR_WRONG_GRAPH,
                           "pagerank: rank map does not belong to this graph")
    keep = []
    wp = _prop(weight if weight is not None else Unity.NEW, gid, keep)
    if pers is None and nv != bound:
        pers = Constant(0.0 if nv == 0 else 1.0 / nv)
    pp = _prop(pers, gid, keep)
    out = _vertex_out(rank, gid, bound)
    iters = C.c_uint64()
    call(fn, dev.handle, C.byref(view), C.byref(wp), C.byref(pp) if pp else None,
         float(damping), C.byref(params._c()), C.byref(out), C.byref(iters))
    return int(iters.value)


def pagerank_batched(g, weight, seeds, damping: float, params: PowerIteration, out=None):
    """k personalised PageRank vectors in one SpMM per sweep (extension).

    Column c equals ``pagerank_with`` with the one-hot personalisation on
    ``seeds[c]``, started from it and stopped on its own L1 change.  Returns
    ``(X, iters)``: X is [n, k] float32 (a CUDA tensor if ``out`` is one, else
    numpy), iters the per-column sweep counts."""
    dev, view, gid, bound, nv = _resolve(g)
    seeds = np.ascontiguousarray(seeds, dtype=np.uint32)
    k = int(seeds.shape[0])
    keep = []
    wp = _prop(weight if weight is not None else Unity.NEW, gid, keep)
    iters = np.zeros(max(k, 1), np.uint64)
    if out is not None and _is_torch_cuda(out):
        import torch
        if out.dtype != torch.float32 or out.dim() != 2 or out.shape[0] != bound \
                or out.shape[1] < k or out.stride(1) != 1:
            raise GtValueError(_lib.GT_ERR_INVALID_ARG, "out must be [n, >=k] float32, row-major")
        _bind_stream(dev, out)
        ptr, ldx, loc, X = C.c_void_p(out.data_ptr()), out.stride(0), GT_DEVICE, out
    else:
        X = np.zeros((bound, max(k, 1)), np.float32)
        dev.set_stream(None)
        ptr, ldx, loc = X.ctypes.data_as(C.c_void_p), X.shape[1], GT_HOST
    call("gt_pagerank_batched", dev.handle, C.byref(view), C.byref(wp), _ptr(seeds), k,
         float(damping), C.byref(params._c()), ptr, ldx, loc, _ptr(iters, C.POINTER(C.c_uint64)))
    return X, iters[:k].astype(np.int64)


def pagerank_grad(g, weight, pers, damping: float, params: PowerIteration, x, dl_dx,
                  precision: str = "f64", grad_w_dtype=None):
    """Gradient of a loss through the PageRank fixed point (extension).

    ``x`` is the forward solution, ``dl_dx`` = dL/dx.  Returns ``(grad_w,
    grad_pers, grad_damping, sweeps)``: grad_w indexed by edge id, grad_pers by
    vertex.  CUDA tensors in, CUDA tensors out; otherwise numpy."""
    dev, view, gid, bound, nv = _resolve(g)
    _, _, eb, _ = dev.info()
    keep = []
    wp = _prop(weight if weight is not None else Unity.NEW, gid, keep)
    pp = _prop(pers, gid, keep)
    xp = _prop(x, gid, keep)
    cp = _prop(dl_dx, gid, keep)
    prec = {"f64": GT_F64, "f32": GT_F32}[precision]
    on_dev = _is_torch_cuda(x.data if isinstance(x, DenseProp) else x)
    if on_dev:
        import torch
        _bind_stream(dev, x)
        tdt = torch.float64 if (grad_w_dtype or "f64") == "f64" else torch.float32
        gw = torch.empty(max(eb, 1), dtype=tdt, device="cuda")
        gpv = torch.empty(max(bound, 1), dtype=torch.float64, device="cuda")
        gw_ptr, gp_ptr, loc = C.c_void_p(gw.data_ptr()), C.c_void_p(gpv.data_ptr()), GT_DEVICE
    else:
        dev.set_stream(None)
        ndt = np.float64 if (grad_w_dtype or "f64") == "f64" else np.float32
        gw = np.zeros(max(eb, 1), ndt)
        gpv = np.zeros(max(bound, 1), np.float64)
        gw_ptr, gp_ptr, loc = gw.ctypes.data_as(C.c_void_p), gpv.ctypes.data_as(C.c_void_p), GT_HOST
    wdt = GT_F64 if (grad_w_dtype or "f64") == "f64" else GT_F32
    gd = C.c_double()
    it = C.c_uint64()
    call("gt_pagerank_backward", dev.handle, C.byref(view), C.byref(wp),
         C.byref(pp) if pp else None, float(damping), C.byref(params._c()), C.byref(xp),
         C.byref(cp), prec, gw_ptr, wdt, loc, gp_ptr, C.byref(gd), C.byref(it))
    return gw[:eb], gpv[:bound], float(gd.value), int(it.value)


def eigenvector(g, weight, params: PowerIteration, x: DenseProp) -> Tuple[float, int]:
    """Eigenvector centrality (centrality.rs:405-497): ``(eigenvalue, sweeps)``."""
    dev, view, gid, bound, nv = _resolve(g)
    _bind_stream(dev, x, weight)
    if x.graph_id != gid:
        raise GtValueError(_lib.GT_ERR_WRONG_GRAPH,
                           "eigenvector: centrality map does not belong to this graph")
    keep = []
    wp = _prop(weight if weight is not None else Unity.NEW, gid, keep)
    out = _vertex_out(x, gid, bound)
    eig = C.c_double()
    iters = C.c_uint64()
    call("gt_eigenvector_f64", dev.handle, C.byref(view), C.byref(wp), C.byref(params._c()),
         C.byref(out), C.byref(eig), C.byref(iters))
    return float(eig.value), int(iters.value)


def _bc_map(m: DenseProp, gid: int, bound: int, what: str):
    if m.graph_id != gid:
        raise GtValueError(_lib.GT_ERR_WRONG_GRAPH,
                           f"betweenness: {what} map does not belong to this graph")
    m.sized_for(gid, bound)
    d = m.data
    if _is_torch_cuda(d):
        import torch
        if d.dtype != torch.float64 or not d.is_contiguous():
            raise GtValueError(_lib.GT_ERR_INVALID_ARG, f"{what} map must be contiguous float64")
        return C.c_void_p(d.data_ptr()), int(d.numel()), GT_DEVICE
    if d.dtype != np.float64 or not d.flags.c_contiguous:
        m.data = d = np.ascontiguousarray(d, dtype=np.float64)
    return d.ctypes.data_as(C.c_void_p), int(d.shape[0]), GT_HOST


def betweenness(g, vertex_bc: DenseProp, edge_bc: Optional[DenseProp], pivots=None) -> dict:
    """Brandes betweenness (crates/gt-algo/src/centrality.rs:614-689) over any
    view: both maps are grown to their bounds and overwritten with the
    unnormalised vertex and edge betweenness (halved on an undirected view).
    ``edge_bc=None`` skips the edge map and ``pivots`` restricts the sources
    (both extensions).  The maps are numpy arrays or CUDA tensors, both in the
    same memory.  Returns the call's work counters."""
    dev, view, gid, bound, nv = _resolve(g)
    _, _, eb, _ = dev.info()
    vptr, vlen, vloc = _bc_map(vertex_bc, gid, bound, "vertex")
    out = _lib.GtBcOut(vptr, vlen, None, 0, vloc, 0, gid, 0)
    if edge_bc is not None:
        eptr, elen, eloc = _bc_map(edge_bc, gid, eb, "edge")
        if eloc != vloc:
            raise GtValueError(_lib.GT_ERR_INVALID_ARG, "both maps must be in the same memory")
        out.edge, out.edge_len, out.edge_graph_id = eptr, elen, gid
    _bind_stream(dev, vertex_bc)
    pv = None if pivots is None else np.ascontiguousarray(pivots, dtype=np.uint32)
    st = _lib.GtBcStats()
    call("gt_betweenness", dev.handle, C.byref(view), None if pv is None else _ptr(pv),
         0 if pv is None else int(pv.shape[0]), C.byref(out), C.byref(st))
    return {k: getattr(st, k) for k, _ in _lib.GtBcStats._fields_}


# ------------------------------------------------------------- generators

def gen_er(n: int, m: int, seed: int) -> Tuple[np.ndarray, np.ndarray]:
    src = np.zeros(max(m, 1), np.uint32)
    tgt = np.zeros(max(m, 1), np.uint32)
    call("gt_gen_er", n, m, seed & (2**64 - 1), _ptr(src), _ptr(tgt))
    return src[:m], tgt[:m]


def gen_rmat(scale: int, m: int, a: float, b: float, c: float,
             seed: int) -> Tuple[np.ndarray, np.ndarray]:
    src = np.zeros(max(m, 1), np.uint32)
    tgt = np.zeros(max(m, 1), np.uint32)
    call("gt_gen_rmat", scale, m, a, b, c, seed & (2**64 - 1), _ptr(src), _ptr(tgt))
    return src[:m], tgt[:m]


def gen_chung_lu_batch(num_graphs: int, seed: int, n_min: int = 64, n_max: int = 65536,
                       mean_degree: float = 8.0, exponent: float = 2.1, d_lo: float = 0.5,
                       d_hi: float = 0.95):
    """A batch of directed Chung-Lu graphs with global vertex and edge ids
    (see gt_gen_chung_lu_batch).  Returns (vertex_offsets [G+1] uint32,
    edge_offsets [G+1] int64, damping [G], src, tgt, weight float32)."""
    G = int(num_graphs)
    ng = np.zeros(max(G, 1), np.uint32)
    mg = np.zeros(max(G, 1), np.uint64)
    dmp = np.zeros(max(G, 1), np.float64)
    args = (G, seed & (2**64 - 1), n_min, n_max, float(mean_degree), float(exponent),
            float(d_lo), float(d_hi), _ptr(ng), _ptr(mg, C.POINTER(C.c_uint64)),
            _ptr(dmp, C.POINTER(C.c_double)))
    call("gt_gen_chung_lu_batch", *args, None, None, None)
    m = int(mg[:G].sum())
    src = np.zeros(max(m, 1), np.uint32)
    tgt = np.zeros(max(m, 1), np.uint32)
    w = np.zeros(max(m, 1), np.float32)
    call("gt_gen_chung_lu_batch", *args, _ptr(src), _ptr(tgt), _ptr(w, C.POINTER(C.c_float)))
    voff = np.zeros(G + 1, np.uint32)
    voff[1:] = np.cumsum(ng[:G], dtype=np.uint64).astype(np.uint32)
    eoff = np.zeros(G + 1, np.int64)
    eoff[1:] = np.cumsum(mg[:G].astype(np.int64))
    return voff, eoff, dmp[:G].copy(), src[:m], tgt[:m], w[:m]


def _uneven_args(dev: "DeviceGraph", vertex_offsets, damping):
    voff = np.ascontiguousarray(vertex_offsets, dtype=np.uint32)
    dmp = np.ascontiguousarray(damping, dtype=np.float64)
    G = int(voff.shape[0]) - 1
    if G < 1 or dmp.shape[0] != G:
        raise GtValueError(_lib.GT_ERR_INVALID_ARG,
                           "vertex_offsets needs G+1 entries and damping G")
    n = dev.info()[0]
    return voff, dmp, G, n


def pagerank_uneven(dev: "DeviceGraph", vertex_offsets, damping, weight=None,
                    epsilon: float = 1e-6, max_iter: int = 0, out=None):
    """PageRank of every graph of a block-diagonal batch (extension; see
    gt_pagerank_uneven).  Returns (rank [n] float64, iters [G]); rank is the
    CUDA tensor ``out`` when one is given, else numpy."""
    voff, dmp, G, n = _uneven_args(dev, vertex_offsets, damping)
    keep = []
    wp = _prop(weight if weight is not None else Unity.NEW, dev.graph_id, keep)
    params = PowerIteration(epsilon, max_iter, Plan.from_parts(n, max(1, -(-n // 4096))))
    iters = np.zeros(G, np.uint64)
    if out is not None and _is_torch_cuda(out):
        _bind_stream(dev, out)
        ptr, loc, res = C.c_void_p(out.data_ptr()), GT_DEVICE, out
    else:
        res = np.zeros(max(n, 1), np.float64)
        dev.set_stream(None)
        ptr, loc = res.ctypes.data_as(C.c_void_p), GT_HOST
    call("gt_pagerank_uneven", dev.handle, G, _ptr(voff), _ptr(dmp, C.POINTER(C.c_double)),
         C.byref(wp), C.byref(params._c()), ptr, loc, _ptr(iters, C.POINTER(C.c_uint64)))
    return (res if loc == GT_DEVICE else res[:n]), iters.astype(np.int64)


def pagerank_uneven_grad(dev: "DeviceGraph", vertex_offsets, damping, x, dl_dx, weight=None,
                         epsilon: float = 1e-6, max_iter: int = 0):
    """Gradients of L = dl_dx . x through every graph's fixed point
    (extension; see gt_pagerank_uneven_backward).  Host numpy in and out:
    returns (grad_w [edge_bound] float32, grad_pers [n], grad_damping [G],
    iters [G])."""
    voff, dmp, G, n = _uneven_args(dev, vertex_offsets, damping)
    _, _, eb, _ = dev.info()
    keep = []
    wp = _prop(weight if weight is not None else Unity.NEW, dev.graph_id, keep)
    params = PowerIteration(epsilon, max_iter, Plan.from_parts(n, max(1, -(-n // 4096))))
    xa = np.ascontiguousarray(x, dtype=np.float64)
    ca = np.ascontiguousarray(dl_dx, dtype=np.float64)
    gw = np.zeros(max(eb, 1), np.float32)
    gp = np.zeros(max(n, 1), np.float64)
    gd = np.zeros(G, np.float64)
    iters = np.zeros(G, np.uint64)
    dev.set_stream(None)
    call("gt_pagerank_uneven_backward", dev.handle, G, _ptr(voff),
         _ptr(dmp, C.POINTER(C.c_double)), C.byref(wp), C.byref(params._c()),
         xa.ctypes.data_as(C.c_void_p), ca.ctypes.data_as(C.c_void_p), GT_HOST,
         gw.ctypes.data_as(C.c_void_p), gp.ctypes.data_as(C.c_void_p),
         _ptr(gd, C.POINTER(C.c_double)), _ptr(iters, C.POINTER(C.c_uint64)))
    return gw[:eb], gp[:n], gd, iters.astype(np.int64)


# ------------------------------------------------------------------ formats

# Value-type indices of the .gt format (graph-tool's value_types order).
VALUE_KINDS = ("bool", "int16_t", "int32_t", "int64_t", "double", "long double", "string",
               "vector<bool>", "vector<int16_t>", "vector<int32_t>", "vector<int64_t>",
               "vector<double>", "vector<long double>", "vector<string>", "python::object")
DOMAINS = ("graph", "vertex", "edge")


class Document:
    """A parsed graph file (gt_read_gt / gt_read_edge_list): the edge list in
    edge-id order, the property columns and the vertex names.  ``adjlist()``
    replays it into an AdjList, ``device()`` builds the device graph."""

    def __init__(self, handle: int):
        self._h = C.c_void_p(handle)
        n, m, nn = C.c_uint64(), C.c_uint64(), C.c_uint64()
        d, npr = C.c_int32(), C.c_uint32()
        call("gt_doc_info", self._h, C.byref(n), C.byref(m), C.byref(d), C.byref(npr),
             C.byref(nn))
        self.n, self.m, self.directed = int(n.value), int(m.value), bool(d.value)
        self._nprops, self._nnames = int(npr.value), int(nn.value)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            _lib.lib().gt_doc_free(h)
            self._h = None

    @property
    def comment(self) -> str:
        p, ln = C.c_void_p(), C.c_uint64()
        call("gt_doc_comment", self._h, C.byref(p), C.byref(ln))
        return C.string_at(p.value, ln.value).decode("utf-8", errors="replace") if ln.value else ""

    def edges(self) -> Tuple[np.ndarray, np.ndarray]:
        src = np.zeros(max(self.m, 1), np.uint32)
        tgt = np.zeros(max(self.m, 1), np.uint32)
        call("gt_doc_edges", self._h, _ptr(src), _ptr(tgt))
        return src[:self.m], tgt[:self.m]

    def names(self) -> list:
        out = []
        for v in range(self._nnames):
            p, ln = C.c_void_p(), C.c_uint64()
            call("gt_doc_vertex_name", self._h, v, C.byref(p), C.byref(ln))
            out.append(C.string_at(p.value, ln.value).decode("utf-8") if ln.value else "")
        return out

    def _cell(self, k: int, i: int) -> bytes:
        p, ln = C.c_void_p(), C.c_uint64()
        call("gt_doc_prop_cell", self._h, k, i, C.byref(p), C.byref(ln))
        return C.string_at(p.value, ln.value) if ln.value else b""

    @staticmethod
    def _decode(kind: int, b: bytes):
        name = VALUE_KINDS[kind]
        if name == "string":
            return b.decode("utf-8")
        if name in ("python::object", "long double"):
            return b  # opaque: never unpickled or converted
        if name == "vector<long double>":
            return [b[i:i + 16] for i in range(0, len(b), 16)]
        if name == "vector<string>":
            k = int.from_bytes(b[:8], "little")
            out, pos = [], 8
            for _ in range(k):
                ln = int.from_bytes(b[pos:pos + 8], "little")
                out.append(b[pos + 8:pos + 8 + ln].decode("utf-8"))
                pos += 8 + ln
            return out
        return np.frombuffer(b, np.float64).copy()  # numeric vectors, widened to f64

    def properties(self) -> list:
        """[{name, domain, kind, values}] in file order; numeric scalar columns
        as float64 arrays, the others as lists (see _decode)."""
        out = []
        for k in range(self._nprops):
            p, ln = C.c_void_p(), C.c_uint64()
            dom, kind, cnt = C.c_int32(), C.c_int32(), C.c_uint64()
            call("gt_doc_prop_info", self._h, k, C.byref(p), C.byref(ln), C.byref(dom),
                 C.byref(kind), C.byref(cnt))
            name = C.string_at(p.value, ln.value).decode("utf-8") if ln.value else ""
            kd = int(kind.value)
            if kd in (0, 1, 2, 3, 4):
                vals = np.zeros(max(int(cnt.value), 1), np.float64)
                call("gt_doc_prop_f64", self._h, k, _ptr(vals, C.POINTER(C.c_double)))
                vals = vals[:int(cnt.value)]
            else:
                vals = [self._decode(kd, self._cell(k, i)) for i in range(int(cnt.value))]
            out.append({"name": name, "domain": DOMAINS[int(dom.value)],
                        "kind": VALUE_KINDS[kd], "values": vals})
        return out

    def adjlist(self) -> "AdjList":
        g = AdjList.with_vertices(self.n)
        s, t = self.edges()
        g.add_edges(s, t)
        return g

    def device(self, graph_id: Optional[int] = None) -> DeviceGraph:
        gid = next(_graph_ids) if graph_id is None else graph_id
        h = _lib.GraphP()
        call("gt_graph_from_doc", self._h, gid, C.byref(h))
        return DeviceGraph(h.value, gid)


def _bytes_of(src) -> bytes:
    if isinstance(src, (bytes, bytearray, memoryview)):
        return bytes(src)
    with open(src, "rb") as f:
        return f.read()


def read_gt(src) -> Document:
    """Parse a .gt file (a path or its bytes)."""
    b = _bytes_of(src)
    h = C.c_void_p()
    call("gt_read_gt", C.c_char_p(b), len(b), C.byref(h))
    return Document(h.value)


def read_edge_list(src, delimiter: str = ",", header: bool = False,
                   symbolic: bool = True) -> Document:
    """Parse a delimited edge list (a path or its bytes)."""
    b = _bytes_of(src)
    d = delimiter.encode("latin-1")
    if len(d) != 1:
        raise GtValueError(_lib.GT_ERR_INVALID_ARG, "the delimiter must be one byte")
    h = C.c_void_p()
    call("gt_read_edge_list", C.c_char_p(b), len(b), d[0], int(header), int(symbolic),
         C.byref(h))
    return Document(h.value)


# ------------------------------------------------------------------ degrees

def _degrees(g, which: int, weight) -> np.ndarray:
    dev, view, gid, bound, nv = _resolve(g)
    keep = []
    wp = _prop(weight if weight is not None else Unity.NEW, gid, keep)
    out = np.zeros(max(bound, 1), np.float64)
    dev.set_stream(None)
    call("gt_degrees", dev.handle, C.byref(view), which, C.byref(wp),
         out.ctypes.data_as(C.c_void_p), GT_HOST)
    return out[:bound]


def weighted_out_degrees(g, weight=None) -> np.ndarray:
    """``weighted_out_degree`` of every vertex (crates/gt-algo/src/degree.rs:122-140)."""
    return _degrees(g, 0, weight)


def weighted_in_degrees(g, weight=None) -> np.ndarray:
    """``weighted_in_degree`` of every vertex (degree.rs:146-162)."""
    return _degrees(g, 1, weight)


def weighted_degrees(g, weight=None) -> np.ndarray:
    """``weighted_degree`` of every vertex (degree.rs:172-185)."""
    return _degrees(g, 2, weight)


def weighted_out_degree(g, v: int, weight=None) -> float:
    return float(weighted_out_degrees(g, weight)[v])


def weighted_in_degree(g, v: int, weight=None) -> float:
    return float(weighted_in_degrees(g, weight)[v])


def weighted_degree(g, v: int, weight=None) -> float:
    return float(weighted_degrees(g, weight)[v])


def out_degrees(g, out: Optional[DenseProp] = None) -> DenseProp:
    """``out_degrees`` (degree.rs:199-216): an int64 map grown to the vertex
    bound, every slot rewritten."""
    dev, view, gid, bound, nv = _resolve(g)
    if out is None:
        out = DenseProp(gid, np.zeros(0, np.int64), dtype=np.int64)
    out.sized_for(gid, bound)
    buf = np.zeros(max(bound, 1), np.int64)
    dev.set_stream(None)
    call("gt_out_degrees", dev.handle, C.byref(view), buf.ctypes.data_as(C.c_void_p), GT_HOST)
    out.data = np.ascontiguousarray(out.data).astype(np.int64)
    out.data[:bound] = buf[:bound]
    return out


def degree_histogram(g) -> list:
    """``degree_histogram`` (degree.rs:230-247): hist[d] = number of the view's
    vertices of total degree d."""
    dev, view, gid, bound, nv = _resolve(g)
    ln = C.c_uint64()
    dev.set_stream(None)
    call("gt_degree_histogram", dev.handle, C.byref(view), None, 0, C.byref(ln))
    if ln.value == 0:
        return []
    h = np.zeros(ln.value, np.uint64)
    call("gt_degree_histogram", dev.handle, C.byref(view), _ptr(h, C.POINTER(C.c_uint64)),
         ln.value, C.byref(ln))
    return [int(x) for x in h[:ln.value]]

