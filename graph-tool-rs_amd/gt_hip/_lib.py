"""ctypes binding of libgt_hip.so (the C ABI declared in include/gt_hip.h).

Loading fails loudly: there is no CPU fallback behind this module.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

_LIB_PATH = Path(os.environ.get("GT_HIP_LIB", Path(__file__).resolve().parent / "libgt_hip.so"))

u32p = C.POINTER(C.c_uint32)
u8p = C.POINTER(C.c_uint8)
f64p = C.POINTER(C.c_double)

GT_OK = 0
GT_ERR_INVALID_ARG = 1
GT_ERR_PLAN_COVER = 2
GT_ERR_WRONG_GRAPH = 3
GT_ERR_NOT_DIRECTED = 4
GT_ERR_NO_SUCH_VERTEX = 5
GT_ERR_HIP = 6
GT_ERR_OOM = 7
GT_ERR_UNSUPPORTED = 8

GT_HOST, GT_DEVICE = 0, 1
GT_F64, GT_F32 = 0, 1
GT_MAP_UNITY, GT_MAP_CONSTANT, GT_MAP_DENSE = 0, 1, 2
(GT_VIEW_DIRECTED, GT_VIEW_REVERSED, GT_VIEW_DIRECTED_FILTERED, GT_VIEW_REVERSED_FILTERED,
 GT_VIEW_UNDIRECTED, GT_VIEW_UNDIRECTED_FILTERED) = range(6)


class GtProp(C.Structure):
    _fields_ = [("kind", C.c_int32), ("dtype", C.c_int32), ("loc", C.c_int32),
                ("reserved", C.c_int32), ("value", C.c_double), ("data", C.c_void_p),
                ("len", C.c_uint64), ("graph_id", C.c_uint64)]


class GtPlan(C.Structure):
    _fields_ = [("len", C.c_uint64), ("chunks", C.c_uint64)]


class GtPowerIteration(C.Structure):
    _fields_ = [("epsilon", C.c_double), ("max_iter", C.c_uint64), ("plan", GtPlan)]


class GtVertexOut(C.Structure):
    _fields_ = [("data", C.c_void_p), ("len", C.c_uint64), ("had", C.c_uint64),
                ("loc", C.c_int32), ("reserved", C.c_int32), ("graph_id", C.c_uint64)]


class GtTuning(C.Structure):
    _fields_ = [("sweep_blocks_per_cu", C.c_int32), ("tile_rows", C.c_int32),
                ("short_row", C.c_int32), ("mm_tile_nnz", C.c_int32),
                ("mm_blocks_per_cu", C.c_int32), ("sweep_hot_kb", C.c_int32),
                ("sliced_min_len", C.c_int32), ("sliced_share", C.c_int32),
                ("block_tile_nnz", C.c_int32), ("block_long_min", C.c_int32),
                ("block_per_cu", C.c_int32), ("block_f64", C.c_int32)]


class GtView(C.Structure):
    _fields_ = [("kind", C.c_int32), ("reserved", C.c_int32), ("vmask", C.c_void_p),
                ("vmask_len", C.c_uint64), ("emask", C.c_void_p), ("emask_len", C.c_uint64)]


(GT_SWEEP_PAGERANK_F64, GT_SWEEP_PAGERANK_F32, GT_SWEEP_EIGENVECTOR, GT_SWEEP_OTHER,
 GT_SWEEP_BATCHED, GT_SWEEP_ADJOINT, GT_SWEEP_BC_FORWARD, GT_SWEEP_BC_BACKWARD) = range(8)


class GtBcOut(C.Structure):
    _fields_ = [("vertex", C.c_void_p), ("vertex_len", C.c_uint64), ("edge", C.c_void_p),
                ("edge_len", C.c_uint64), ("loc", C.c_int32), ("reserved", C.c_int32),
                ("vertex_graph_id", C.c_uint64), ("edge_graph_id", C.c_uint64)]


class GtBcStats(C.Structure):
    _fields_ = [("sources", C.c_uint64), ("batches", C.c_uint64),
                ("batches_rerun", C.c_uint64), ("levels", C.c_uint64),
                ("forward_entries", C.c_uint64), ("forward_hits", C.c_uint64),
                ("forward_hit_entries", C.c_uint64),
                ("backward_entries", C.c_uint64), ("backward_hits", C.c_uint64),
                ("backward_hit_entries", C.c_uint64),
                ("forward_ms", C.c_double), ("backward_ms", C.c_double),
                ("streams", C.c_uint64)]

class GtSweepLayout(C.Structure):
    _fields_ = [("slot_bits", C.c_uint32), ("tiles", C.c_uint32), ("split_rows", C.c_uint32),
                ("slices", C.c_uint32), ("sliced_rows", C.c_uint64),
                ("sliced_entries", C.c_uint64), ("block_tiles", C.c_uint32),
                ("block_long_rows", C.c_uint32), ("block_groups", C.c_uint64)]


GraphP = C.c_void_p

# name -> (argtypes); every function returns gt_status (c_int).
_PROTOS = {
    "gt_device_count": [C.POINTER(C.c_int32)],
    "gt_set_device_memory_limit": [C.c_uint64],
    "gt_device_memory_in_use": [C.POINTER(C.c_uint64)],
    "gt_graph_upload": [C.c_uint64, C.c_uint32, C.c_uint64, u32p, u32p, u32p, u32p, u32p, u32p,
                        C.POINTER(GraphP)],
    "gt_graph_build_edges": [C.c_uint64, C.c_uint32, C.c_uint64, C.c_void_p, C.c_void_p,
                             C.c_int32, C.POINTER(GraphP)],
    "gt_graph_free": [GraphP],
    "gt_graph_info": [GraphP, u32p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                      C.POINTER(C.c_uint64)],
    "gt_graph_download": [GraphP, u32p, u32p, u32p, u32p, u32p, u32p],
    "gt_graph_set_stream": [GraphP, C.c_void_p],
    "gt_pagerank_f64": [GraphP, C.POINTER(GtView), C.POINTER(GtProp), C.POINTER(GtProp),
                        C.c_double, C.POINTER(GtPowerIteration), C.POINTER(GtVertexOut),
                        C.POINTER(C.c_uint64)],
    "gt_pagerank_mixed": [GraphP, C.POINTER(GtView), C.POINTER(GtProp), C.POINTER(GtProp),
                          C.c_double, C.POINTER(GtPowerIteration), C.POINTER(GtVertexOut),
                          C.POINTER(C.c_uint64)],
    "gt_pagerank_f64_noalloc": [GraphP, C.POINTER(GtView), C.POINTER(GtProp), C.POINTER(GtProp),
                                C.c_double, C.POINTER(GtPowerIteration),
                                C.POINTER(GtVertexOut), C.POINTER(C.c_uint64)],
    "gt_pagerank_batched": [GraphP, C.POINTER(GtView), C.POINTER(GtProp), u32p, C.c_uint32,
                            C.c_double, C.POINTER(GtPowerIteration), C.c_void_p, C.c_uint64,
                            C.c_int32, C.POINTER(C.c_uint64)],
    "gt_pagerank_backward": [GraphP, C.POINTER(GtView), C.POINTER(GtProp), C.POINTER(GtProp),
                             C.c_double, C.POINTER(GtPowerIteration), C.POINTER(GtProp),
                             C.POINTER(GtProp), C.c_int32, C.c_void_p, C.c_int32, C.c_int32,
                             C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_uint64)],
    "gt_eigenvector_f64": [GraphP, C.POINTER(GtView), C.POINTER(GtProp),
                           C.POINTER(GtPowerIteration), C.POINTER(GtVertexOut),
                           C.POINTER(C.c_double), C.POINTER(C.c_uint64)],
    "gt_graph_set_profiling": [GraphP, C.c_int32],
    "gt_sweep_stats": [GraphP, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_uint64),
                       C.POINTER(C.c_uint64)],
    "gt_last_layout": [GraphP, C.POINTER(GtSweepLayout)],
    "gt_gen_er": [C.c_uint32, C.c_uint64, C.c_uint64, u32p, u32p],
    "gt_pagerank_uneven": [GraphP, C.c_uint32, u32p, C.POINTER(C.c_double), C.POINTER(GtProp),
                           C.POINTER(GtPowerIteration), C.c_void_p, C.c_int32,
                           C.POINTER(C.c_uint64)],
    "gt_pagerank_uneven_backward": [GraphP, C.c_uint32, u32p, C.POINTER(C.c_double),
                                    C.POINTER(GtProp), C.POINTER(GtPowerIteration), C.c_void_p,
                                    C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p,
                                    C.POINTER(C.c_double), C.POINTER(C.c_uint64)],
    "gt_betweenness": [GraphP, C.POINTER(GtView), u32p, C.c_uint64, C.POINTER(GtBcOut),
                       C.POINTER(GtBcStats)],
    "gt_degrees": [GraphP, C.POINTER(GtView), C.c_int32, C.POINTER(GtProp), C.c_void_p,
                   C.c_int32],
    "gt_out_degrees": [GraphP, C.POINTER(GtView), C.c_void_p, C.c_int32],
    "gt_degree_histogram": [GraphP, C.POINTER(GtView), C.POINTER(C.c_uint64), C.c_uint64,
                            C.POINTER(C.c_uint64)],
    "gt_read_gt": [C.c_void_p, C.c_uint64, C.POINTER(C.c_void_p)],
    "gt_read_edge_list": [C.c_void_p, C.c_uint64, C.c_uint8, C.c_int32, C.c_int32,
                          C.POINTER(C.c_void_p)],
    "gt_doc_free": [C.c_void_p],
    "gt_doc_info": [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                    C.POINTER(C.c_int32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint64)],
    "gt_doc_edges": [C.c_void_p, u32p, u32p],
    "gt_doc_comment": [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_uint64)],
    "gt_doc_prop_info": [C.c_void_p, C.c_uint32, C.POINTER(C.c_void_p), C.POINTER(C.c_uint64),
                         C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_uint64)],
    "gt_doc_prop_f64": [C.c_void_p, C.c_uint32, C.POINTER(C.c_double)],
    "gt_doc_prop_cell": [C.c_void_p, C.c_uint32, C.c_uint64, C.POINTER(C.c_void_p),
                         C.POINTER(C.c_uint64)],
    "gt_doc_vertex_name": [C.c_void_p, C.c_uint64, C.POINTER(C.c_void_p),
                           C.POINTER(C.c_uint64)],
    "gt_graph_from_doc": [C.c_void_p, C.c_uint64, C.POINTER(GraphP)],
    "gt_graph_set_tuning": [GraphP, C.POINTER(GtTuning)],
    "gt_graph_get_tuning": [GraphP, C.POINTER(GtTuning)],
    "gt_gen_chung_lu_batch": [C.c_uint32, C.c_uint64, C.c_uint32, C.c_uint32, C.c_double,
                              C.c_double, C.c_double, C.c_double, u32p,
                              C.POINTER(C.c_uint64), C.POINTER(C.c_double), u32p, u32p,
                              C.POINTER(C.c_float)],
    "gt_gen_rmat": [C.c_uint32, C.c_uint64, C.c_double, C.c_double, C.c_double, C.c_uint64,
                    u32p, u32p],
    "gt_host_adjacency": [C.c_uint32, C.c_uint64, u32p, u32p, u32p, u32p, u32p, u32p, u32p,
                          u32p],
}

EXPORTED = tuple(_PROTOS) + ("gt_last_error",)

_lib = None


def lib() -> C.CDLL:
    """The loaded library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            raise RuntimeError(
                f"{_LIB_PATH} is missing: build it with __graft_entry__.build() "
                "(the HIP engine has no CPU fallback)")
        # torch ships its own HIP runtime under the same soname; loading it
        # first makes this library bind to it, so one runtime serves both
        # (two runtimes in one process leave the second unable to see the GPU).
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(str(_LIB_PATH))
        for name, args in _PROTOS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_int
        L.gt_last_error.argtypes = []
        L.gt_last_error.restype = C.c_char_p
        _lib = L
    return _lib


class GtError(RuntimeError):
    """A device-side failure (HIP error, out of memory, unsupported request)."""

    def __init__(self, status: int, msg: str):
        super().__init__(msg)
        self.status = status


class GtValueError(ValueError):
    """A caller error (the reference raises ValueError / panics for these)."""

    def __init__(self, status: int, msg: str):
        super().__init__(msg)
        self.status = status


def check(status: int) -> None:
    if status == GT_OK:
        return
    msg = lib().gt_last_error().decode(errors="replace")
    if status in (GT_ERR_INVALID_ARG, GT_ERR_PLAN_COVER, GT_ERR_WRONG_GRAPH,
                  GT_ERR_NOT_DIRECTED, GT_ERR_NO_SUCH_VERTEX):
        raise GtValueError(status, msg)
    raise GtError(status, msg)


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args))
