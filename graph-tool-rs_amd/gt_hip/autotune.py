"""Autotuner for the launch and schedule parameters of one device graph.

The pull sweeps (PageRank, eigenvector, adjoint, uneven batch) depend on
three parameters: the resident workgroups per compute unit, the rows per wave
tile, and the row length below which one lane sums a row (above it the whole
wavefront does).  The FP64 PageRank sweep on a relabelled graph runs one
workgroup per compute unit and also depends on how many of the most gathered
contributions it keeps in LDS.  The FP32 PageRank sweep on a relabelled,
unweighted graph runs on block tiles, whose size and resident workgroups per
compute unit are tuned too; the PageRank probe therefore times an FP32 sweep
and an FP64 sweep, and minimises their sum.  The batched SpMM depends on its tile size and its resident
workgroups per compute unit.  ``autotune`` times candidates on the caller's
own graph -- the sweep kernels alone, measured with the HIP events the library
records on its stream (gt_graph_set_profiling / gt_sweep_stats) -- walks each
parameter in turn from the current best (coordinate descent, two passes), and
leaves the winner set on the graph.  Results are cached in a JSON file keyed
by device, workload, graph size and weighting, so a later process can apply
them without timing again (``apply_cached``).
"""
from __future__ import annotations

import ctypes as C
import json
import os
from pathlib import Path
from typing import Dict, Optional

import numpy as np

from . import _lib
from ._lib import call

DEFAULT_CACHE = Path(os.environ.get("GT_HIP_TUNING_CACHE",
                                    str(Path.home() / ".cache" / "gt_hip" / "tuning.json")))

SWEEP_SPACE = {
    "sweep_blocks_per_cu": (2, 3, 4, 6, 8),
    "tile_rows": (64, 128, 256),
    "short_row": (32, 64, 256),
    "sweep_hot_kb": (128, 64, -1),
    "sliced_min_len": (-1, 64, 256),
    "block_tile_nnz": (131072, 65536),
    "block_per_cu": (2, 1),
}
BATCHED_SPACE = {
    "mm_tile_nnz": (512, 1024, 2048, 4096),
    "mm_blocks_per_cu": (2, 4, 8),
}


def _device_name() -> str:
    try:
        import torch
        if torch.cuda.is_available():
            return torch.cuda.get_device_name(0).replace("|", "/")
    except Exception:  # torch is optional here
        pass
    return "hip"


def _key(dev, workload: str, weighted: bool) -> str:
    n, m, eb, _ = dev.info()
    return f"{_device_name()}|{workload}|n={n}|m={m}|eb={eb}|weighted={int(weighted)}"


def _load(path: Path) -> dict:
    try:
        return json.loads(Path(path).read_text())
    except (OSError, ValueError):
        return {}


def _store(path: Path, key: str, rec: dict) -> None:
    path = Path(path)
    data = _load(path)
    data[key] = rec
    path.parent.mkdir(parents=True, exist_ok=True)
    tmp = path.with_suffix(".tmp")
    tmp.write_text(json.dumps(data, indent=1, sort_keys=True))
    os.replace(tmp, path)


class _PagerankProbe:
    """Runs capped FP32-form and FP64 PageRank solves and returns the mean FP32
    sweep kernel time plus the mean FP64 sweep kernel time (seconds)."""

    def __init__(self, dev, weight, sweeps: int, reps: int):
        from . import _prop, Unity
        self.dev, self.sweeps, self.reps = dev, sweeps, reps
        n = dev.info()[0]
        self.keep = []
        self.wp = _prop(weight if weight is not None else Unity.NEW, dev.graph_id, self.keep)
        self.params = _lib.GtPowerIteration(1e-300, sweeps, _lib.GtPlan(n, max(1, -(-n // 4096))))
        self.view = _lib.GtView(_lib.GT_VIEW_DIRECTED, 0, None, 0, None, 0)
        self.rank = np.zeros(max(n, 1), np.float64)
        self.out = _lib.GtVertexOut(self.rank.ctypes.data_as(C.c_void_p), n, 0, _lib.GT_HOST, 0,
                                    dev.graph_id)
        self.iters = C.c_uint64()

    def _solve(self, fn: str):
        self.out.had = 0
        call(fn, self.dev.handle, C.byref(self.view), C.byref(self.wp), None,
             0.85, C.byref(self.params), C.byref(self.out), C.byref(self.iters))

    def __call__(self) -> float:
        self._solve("gt_pagerank_mixed")  # warm-up: rebuilds the schedule after a tile change
        self.dev.set_profiling(True)
        for _ in range(self.reps):
            self._solve("gt_pagerank_mixed")
            self._solve("gt_pagerank_f64")
        ms32, la32, _ = self.dev.sweep_stats(_lib.GT_SWEEP_PAGERANK_F32)
        ms64, la64, _ = self.dev.sweep_stats(_lib.GT_SWEEP_PAGERANK_F64)
        self.dev.set_profiling(False)
        # The mixed solve's FP64 refinement sweep counts among the FP64 launches.
        return ms32 * 1e-3 / max(la32, 1) + ms64 * 1e-3 / max(la64, 1)


class _BatchedProbe:
    """Runs capped k = 64 batched solves and returns the mean SpMM sweep time."""

    def __init__(self, dev, weight, sweeps: int, reps: int, k: int = 64):
        from . import _prop, Unity
        self.dev, self.reps = dev, reps
        n = dev.info()[0]
        self.keep = []
        self.wp = _prop(weight if weight is not None else Unity.NEW, dev.graph_id, self.keep)
        self.params = _lib.GtPowerIteration(1e-300, sweeps, _lib.GtPlan(n, max(1, -(-n // 4096))))
        self.view = _lib.GtView(_lib.GT_VIEW_DIRECTED, 0, None, 0, None, 0)
        rng = np.random.default_rng(0)
        self.k = min(k, n)
        self.seeds = rng.choice(n, self.k, replace=False).astype(np.uint32)
        self.x = np.zeros((n, self.k), np.float32)
        self.iters = np.zeros(self.k, np.uint64)

    def _solve(self):
        call("gt_pagerank_batched", self.dev.handle, C.byref(self.view), C.byref(self.wp),
             self.seeds.ctypes.data_as(C.POINTER(C.c_uint32)), self.k, 0.85,
             C.byref(self.params), self.x.ctypes.data_as(C.c_void_p), self.k, _lib.GT_HOST,
             self.iters.ctypes.data_as(C.POINTER(C.c_uint64)))

    def __call__(self) -> float:
        self._solve()
        self.dev.set_profiling(True)
        for _ in range(self.reps):
            self._solve()
        ms, launches, _ = self.dev.sweep_stats(_lib.GT_SWEEP_BATCHED)
        self.dev.set_profiling(False)
        return ms * 1e-3 / max(launches, 1)


def _descend(dev, space: Dict[str, tuple], probe, passes: int = 2):
    best = {k: dev.tuning()[k] for k in space}
    dev.set_tuning(**best)
    best_t = probe()
    trials = [(dict(best), best_t)]
    for _ in range(passes):
        moved = False
        for name, values in space.items():
            for v in values:
                if v == best[name]:
                    continue
                cand = dict(best, **{name: v})
                dev.set_tuning(**cand)
                t = probe()
                trials.append((cand, t))
                if t < best_t * 0.99:  # a change must win by more than the noise
                    best, best_t, moved = cand, t, True
        if not moved:
            break
    dev.set_tuning(**best)
    return best, best_t, trials


def autotune(dev, workload: str = "pagerank", weight=None, cache: Optional[Path] = DEFAULT_CACHE,
             sweeps: int = 6, reps: int = 3) -> dict:
    """Tune ``dev`` for ``workload`` ("pagerank": the pull sweeps; "batched":
    the k = 64 SpMM), leave the best parameters set on it, cache them, and
    return {"tuning", "sweep_us", "default_sweep_us", "trials"}; "sweep_us" is the
    probe's objective (for "pagerank" the FP32 plus the FP64 sweep time)."""
    weighted = weight is not None
    if workload == "pagerank":
        probe, space = _PagerankProbe(dev, weight, sweeps, reps), SWEEP_SPACE
    elif workload == "batched":
        probe, space = _BatchedProbe(dev, weight, max(2, sweeps // 2), reps), BATCHED_SPACE
    else:
        raise ValueError(f"unknown workload {workload!r}")
    dev.set_tuning(**{k: 0 for k in space})  # start from the defaults
    default_t = probe()
    best, best_t, trials = _descend(dev, space, probe)
    rec = {"tuning": best, "sweep_us": round(best_t * 1e6, 2),
           "default_sweep_us": round(default_t * 1e6, 2),
           "trials": [{"tuning": c, "sweep_us": round(t * 1e6, 2)} for c, t in trials]}
    if cache is not None:
        _store(Path(cache), _key(dev, workload, weighted), rec)
    return rec


def apply_cached(dev, workload: str = "pagerank", weighted: bool = False,
                 cache: Path = DEFAULT_CACHE) -> Optional[dict]:
    """Set the cached tuning for this graph and workload, if there is one."""
    rec = _load(Path(cache)).get(_key(dev, workload, weighted))
    if rec:
        dev.set_tuning(**rec["tuning"])
        return rec["tuning"]
    return None
