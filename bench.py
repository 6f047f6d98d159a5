"""PageRank / eigenvector benchmark on one MI355X.

Workload: R-MAT scale 22 (4,194,304 vertices), edge factor 16 (67,108,864
directed edges), a=0.57 b=0.19 c=0.19 d=0.05, SplitMix64-seeded, unweighted;
PageRank with damping 0.85, epsilon 1e-9, uniform personalisation, cold start
per step (a fresh rank map, as the reference's criterion bench does).

One step = one full PageRank solve with the graph resident in HBM.  `value` is
the throughput of the FP32-value / FP64-residual form in GTEPS (edges x sweeps
/ second); the plain FP64 drop-in is timed beside it, and eigenvector
centrality on the symmetrised graph after it.  The CPU baseline is the oracle
(the C restatement of the reference's rayon path) on the host cores, which
also provides the FP64 answer every GPU result's L1 error is reported against.

Algorithmic bytes per sweep (pull form, 4-byte indices, value width s):
  PageRank  B = 4m + 4(n+1) + n(4s + 4)        (s = 8: FP64 drop-in, 4: FP32 form)
  Eigenvector (symmetrised, m_s = 2m)  B = 4 m_s + 4(n+1) + 24 n
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
for p in (ROOT, ROOT / "graph-tool-rs_amd"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

HBM_PEAK_GBS = 8000.0
METRIC = "PageRank GTEPS + HBM GB/s, FP32/batched vs FP64 plain GPU at L1≤1e-6"
SEED = 0x243F6A8885A308D3


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def cpu_share() -> int:
    """Host threads the CPU baseline runs on: every core this process may
    use (its affinity mask), capped by OMP_NUM_THREADS when that is set -- a
    GPU box gives one GPU's job a share of its cores and says so there."""
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        usable = min(usable, int(omp))
    return max(1, usable)


def host_info() -> dict:
    """nproc, the cores this process may use, OMP_NUM_THREADS and the CPU model."""
    model = "unknown"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = None
    return {"nproc": os.cpu_count(), "affinity": usable,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS"), "cpu_model": model}


def step_times(torch, steps: int, warmup: int, solve):
    """`warmup` untimed calls, then `steps` calls each bracketed by a device
    synchronisation; returns (per-step seconds, last result)."""
    out = None
    for _ in range(warmup):
        out = solve()
    times = []
    for _ in range(steps):
        torch.cuda.synchronize()
        t = time.perf_counter()
        out = solve()
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t)
    return times, out


def timing_stats(times) -> dict:
    a = np.asarray(times)
    return {"median_ms": round(float(np.median(a)) * 1e3, 4),
            "mean_ms": round(float(a.mean()) * 1e3, 4),
            "min_ms": round(float(a.min()) * 1e3, 4), "max_ms": round(float(a.max()) * 1e3, 4),
            "steps": int(a.shape[0])}


def copy_calibration(torch, mib: int = 1024, reps: int = 5) -> dict:
    """Time a plain streaming kernel (torch.neg on `mib` MiB of f32: reads and
    writes every byte once) on the current stream.  Its achieved GB/s is the
    practical HBM ceiling on this box; under rocprofv3 --pmc the same dispatch
    calibrates FETCH_SIZE / WRITE_SIZE against a known byte count."""
    x = torch.ones(mib * (1 << 20) // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    torch.neg(x, out=y)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        torch.neg(x, out=y)
    e1.record()
    torch.cuda.synchronize()
    sec = e0.elapsed_time(e1) * 1e-3 / reps
    nbytes = 2 * x.numel() * 4
    del x, y
    return {"kernel": "torch.neg f32 (read + write)", "bytes": nbytes,
            "us": round(sec * 1e6, 1), "gbs": round(nbytes / sec / 1e9, 1)}


def profiled_traffic(workload: str, kernel: str):
    """HBM bytes per launch of `kernel` on `workload`, from the committed
    rocprofv3 --pmc summary (FETCH_SIZE and WRITE_SIZE, each scaled by what
    the counter reads for the copy kernel of known size), or None."""
    f = ROOT / "profiles" / f"traffic_{workload}.json"
    if not f.exists():
        return None, None
    rec = json.loads(f.read_text()).get("kernels", {}).get(kernel)
    if not rec:
        return None, None
    return int(rec["traffic_bytes"]), f"profiles/{f.name}"


def gather_rate() -> dict:
    """The random-gather rate of this device (libgt_probe.so, one workgroup of
    eight wavefronts per CU walking 65,536 tiles of 1,024 labels with every
    lane gathering): nanoseconds per gathered lane per compute unit, with the
    1 MB window resident in every XCD's L2, and with labels spread over a
    16 MB window (the FP32 contribution vector of R-MAT 22)."""
    lib = C.CDLL(str(ROOT / "graph-tool-rs_amd" / "gt_hip" / "libgt_probe.so"))
    lib.gt_probe_gather.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                    C.POINTER(C.c_double), C.POINTER(C.c_uint32)]
    lib.gt_probe_error.restype = C.c_char_p
    tiles, reps = 65536, 5
    out = {"tiles": tiles, "entries_per_tile": 1024, "kernel": "gather_walk (libgt_probe.so)"}
    for name, window in (("l2_resident_1mb", 1 << 18), ("uniform_16mb", 1 << 22)):
        us, cus = C.c_double(), C.c_uint32()
        if lib.gt_probe_gather(window, tiles, 1024, reps, C.byref(us), C.byref(cus)) != 0:
            raise RuntimeError(lib.gt_probe_error().decode())
        out[name] = {"us_per_launch": round(us.value, 2),
                     "ns_per_lane_per_cu": round(us.value * 1e3 * cus.value / (tiles * 1024), 4)}
        out["compute_units"] = cus.value
    return out


def cold_gathers(src, n: int, nhot: int) -> int:
    """Entries whose source is outside the nhot most gathered vertices (the
    relabelled order): the gathers a hot-slice sweep sends to memory."""
    deg = np.bincount(src, minlength=n)
    top = np.sort(deg)[::-1][:nhot]
    return int(deg.sum() - top.sum())


def hot_len(kb: int, vbytes: int, n: int) -> int:
    """The hot slice the library uses for a tuning value (0: 128 KB)."""
    if kb < 0:
        return 0
    kb = 128 if kb == 0 else min(kb, 128)
    cap = 32704 if vbytes == 4 else 12224
    return min(kb * 1024 // vbytes, cap, n) & ~3


def profiled_misses(workload: str, kernel: str):
    """L2 misses per launch of `kernel` from the committed counter profile."""
    f = ROOT / "profiles" / f"traffic_{workload}.json"
    if not f.exists():
        return None
    rec = json.loads(f.read_text()).get("kernels", {}).get(kernel)
    return rec.get("tcc_miss") if rec else None


def pr_bytes(n: int, m: int, s: int) -> int:
    return 4 * m + 4 * (n + 1) + n * (4 * s + 4)


def ev_bytes(n: int, ms: int) -> int:
    return 4 * ms + 4 * (n + 1) + 24 * n


def _upload_rmat(gt_hip, scale: int, edgefactor: int, gid: int):
    n = 1 << scale
    m = edgefactor * n
    t0 = time.perf_counter()
    src, tgt = gt_hip.gen_rmat(scale, m, 0.57, 0.19, 0.19, SEED)
    adj = gt_hip.host_adjacency(n, src, tgt)
    dev = gt_hip.DeviceGraph.upload(gid, n, m, *adj)
    log(f"[bench] R-MAT scale {scale}: n={n} m={m} built and uploaded in "
        f"{time.perf_counter()-t0:.1f}s")
    return n, m, src, tgt, dev


def _tune(args, dev, workload: str, weight) -> dict:
    """Autotune the graph's launch parameters before the timed region (the
    sweeps timed afterwards run with the winner); report what was chosen."""
    if args.no_tune:
        return {"autotuned": False, "params": dev.tuning()}
    from gt_hip import autotune
    t0 = time.perf_counter()
    rec = autotune.autotune(dev, workload, weight=weight, cache=None)
    what = "FP32 + FP64 sweep" if workload == "pagerank" else "SpMM sweep"
    log(f"[bench] autotune ({workload}): {rec['tuning']} -> {rec['sweep_us']:.1f} us {what} "
        f"(defaults {rec['default_sweep_us']:.1f}), {len(rec['trials'])} trials in "
        f"{time.perf_counter()-t0:.1f}s")
    return {"autotuned": True, "params": dev.tuning(), "sweep_us": rec["sweep_us"],
            "default_sweep_us": rec["default_sweep_us"], "trials": len(rec["trials"])}


def _timed_solves(torch, dev, kind: int, steps: int, warmup: int, solve):
    """Per-step timed solves, then the same solves again with the library's
    sweep events on for the kernel times (an event around every sweep costs
    the solve a few microseconds per sweep, so the timed solves run without
    them); returns (median step x steps, outputs, summed kernel ms, launches,
    step stats)."""
    for _ in range(warmup):
        solve()
    times, outs = [], []
    for _ in range(steps):
        torch.cuda.synchronize()
        t = time.perf_counter()
        outs.append(solve())
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t)
    dev.set_profiling(True)
    for _ in range(steps):
        solve()
    torch.cuda.synchronize()
    ms_k, launches, _ = dev.sweep_stats(kind)
    dev.set_profiling(False)
    return float(np.median(times)) * steps, outs, ms_k, launches, timing_stats(times)


def bench_batched(args, torch, gt_hip, _lib) -> int:
    """k = 64 personalised PageRank vectors as one SpMM per sweep.

    R-MAT scale 20 (1,048,576 vertices, 16,777,216 edges), FP32 weights
    exp(N(0,1)), 64 distinct uniform seed vertices, damping 0.85, epsilon 1e-6
    per column.  GTEPS counts m per active column per sweep.  Algorithmic bytes
    per sweep: B = 8m + 4(n+1) + 16nk + 4n (index + weight stream, offsets,
    state gather / read / write and contribution write at [n][64] FP32,
    out-degree)."""
    scale = 20 if args.scale is None else args.scale
    k = 64
    n, m, src, tgt, dev = _upload_rmat(gt_hip, scale, args.edgefactor, 1)
    rng = np.random.default_rng(SEED & 0xFFFFFFFF)
    w_host = np.exp(rng.standard_normal(m)).astype(np.float32)
    seeds = rng.choice(n, k, replace=False).astype(np.uint32)
    stream = torch.cuda.current_stream()
    dev.set_stream(stream.cuda_stream)
    copy = copy_calibration(torch)
    w = torch.from_numpy(w_host).cuda()
    X = torch.empty((n, k), dtype=torch.float32, device="cuda")
    wprop = _lib.GtProp(_lib.GT_MAP_DENSE, _lib.GT_F32, _lib.GT_DEVICE, 0, 0.0,
                        C.c_void_p(w.data_ptr()), m, 1)
    tuning = _tune(args, dev, "batched", w)
    eps = 1e-6
    params = _lib.GtPowerIteration(eps, 0, _lib.GtPlan(n, -(-n // 4096)))
    view = _lib.GtView(_lib.GT_VIEW_DIRECTED, 0, None, 0, None, 0)
    iters = np.zeros(k, np.uint64)

    def solve():
        _lib.call("gt_pagerank_batched", dev.handle, C.byref(view), C.byref(wprop),
                  seeds.ctypes.data_as(C.POINTER(C.c_uint32)), k, 0.85, C.byref(params),
                  C.c_void_p(X.data_ptr()), k, _lib.GT_DEVICE,
                  iters.ctypes.data_as(C.POINTER(C.c_uint64)))
        return iters.copy()

    dt, outs, ms_k, launches, tstats = _timed_solves(torch, dev, _lib.GT_SWEEP_BATCHED,
                                                     args.steps, args.warmup, solve)
    it = outs[-1].astype(np.int64)
    col_sweeps = int(it.sum())
    sweeps = int(it.max())
    gteps = m * col_sweeps * args.steps / dt / 1e9
    # Algorithmic bytes of one solve: every sweep streams the tiles' entries
    # and weights (8m), the row offsets and out-degrees (8n); each (column,
    # sweep) with the column still iterating reads and writes its x and c
    # (16n); a column that stops before the last sweep is copied once (16n).
    copied = int((it < sweeps).sum())
    B_solve = sweeps * (8 * m + 4 * (n + 1) + 4 * n) + 16 * n * (col_sweeps + copied)
    B = B_solve // max(sweeps, 1)
    avg = ms_k / max(launches, 1) * 1e-3
    ach = B_solve * args.steps / (ms_k * 1e-3) / 1e9 if launches else 0.0
    log(f"[bench] batched: {dt/args.steps*1e3:.2f} ms/solve, sweeps {sweeps} (columns "
        f"{it.min()}..{it.max()}), {launches} launches ({avg*1e6:.1f} us avg)")
    result = {
        "metric": "Batched personalised PageRank GTEPS (k=64 SpMM, FP32 state, FP64 residuals)",
        "value": round(gteps, 3), "unit": "GTEPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dt / args.steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 state/weights/SpMM, f64 teleport/dangling/L1", "data": "synthetic",
        "config": {"workload": f"batched_ppr_rmat{scale}_ef{args.edgefactor}_k{k}",
                   "vertices": n, "edges": m, "k": k, "weights": "f32 exp(N(0,1))",
                   "seeds": "64 distinct uniform", "damping": 0.85, "epsilon": eps},
        "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                     "traffic": profiled_traffic(
                         f"batched_ppr_rmat{scale}_ef{args.edgefactor}_k{k}", "ppr_sweep_w")[0],
                     "traffic_source": profiled_traffic(
                         f"batched_ppr_rmat{scale}_ef{args.edgefactor}_k{k}", "ppr_sweep_w")[1],
                     "kernel": "ppr_sweep_w<true, 16> (batched SpMM sweep, wave tiles)",
                     "bytes_per_sweep": B, "bytes_per_solve": B_solve,
                     "avg_launch_us": round(avg * 1e6, 2), "copy_calibration": copy},
        "column_sweeps": {"min": int(it.min()), "max": int(it.max()), "sum": col_sweeps},
        "timing": tstats,
        "tuning": tuning,
    }
    if not args.no_cpu:
        import oracle
        threads = cpu_share()
        o = oracle.Adj(n, src, tgt)
        w64 = w_host.astype(np.float64)
        Xh = X.cpu().numpy()
        t0 = time.perf_counter()
        errs, sw_ok, done = [], 0, 0
        for c in range(k):
            p = np.zeros(n)
            p[seeds[c]] = 1.0
            want, ito = o.pagerank(weight=w64, pers=p, epsilon=eps, max_iter=0, grain=4096,
                                   threads=threads)
            errs.append(float(np.abs(Xh[:, c].astype(np.float64) - want).sum()))
            sw_ok += int(abs(ito - int(it[c])) <= 1)
            done += 1
            if time.perf_counter() - t0 >= 10.0:
                break
        dtc = time.perf_counter() - t0
        cpu_sweeps = sum(int(it[c]) for c in range(done))
        result["cpu_baseline"] = {
            "value": round(m * cpu_sweeps / dtc / 1e9, 4), "unit": "GTEPS", "cores": threads,
            "kind": "port", "host": host_info(),
            "sample": f"{done} of the 64 columns solved one at a time by the oracle's "
                      f"pagerank_with (one-hot pers, Plan grain 4096), {dtc:.1f}s"}
        result["accuracy_vs_fp64_oracle"] = {
            "columns_checked": done, "max_l1": max(errs),
            "sweeps_within_1": f"{sw_ok}/{done}"}
    print(json.dumps(result), flush=True)
    dev.close()
    return 0


def bench_adjoint(args, torch, gt_hip, _lib) -> int:
    """Weighted PageRank forward (FP32 form) plus its adjoint.

    R-MAT scale 22 (unless --scale), FP32 weights 1 - 0.9 U[0,1) in (0.1, 1],
    loss L = c . x with c ~ N(0,1); gradients for every edge weight, the
    personalisation and the damping.  One step = one forward solve plus one
    backward (transposed solve and the fused gradient pass); the value is the
    backward / forward time ratio.  Forward algorithmic bytes per FP32 sweep:
    8m + 4(n+1) + 20n; adjoint sweep the same on the CSR; gradient pass
    12m + 16n once."""
    scale = 22 if args.scale is None else args.scale
    n, m, src, tgt, dev = _upload_rmat(gt_hip, scale, args.edgefactor, 1)
    rng = np.random.default_rng((SEED >> 7) & 0xFFFFFFFF)
    w_host = (1.0 - 0.9 * rng.random(m)).astype(np.float32)
    c_host = rng.standard_normal(n)
    stream = torch.cuda.current_stream()
    dev.set_stream(stream.cuda_stream)
    w = torch.from_numpy(w_host).cuda()
    cvec = torch.from_numpy(c_host).cuda()
    x = torch.empty(n, dtype=torch.float64, device="cuda")
    gw = torch.empty(m, dtype=torch.float32, device="cuda")
    gp = torch.empty(n, dtype=torch.float64, device="cuda")
    gd = C.c_double()
    wprop = _lib.GtProp(_lib.GT_MAP_DENSE, _lib.GT_F32, _lib.GT_DEVICE, 0, 0.0,
                        C.c_void_p(w.data_ptr()), m, 1)
    xprop = _lib.GtProp(_lib.GT_MAP_DENSE, _lib.GT_F64, _lib.GT_DEVICE, 0, 0.0,
                        C.c_void_p(x.data_ptr()), n, 1)
    cprop = _lib.GtProp(_lib.GT_MAP_DENSE, _lib.GT_F64, _lib.GT_DEVICE, 0, 0.0,
                        C.c_void_p(cvec.data_ptr()), n, 1)
    eps = 1e-9
    params = _lib.GtPowerIteration(eps, 0, _lib.GtPlan(n, -(-n // 4096)))
    view = _lib.GtView(_lib.GT_VIEW_DIRECTED, 0, None, 0, None, 0)
    out = _lib.GtVertexOut(C.c_void_p(x.data_ptr()), n, 0, _lib.GT_DEVICE, 0, 1)
    iters = C.c_uint64()
    biters = C.c_uint64()

    def forward():
        out.had = 0
        _lib.call("gt_pagerank_mixed", dev.handle, C.byref(view), C.byref(wprop), None, 0.85,
                  C.byref(params), C.byref(out), C.byref(iters))
        return int(iters.value)

    def backward():
        _lib.call("gt_pagerank_backward", dev.handle, C.byref(view), C.byref(wprop), None, 0.85,
                  C.byref(params), C.byref(xprop), C.byref(cprop), _lib.GT_F32,
                  C.c_void_p(gw.data_ptr()), _lib.GT_F32, _lib.GT_DEVICE,
                  C.c_void_p(gp.data_ptr()), C.byref(gd),
                  C.byref(biters))
        return int(biters.value)

    dtf, fo, ms_f, la_f, tsf = _timed_solves(torch, dev, _lib.GT_SWEEP_PAGERANK_F32,
                                             args.steps, args.warmup, forward)
    dtb, bo, ms_b, la_b, tsb = _timed_solves(torch, dev, _lib.GT_SWEEP_ADJOINT, args.steps,
                                             args.warmup, backward)
    ratio = dtb / dtf
    blk_b = dev.layout()["block_tiles"] > 0
    kb_ = ("pr_sweep_blk<false, 4, true>" if blk_b
           else "adj_sweep_hot<true, false>" if dev.tuning()["sweep_hot_kb"] > 0
           else "adj_sweep<float, true, false>")
    Bf = 8 * m + 4 * (n + 1) + 20 * n
    ach_b = Bf * bo[-1] * args.steps / (ms_b * 1e-3) / 1e9 if la_b else 0.0
    log(f"[bench] adjoint: forward {dtf/args.steps*1e3:.2f} ms ({fo[-1]} sweeps), backward "
        f"{dtb/args.steps*1e3:.2f} ms ({bo[-1]} sweeps), ratio {ratio:.3f}")
    result = {
        "metric": "Adjoint cost / forward solve, weighted PageRank (FP32 form)",
        "value": round(ratio, 4), "unit": "x forward", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round((dtf + dtb) / args.steps * 1e3, 4),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 values/weights/SpMV, f64 residuals and sums", "data": "synthetic",
        "config": {"workload": f"adjoint_rmat{scale}_ef{args.edgefactor}", "vertices": n,
                   "edges": m, "weights": "f32 1 - 0.9 U[0,1)", "loss": "c . x, c ~ N(0,1)",
                   "damping": 0.85, "epsilon": eps},
        "forward_ms": round(dtf / args.steps * 1e3, 4), "forward_sweeps": fo[-1],
        "backward_ms": round(dtb / args.steps * 1e3, 4), "backward_sweeps": bo[-1],
        "roofline": {"bound": "hbm", "achieved": round(ach_b, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(ach_b / HBM_PEAK_GBS, 4),
                     "traffic": profiled_traffic(f"adjoint_rmat{scale}_ef{args.edgefactor}",
                                                 kb_)[0],
                     "kernel": kb_ + (" (adjoint sweep on block tiles, weights in block order)"
                                      if blk_b else " (adjoint sweep, wave engine)"),
                     "bytes_per_sweep": Bf,
                     "avg_launch_us": round(ms_b / max(la_b, 1) * 1e3, 2)},
        "grad_damping": gd.value,
        "timing": {"forward": tsf, "backward": tsb},
    }
    if args.fd_edges > 0:
        result["finite_differences"] = _adjoint_fd(args, torch, _lib, dev, n, m, w, cvec, view,
                                                   gw)
    print(json.dumps(result), flush=True)
    dev.close()
    return 0


def _adjoint_fd(args, torch, _lib, dev, n, m, w32, cvec, view, gw32) -> dict:
    """Central differences of L = c . x over `args.fd_edges` uniformly drawn
    edges (h = 1e-4 on the weight, FP64 forward solves at epsilon 1e-14)
    against the FP64 adjoint's gradient, and the FP32 adjoint's against the
    FP64 one.  The noise floor is what the forward's stopping leaves in
    L(w + h) - L(w - h): |L at 1e-14 - L converged further| / h."""
    h, eps_fd = 1e-4, 1e-14
    w64 = w32.double()
    x = torch.empty(n, dtype=torch.float64, device="cuda")
    wprop = _lib.GtProp(_lib.GT_MAP_DENSE, _lib.GT_F64, _lib.GT_DEVICE, 0, 0.0,
                        C.c_void_p(w64.data_ptr()), m, 1)
    out = _lib.GtVertexOut(C.c_void_p(x.data_ptr()), n, 0, _lib.GT_DEVICE, 0, 1)
    it = C.c_uint64()

    def loss(eps=eps_fd, cap=0) -> float:
        params = _lib.GtPowerIteration(eps, cap, _lib.GtPlan(n, -(-n // 4096)))
        out.had = 0
        _lib.call("gt_pagerank_f64", dev.handle, C.byref(view), C.byref(wprop), None, 0.85,
                  C.byref(params), C.byref(out), C.byref(it))
        return float(torch.dot(cvec, x).item())

    t0 = time.perf_counter()
    base = loss()
    sweeps_fd = int(it.value)
    deeper = loss(1e-300, sweeps_fd + 40)
    noise = abs(base - deeper) / h
    # The FP64 adjoint at the FP64 fixed point.
    loss()
    gw = torch.empty(m, dtype=torch.float64, device="cuda")
    gp = torch.empty(n, dtype=torch.float64, device="cuda")
    gd = C.c_double()
    bit = C.c_uint64()
    xprop = _lib.GtProp(_lib.GT_MAP_DENSE, _lib.GT_F64, _lib.GT_DEVICE, 0, 0.0,
                        C.c_void_p(x.data_ptr()), n, 1)
    cprop = _lib.GtProp(_lib.GT_MAP_DENSE, _lib.GT_F64, _lib.GT_DEVICE, 0, 0.0,
                        C.c_void_p(cvec.data_ptr()), n, 1)
    bparams = _lib.GtPowerIteration(1e-14, 0, _lib.GtPlan(n, -(-n // 4096)))
    _lib.call("gt_pagerank_backward", dev.handle, C.byref(view), C.byref(wprop), None, 0.85,
              C.byref(bparams), C.byref(xprop), C.byref(cprop), _lib.GT_F64,
              C.c_void_p(gw.data_ptr()), _lib.GT_F64, _lib.GT_DEVICE,
              C.c_void_p(gp.data_ptr()), C.byref(gd), C.byref(bit))
    g64 = gw.cpu().numpy()
    g32 = gw32.cpu().numpy().astype(np.float64)
    rng = np.random.default_rng(20240601)
    edges = rng.choice(m, args.fd_edges, replace=False)
    fd = np.zeros(len(edges))
    for i, e in enumerate(edges.tolist()):
        w0 = float(w64[e].item())
        w64[e] = w0 + h
        lp = loss()
        w64[e] = w0 - h
        lm = loss()
        w64[e] = w0
        fd[i] = (lp - lm) / (2 * h)
    ga = g64[edges]
    scale = float(np.abs(g64).max())
    abs_err = np.abs(fd - ga)
    rel = abs_err / np.maximum(np.abs(ga), 1e-300)
    dt = time.perf_counter() - t0
    log(f"[bench] adjoint FD: {len(edges)} edges, median rel err {np.median(rel):.3e}, "
        f"max abs err {abs_err.max():.3e} (noise floor {noise:.3e}, max |grad| {scale:.3e}), "
        f"{dt:.1f}s")
    return {"edges": int(len(edges)), "h": h, "forward_epsilon": eps_fd,
            "forward_sweeps": sweeps_fd, "adjoint_sweeps": int(bit.value),
            "fd_vs_fp64_adjoint": {
                "median_rel": float(np.median(rel)), "p90_rel": float(np.percentile(rel, 90)),
                "max_abs": float(abs_err.max()), "max_abs_over_max_grad": float(abs_err.max() / scale),
                "noise_floor_abs": noise,
                "within_noise_floor": int((abs_err <= 2.0 * noise + 1e-6 * np.abs(ga)).sum())},
            "fp32_vs_fp64_adjoint_max_rel": float(np.abs(g32 - g64).max() / scale),
            "seconds": round(dt, 1)}



def bench_uneven(args, torch, gt_hip, _lib) -> int:
    """An uneven batch: 4096 directed Chung-Lu graphs, vertex counts
    log-uniform in [64, 65536], expected out-degrees a power law (exponent
    2.1) with mean 8, FP32 weights in (0, 1], damping uniform in [0.5, 0.95],
    epsilon 1e-6; PageRank forward plus the adjoint of L = c . x (c ~ N(0,1))
    for every graph, each call one hipGraph-captured sweep loop.  One step =
    forward + backward.  GTEPS counts each graph's edges once per sweep it ran
    (forward and adjoint sweeps); algorithmic bytes per graph sweep:
    8 m_g + 4 (n_g + 1) + 20 n_g."""
    G = 4096
    t0 = time.perf_counter()
    voff, eoff, damp, src, tgt, w_host = gt_hip.gen_chung_lu_batch(
        G, SEED, 64, 65536, 8.0, 2.1, 0.5, 0.95)
    n, m = int(voff[-1]), int(src.shape[0])
    t_gen = time.perf_counter() - t0
    t0 = time.perf_counter()
    dev = gt_hip.DeviceGraph.build(n, src, tgt, graph_id=1)
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t0
    log(f"[bench] uneven batch: {G} graphs, n={n} m={m}, generated in {t_gen:.1f}s, "
        f"device build {t_build:.2f}s")
    rng = np.random.default_rng(SEED & 0xFFFFFFFF)
    c_host = rng.standard_normal(n)
    w = torch.from_numpy(w_host).cuda()
    c = torch.from_numpy(c_host).cuda()
    x = torch.empty(n, dtype=torch.float64, device="cuda")
    gw = torch.empty(m, dtype=torch.float32, device="cuda")
    gp = torch.empty(n, dtype=torch.float64, device="cuda")
    gd = np.zeros(G)
    wprop = _lib.GtProp(_lib.GT_MAP_DENSE, _lib.GT_F32, _lib.GT_DEVICE, 0, 0.0,
                        C.c_void_p(w.data_ptr()), m, 1)
    eps = 1e-6
    params = _lib.GtPowerIteration(eps, 0, _lib.GtPlan(n, -(-n // 4096)))
    it_f = np.zeros(G, np.uint64)
    it_b = np.zeros(G, np.uint64)
    u32 = lambda a: a.ctypes.data_as(C.POINTER(C.c_uint32))
    f64 = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    u64 = lambda a: a.ctypes.data_as(C.POINTER(C.c_uint64))

    def forward():
        _lib.call("gt_pagerank_uneven", dev.handle, G, u32(voff), f64(damp), C.byref(wprop),
                  C.byref(params), C.c_void_p(x.data_ptr()), _lib.GT_DEVICE, u64(it_f))

    def backward():
        _lib.call("gt_pagerank_uneven_backward", dev.handle, G, u32(voff), f64(damp),
                  C.byref(wprop), C.byref(params), C.c_void_p(x.data_ptr()),
                  C.c_void_p(c.data_ptr()), _lib.GT_DEVICE, C.c_void_p(gw.data_ptr()),
                  C.c_void_p(gp.data_ptr()), f64(gd), u64(it_b))

    for _ in range(args.warmup):
        forward()
        backward()
    tf = tb = 0.0
    for _ in range(args.steps):
        torch.cuda.synchronize()
        t = time.perf_counter()
        forward()
        torch.cuda.synchronize()
        tf += time.perf_counter() - t
        t = time.perf_counter()
        backward()
        torch.cuda.synchronize()
        tb += time.perf_counter() - t
    # Block tiles cut at the graphs' bounds unless they cannot hold the batch.
    on_blocks = dev.layout()["block_tiles"] > 0
    tkey = "pr_sweep_blk_forward_call" if on_blocks else "u_sweep_forward_call"
    mg = np.diff(eoff).astype(np.float64)
    ng = np.diff(voff.astype(np.int64)).astype(np.float64)
    itf, itb = it_f.astype(np.float64), it_b.astype(np.float64)
    edges_f = float((mg * itf).sum())
    edges_b = float((mg * itb).sum())
    bytes_f = float(((8 * mg + 4 * (ng + 1) + 20 * ng) * itf).sum())
    dt = tf + tb
    gteps = (edges_f + edges_b) * args.steps / dt / 1e9
    ach = bytes_f * args.steps / tf / 1e9
    log(f"[bench] uneven: forward {tf/args.steps*1e3:.2f} ms (sweeps {int(itf.min())}.."
        f"{int(itf.max())}), backward {tb/args.steps*1e3:.2f} ms (sweeps {int(itb.min())}.."
        f"{int(itb.max())}), {gteps:.1f} GTEPS")
    result = {
        "metric": "Uneven batch PageRank forward + adjoint GTEPS (4096 Chung-Lu graphs)",
        "value": round(gteps, 3), "unit": "GTEPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dt / args.steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 values/weights/SpMV, f64 residuals", "data": "synthetic",
        "config": {"workload": "uneven_chung_lu_4096", "graphs": G, "vertices": n,
                   "edges": m, "n_range": [64, 65536], "exponent": 2.1, "mean_degree": 8,
                   "damping": [0.5, 0.95], "weights": "f32 1 - U[0,1)", "epsilon": eps},
        "forward_ms": round(tf / args.steps * 1e3, 4),
        "backward_ms": round(tb / args.steps * 1e3, 4),
        "backward_over_forward": round(tb / tf, 4),
        "forward_sweeps": {"min": int(itf.min()), "max": int(itf.max()),
                           "mean": round(float(itf.mean()), 2)},
        "adjoint_sweeps": {"min": int(itb.min()), "max": int(itb.max()),
                           "mean": round(float(itb.mean()), 2)},
        "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                     "traffic": profiled_traffic("uneven_chung_lu_4096", tkey)[0],
                     "traffic_unit": "bytes per forward call (its sweeps)",
                     "traffic_source": profiled_traffic("uneven_chung_lu_4096", tkey)[1],
                     "kernel": ("pr_sweep_blk<false, 6, true> (forward sweeps on block tiles cut at "
                                "the graphs' bounds; wall time of the call)" if on_blocks else
                                "u_sweep<UFwdFin<true>> (forward, wall time of the call)"),
                     "bytes_per_solve": int(bytes_f)},
    }
    if not args.no_cpu:
        import oracle
        threads = cpu_share()
        xh = x.cpu().numpy()
        t0 = time.perf_counter()
        done, edges, worst, sw_ok = 0, 0.0, 0.0, 0
        order = np.random.default_rng(1).permutation(G)
        for k in order:
            lo, hi = int(eoff[k]), int(eoff[k + 1])
            o = oracle.Adj(int(ng[k]), src[lo:hi] - voff[k], tgt[lo:hi] - voff[k])
            want, ito = o.pagerank(weight=w_host[lo:hi].astype(np.float64),
                                   damping=float(damp[k]), epsilon=eps, grain=4096,
                                   threads=threads)
            worst = max(worst, float(np.abs(xh[voff[k]:voff[k + 1]] - want).sum()))
            sw_ok += int(abs(ito - int(it_f[k])) <= 1)
            edges += mg[k] * ito
            done += 1
            if time.perf_counter() - t0 >= 10.0:
                break
        dtc = time.perf_counter() - t0
        result["cpu_baseline"] = {
            "value": round(edges / dtc / 1e9, 4), "unit": "GTEPS (forward only)",
            "cores": threads, "kind": "port", "host": host_info(),
            "sample": f"{done} random graphs of the batch, each solved alone by the oracle's "
                      f"pagerank_with (forward only, graph build included), {dtc:.1f}s"}
        result["accuracy_vs_fp64_oracle"] = {"graphs_checked": done, "max_l1": worst,
                                             "sweeps_within_1": f"{sw_ok}/{done}"}
    print(json.dumps(result), flush=True)
    dev.close()
    return 0



def bench_betweenness(args, torch, gt_hip, _lib) -> int:
    """All-pivot Brandes betweenness (vertex and edge maps).

    R-MAT scale 16 (unless --scale): 65,536 vertices, 1,048,576 edges, the
    same generator and seed as the PageRank workload, on the undirected view;
    every vertex with an incident edge is a source.  One step = one call.
    GTEPS = sources x edges / time (each BFS traverses the graph once).
    Algorithmic bytes, from the call's own work counters: 12 B per adjacency
    entry a level kernel reads (4 B index + 8 B lane mask), 8 B per forward
    (entry, source) hit (a path count), 16 B per backward hit (path count and
    dependency), and 32 B per vertex per level (lane masks in and out, level
    range, row offsets)."""
    scale = 16 if args.scale is None else args.scale
    n, m, src, tgt, dev = _upload_rmat(gt_hip, scale, args.edgefactor, 1)
    stream = torch.cuda.current_stream()
    dev.set_stream(stream.cuda_stream)
    copy = copy_calibration(torch)
    vb = torch.zeros(n, dtype=torch.float64, device="cuda")
    eb = torch.zeros(m, dtype=torch.float64, device="cuda")
    out = _lib.GtBcOut(C.c_void_p(vb.data_ptr()), n, C.c_void_p(eb.data_ptr()), m,
                       _lib.GT_DEVICE, 0, 1, 1)
    view = _lib.GtView(_lib.GT_VIEW_UNDIRECTED, 0, None, 0, None, 0)
    st = _lib.GtBcStats()

    def solve(pivots=None):
        pp = None if pivots is None else pivots.ctypes.data_as(C.POINTER(C.c_uint32))
        _lib.call("gt_betweenness", dev.handle, C.byref(view), pp,
                  0 if pivots is None else int(pivots.shape[0]), C.byref(out), C.byref(st))
        return {k: getattr(st, k) for k, _ in _lib.GtBcStats._fields_}

    steps = args.steps
    for _ in range(args.warmup):
        solve()
    torch.cuda.synchronize()
    t = time.perf_counter()
    timed = {"streams": 0}
    for _ in range(steps):
        timed = solve()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t
    dev.set_profiling(True)
    prof = solve()
    dev.set_profiling(False)
    sources = prof["sources"]
    gteps = sources * m * steps / dt / 1e9
    kern_ms = prof["forward_ms"] + prof["backward_ms"]
    B = (12 * (prof["forward_entries"] + prof["backward_entries"]) + 8 * prof["forward_hits"]
         + 16 * prof["backward_hits"] + 32 * n * prof["levels"])
    ach = B / (kern_ms * 1e-3) / 1e9 if kern_ms > 0 else 0.0
    log(f"[bench] betweenness: {dt/steps*1e3:.1f} ms/call, {sources} sources in "
        f"{prof['batches']} batches, {prof['levels']} levels, kernels {kern_ms:.1f} ms "
        f"(forward {prof['forward_ms']:.1f}, backward {prof['backward_ms']:.1f}), "
        f"{gteps:.2f} GTEPS")
    result = {
        "metric": "Betweenness GTEPS (all-pivot Brandes, vertex + edge maps, f64)",
        "value": round(gteps, 3), "unit": "GTEPS", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": round(dt / steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 path counts and dependencies, u32 levels", "data": "synthetic",
        "config": {"workload": f"betweenness_rmat{scale}_ef{args.edgefactor}_undirected",
                   "vertices": n, "edges": m, "view": "undirected", "pivots": "all",
                   "sources": sources, "rmat": [0.57, 0.19, 0.19, 0.05], "seed": SEED},
        "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "bc_forward + bc_backward level launches (one call, timed "
                               "per level group on one stream)",
                     "bytes_per_call": int(B), "kernel_ms_per_call": round(kern_ms, 3),
                     "achieved_over_call_wall": round(B / (dt / steps) / 1e9, 1),
                     "streams_per_timed_call": int(timed["streams"]),
                     "copy_calibration": copy},
        "work": prof,
    }
    if not args.no_cpu:
        import oracle
        threads = cpu_share()
        o = oracle.Adj(n, src, tgt)
        deg = np.bincount(src, minlength=n) + np.bincount(tgt, minlength=n)
        cand = np.flatnonzero(deg > 0).astype(np.uint32)
        rng = np.random.default_rng(5)
        k = 64
        while True:  # grow the pivot sample until the oracle takes ~10 s
            piv = np.sort(rng.choice(cand, min(k, cand.shape[0]), replace=False)).astype(np.uint32)
            t0 = time.perf_counter()
            wv, we = oracle.betweenness(o, undirected=True, pivots=piv, threads=threads)
            dtc = time.perf_counter() - t0
            if dtc >= 10.0 or k >= cand.shape[0] or dtc * (2 * k) / k > 40.0:
                break
            k *= 2
        # The device on the same pivot sample, for the accuracy line.
        solve(piv)
        torch.cuda.synchronize()
        gv, ge = vb.cpu().numpy(), eb.cpu().numpy()
        nz_v, nz_e = wv != 0, we != 0
        result["cpu_baseline"] = {
            "value": round(piv.shape[0] * m / dtc / 1e9, 4), "unit": "GTEPS", "cores": threads,
            "kind": "port", "host": host_info(),
            "sample": f"{piv.shape[0]} random pivots of the same graph and view through the "
                      f"oracle's Brandes (16 chunks, C restatement of the reference), {dtc:.1f}s"}
        result["accuracy_vs_oracle"] = {
            "pivots": int(piv.shape[0]),
            "vertex_max_rel": float((np.abs(gv - wv)[nz_v] / np.abs(wv[nz_v])).max()),
            "edge_max_rel": float((np.abs(ge - we)[nz_e] / np.abs(we[nz_e])).max()),
            "zeros_match": bool(np.all(gv[~nz_v] == 0) and np.all(ge[~nz_e] == 0))}
    print(json.dumps(result), flush=True)
    dev.close()
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload",
                    choices=("pagerank", "batched", "adjoint", "uneven", "betweenness"),
                    default="pagerank",
                    help="pagerank: R-MAT 22 FP32/FP64 solves (the headline); batched: k=64 "
                         "personalised SpMM on weighted R-MAT 20; adjoint: weighted R-MAT 22 "
                         "forward + gradient; uneven: 4096 Chung-Lu graphs, forward + adjoint; "
                         "betweenness: all-pivot Brandes on undirected R-MAT 16")
    ap.add_argument("--scale", type=int, default=None)
    ap.add_argument("--edgefactor", type=int, default=16)
    ap.add_argument("--epsilon", type=float, default=1e-9)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline / oracle")
    ap.add_argument("--fd-edges", type=int, default=1000,
                    help="adjoint workload: edges checked by central finite differences "
                         "(0: none)")
    ap.add_argument("--no-eigen", action="store_true")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the built-in launch parameters instead of autotuning them")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on exactly one GPU: the library has no multi-GPU path "
              "(a row-sharded design exists on paper only)", file=sys.stderr)
        return 2

    import torch

    import gt_hip
    from gt_hip import _lib

    if not torch.cuda.is_available() or gt_hip.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        return 1
    if args.workload == "batched":
        return bench_batched(args, torch, gt_hip, _lib)
    if args.workload == "adjoint":
        return bench_adjoint(args, torch, gt_hip, _lib)
    if args.workload == "uneven":
        return bench_uneven(args, torch, gt_hip, _lib)
    if args.workload == "betweenness":
        return bench_betweenness(args, torch, gt_hip, _lib)
    if args.scale is None:
        args.scale = 22

    n = 1 << args.scale
    m = args.edgefactor * n
    t0 = time.perf_counter()
    src, tgt = gt_hip.gen_rmat(args.scale, m, 0.57, 0.19, 0.19, SEED)
    log(f"[bench] R-MAT scale {args.scale}: n={n} m={m} generated in {time.perf_counter()-t0:.1f}s")
    t0 = time.perf_counter()
    adj = gt_hip.host_adjacency(n, src, tgt)
    t_adj = time.perf_counter() - t0
    t0 = time.perf_counter()
    gid = 1
    dev = gt_hip.DeviceGraph.upload(gid, n, m, *adj)
    t_up = time.perf_counter() - t0
    log(f"[bench] host adjacency {t_adj:.1f}s, upload {t_up:.2f}s")
    del adj

    stream = torch.cuda.current_stream()
    dev.set_stream(stream.cuda_stream)
    copy = copy_calibration(torch)
    log(f"[bench] copy calibration: {copy['bytes']/2**30:.0f} GiB in {copy['us']:.0f} us "
        f"= {copy['gbs']:.0f} GB/s")
    tuning = _tune(args, dev, "pagerank", None)
    rank = torch.empty(n, dtype=torch.float64, device="cuda")
    params = _lib.GtPowerIteration(args.epsilon, 0, _lib.GtPlan(n, -(-n // 4096)))
    view = _lib.GtView(_lib.GT_VIEW_DIRECTED, 0, None, 0, None, 0)
    unity = _lib.GtProp(_lib.GT_MAP_UNITY, 0, 0, 0, 1.0, None, 0, 0)
    out = _lib.GtVertexOut(C.c_void_p(rank.data_ptr()), n, 0, _lib.GT_DEVICE, 0, gid)
    iters = C.c_uint64()

    def solve(fn: str) -> int:
        out.had = 0  # cold start: every slot seeded with 1/n
        _lib.call(fn, dev.handle, C.byref(view), C.byref(unity), None, 0.85, C.byref(params),
                  C.byref(out), C.byref(iters))
        return int(iters.value)

    def timed(fn: str, kind: int):
        # Per-step times (each solve between two synchronisations); the
        # median step is the reported one, the mean is kept beside it.  The
        # kernel times come from the same solves run again with the
        # library's sweep events on (events around every sweep cost a solve
        # about 140 us, so the timed solves run without them).
        for _ in range(args.warmup):
            solve(fn)
        times, sweeps = [], []
        for _ in range(args.steps):
            torch.cuda.synchronize()
            t = time.perf_counter()
            sweeps.append(solve(fn))
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t)
        dev.set_profiling(True)
        for _ in range(args.steps):
            solve(fn)
        torch.cuda.synchronize()
        ms_k, launches, _ = dev.sweep_stats(kind)
        ms_other = {k: dev.sweep_stats(k)[:2] for k in range(4) if k != kind}
        dev.set_profiling(False)
        dt = float(np.median(times)) * args.steps
        return dt, sweeps, ms_k, launches, ms_other, timing_stats(times)

    # ---- FP32-value / FP64-residual form (the chosen form)
    dt32, sw32, ms32, la32, other32, st32 = timed("gt_pagerank_mixed", _lib.GT_SWEEP_PAGERANK_F32)
    x32 = rank.cpu().numpy().copy()
    lay32 = dev.layout()
    f64_in_mixed = other32[_lib.GT_SWEEP_PAGERANK_F64]
    log(f"[bench] mixed: {dt32/args.steps*1e3:.3f} ms/solve, sweeps {sw32[0]}, "
        f"f32 launches {la32} ({ms32/max(la32,1)*1e3:.1f} us avg), "
        f"f64 refinement launches {f64_in_mixed[1]}")

    # ---- plain FP64 drop-in
    dt64, sw64, ms64, la64, _, st64 = timed("gt_pagerank_f64", _lib.GT_SWEEP_PAGERANK_F64)
    x64 = rank.cpu().numpy().copy()
    log(f"[bench] fp64: {dt64/args.steps*1e3:.3f} ms/solve, sweeps {sw64[0]}, "
        f"launches {la64} ({ms64/max(la64,1)*1e3:.1f} us avg)")

    lay64 = dev.layout()
    # The same drop-in on the wave engine (tuning block_f64 = -1), beside it.
    dev.set_tuning(block_f64=-1)
    dt64b, sw64b, ms64b, la64b, _, st64b = timed("gt_pagerank_f64", _lib.GT_SWEEP_PAGERANK_F64)
    x64b = rank.cpu().numpy().copy()
    dev.set_tuning(block_f64=0)
    log(f"[bench] fp64 on the wave engine: {dt64b/args.steps*1e3:.3f} ms/solve, "
        f"sweeps {sw64b[0]}, launches {la64b} ({ms64b/max(la64b,1)*1e3:.1f} us avg)")

    # The same solve handing over a host rank map (copy in / out over PCIe
    # inside the call): reported beside `value`, never as it.
    rank_h = np.zeros(n, np.float64)
    out_h = _lib.GtVertexOut(rank_h.ctypes.data_as(C.c_void_p), n, 0, _lib.GT_HOST, 0, gid)

    def solve_host() -> None:
        out_h.had = 0
        _lib.call("gt_pagerank_mixed", dev.handle, C.byref(view), C.byref(unity), None, 0.85,
                  C.byref(params), C.byref(out_h), C.byref(iters))

    solve_host()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(3):
        solve_host()
    host_ms = (time.perf_counter() - t) / 3 * 1e3
    log(f"[bench] mixed with a host rank map: {host_ms:.3f} ms/solve")

    # Work per solve: every sweep of the mixed form is a full pass over m edges.
    gteps32 = m * sw32[0] * args.steps / dt32 / 1e9
    gteps64 = m * sw64[0] * args.steps / dt64 / 1e9
    b32, b64 = pr_bytes(n, m, 4), pr_bytes(n, m, 8)
    avg32 = ms32 / max(la32, 1) * 1e-3
    avg64 = ms64 / max(la64, 1) * 1e-3
    # Launches that found the loop finished move no bytes: spread the
    # algorithmic bytes of the sweeps done over every launch of that kernel.
    n_f32_sweeps = (sw32[0] - f64_in_mixed[1] // max(args.steps, 1)) * args.steps
    ach32 = b32 * n_f32_sweeps / max(la32, 1) / avg32 / 1e9 if la32 else 0.0
    ach64 = b64 * sw64[0] * args.steps / max(la64, 1) / avg64 / 1e9 if la64 else 0.0

    workload = f"pagerank_rmat{args.scale}_ef{args.edgefactor}"
    hot = dev.tuning()["sweep_hot_kb"] > 0
    # The FP32 sweeps ran on block tiles unless the tuning turned them off
    # (the timed solves' layout, read before the FP64 solves).
    blk = {k: lay32[k] for k in ("block_tiles", "block_groups", "block_long_rows")}
    on_blocks = blk["block_tiles"] > 0
    k32 = ("pr_sweep_blk<false, 0, false>" if on_blocks
           else "pr_sweep_hot<float, false, false>" if hot else "pr_sweep<float>")
    traffic32, traffic_src = profiled_traffic(workload, k32)
    result = {
        "metric": METRIC,
        "value": round(gteps32, 3),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt32 / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 values/SpMV, f64 dangling/teleport/L1 + f64 refinement",
        "data": "synthetic",
        "config": {
            "workload": workload,
            "vertices": n, "edges": m, "rmat": [0.57, 0.19, 0.19, 0.05], "seed": SEED,
            "damping": 0.85, "epsilon": args.epsilon, "weights": "unity",
            "personalisation": "uniform", "start": "cold (fresh map per step)",
            "fp32_phase_stop": "delta < max(epsilon, 2^-22 * sum|x|), then one FP64 sweep",
        },
        "roofline": {
            "bound": "hbm", "achieved": round(ach32, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach32 / HBM_PEAK_GBS, 4), "traffic": traffic32,
            "kernel": ("pr_sweep_blk (FP32 PageRank sweep on block tiles: source-sorted tiles, "
                       "64-bit fixed-point LDS row sums)" if on_blocks
                       else "pr_sweep_hot (FP32 PageRank sweep, hot contributions in LDS)" if hot
                       else "pr_sweep<float> (FP32 PageRank sweep)"),
            "bytes_per_sweep": b32, "avg_launch_us": round(avg32 * 1e6, 2),
            "traffic_source": traffic_src, "copy_calibration": copy,
            "layout": blk,
        },
        "tuning": tuning,
        "fp64_dropin": {
            "gteps": round(gteps64, 3), "ms_per_solve": round(dt64 / args.steps * 1e3, 4),
            "sweeps": sw64[0], "achieved_gbs": round(ach64, 1),
            "roofline_frac": round(ach64 / HBM_PEAK_GBS, 4), "bytes_per_sweep": b64,
            "avg_launch_us": round(avg64 * 1e6, 2),
            "kernel": ("pr_sweep_blk<false, 2, false> (block tiles, two-word fixed-point sums)"
                       if lay64["block_tiles"] else "pr_sweep_hot<double> (wave engine)"),
            "traffic": profiled_traffic(workload, "pr_sweep_blk<false, 2, false>"
                                        if lay64["block_tiles"]
                                        else "pr_sweep_hot<double, false, false>")[0],
        },
        "fp64_dropin_wave_engine": {
            "tuning": {"block_f64": -1},
            "gteps": round(m * sw64b[0] * args.steps / dt64b / 1e9, 3),
            "ms_per_solve": round(dt64b / args.steps * 1e3, 4), "sweeps": sw64b[0],
            "avg_launch_us": round(ms64b / max(la64b, 1) * 1e3, 2),
            "roofline_frac": round(b64 * sw64b[0] * args.steps / max(la64b, 1)
                                   / (ms64b / max(la64b, 1) * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
            if la64b else None,
        },
        "timing": {"mixed": st32, "fp64_dropin": st64, "fp64_dropin_wave_engine": st64b,
                   "stat": f"median of {args.steps} solves after {args.warmup} warm-ups, sweep "
                           f"events off; kernel averages from the same solves run again with "
                           f"events"},
        "mixed_vs_fp64_speedup": round(dt64 / dt32, 3),
        # The speedup in two factors: fewer sweeps (the FP32 phase stops at
        # its noise floor, then one FP64 sweep) and a faster sweep kernel.
        "mixed_vs_fp64_split": {
            "sweeps_fp64": sw64[0], "sweeps_mixed": sw32[0],
            "sweep_count_ratio": round(sw64[0] / max(sw32[0], 1), 4),
            "per_sweep_kernel_ratio": round(avg64 / avg32, 4) if avg32 else None,
            "rest": round((dt64 / dt32) / (sw64[0] / max(sw32[0], 1)) / (avg64 / avg32), 4)
            if avg32 else None},
        "host_rank_map_ms_per_solve": round(host_ms, 4),
        "mixed_sweeps": {"total": sw32[0], "fp64_refinement": f64_in_mixed[1] // max(args.steps, 1)},
        "host_adjacency_s": round(t_adj, 2), "upload_s": round(t_up, 3),
    }

    # ---- the gather ceiling: the time the sweeps' own memory gathers take
    # at the device's random-gather rate (every gather an L2 hit, nothing
    # else costing anything), beside the measured sweep times.
    # A second floor adds the L2 misses the committed counter profile
    # measured for the kernel, each at the probe's extra cost of a miss (the
    # 16 MB window's rate, where nearly every gather misses, less the
    # L2-resident rate).
    rate = gather_rate()
    ns_l2 = rate["l2_resident_1mb"]["ns_per_lane_per_cu"]
    ns_miss = rate["uniform_16mb"]["ns_per_lane_per_cu"] - ns_l2
    cus = rate["compute_units"]
    kb = dev.tuning()["sweep_hot_kb"]
    ceiling = {"probe": rate, "miss_extra_ns_per_cu": round(ns_miss, 4)}

    def floors(key, cold, nh, avg, pmc_kernel):
        misses = profiled_misses(workload, pmc_kernel)
        floor_us = cold * ns_l2 / cus / 1e3
        rec = {"hot_slice": nh, "gathers_from_memory": cold, "floor_us": round(floor_us, 1),
               "sweep_us": round(avg * 1e6, 1),
               "floor_over_sweep": round(floor_us / (avg * 1e6), 4) if avg else None}
        if misses is not None:
            fm = (cold * ns_l2 + misses * ns_miss) / cus / 1e3
            rec.update({"l2_misses_profiled": int(misses), "floor_with_misses_us": round(fm, 1),
                        "floor_with_misses_over_sweep": round(fm / (avg * 1e6), 4) if avg else None})
        ceiling[key] = rec

    # The floors model the wave engine's sweeps (hot slice in LDS, one
    # request per gathered lane outside it): the FP64 and eigenvector sweeps,
    # and the FP32 sweep when it ran there.
    avg64w = ms64b / max(la64b, 1) * 1e-3
    sweeps = [("fp64_sweep", 8, avg64w, "pr_sweep_hot<double, false, false>")]
    if not on_blocks:
        sweeps.insert(0, ("fp32_sweep", 4, avg32, "pr_sweep_hot<float, false, false>"))
    for key, vbytes, avg, pk in sweeps:
        nh = hot_len(kb, vbytes, n)
        floors(key, cold_gathers(src, n, nh), nh, avg, pk)
    result["gather_ceiling"] = ceiling
    log(f"[bench] gather ceiling: {ns_l2:.3f} ns per L2-hit lane per CU; FP64 sweep floor "
        f"{ceiling['fp64_sweep']['floor_us']:.0f} us "
        f"({ceiling['fp64_sweep'].get('floor_with_misses_us', float('nan')):.0f} us with its "
        f"L2 misses) vs {avg64w*1e6:.0f} us (wave engine)")

    # ---- eigenvector on the symmetrised graph
    if not args.no_eigen:
        S = np.concatenate([src, tgt])
        T = np.concatenate([tgt, src])
        ms_ = int(S.shape[0])
        adj_s = gt_hip.host_adjacency(n, S, T)
        devs = gt_hip.DeviceGraph.upload(2, n, ms_, *adj_s)
        del adj_s
        devs.set_stream(stream.cuda_stream)
        xs = torch.empty(n, dtype=torch.float64, device="cuda")
        outs = _lib.GtVertexOut(C.c_void_p(xs.data_ptr()), n, 0, _lib.GT_DEVICE, 0, 2)
        eig = C.c_double()

        def esolve() -> int:
            outs.had = 0
            _lib.call("gt_eigenvector_f64", devs.handle, C.byref(view), C.byref(unity),
                      C.byref(params), C.byref(outs), C.byref(eig), C.byref(iters))
            return int(iters.value)

        for _ in range(args.warmup):
            esolve()
        torch.cuda.synchronize()
        t = time.perf_counter()
        swe = [esolve() for _ in range(args.steps)]
        torch.cuda.synchronize()
        dte = time.perf_counter() - t
        devs.set_profiling(True)
        for _ in range(args.steps):
            esolve()
        torch.cuda.synchronize()
        mse, lae, _ = devs.sweep_stats(_lib.GT_SWEEP_EIGENVECTOR)
        devs.set_profiling(False)
        lay_e = devs.layout()
        ev_blocks = lay_e["block_tiles"] > 0
        ke = "pr_sweep_blk<false, 3, false>" if ev_blocks else "ev_sweep_hot"
        be = ev_bytes(n, ms_)
        avge = mse / max(lae, 1) * 1e-3
        ache = be * lae / max(lae, 1) / avge / 1e9 if lae else 0.0
        result["eigenvector"] = {
            "gteps": round(ms_ * swe[0] * args.steps / dte / 1e9, 3),
            "ms_per_solve": round(dte / args.steps * 1e3, 4), "sweeps": swe[0],
            "launches_per_solve": lae // max(args.steps, 1),
            "eigenvalue": eig.value, "achieved_gbs": round(ache, 1),
            "roofline_frac": round(ache / HBM_PEAK_GBS, 4), "bytes_per_sweep": be,
            "kernel": ("pr_sweep_blk<false, 3, false> (eigenvector sweep on block tiles, two-word "
                       "fixed-point row sums)" if ev_blocks
                       else "ev_sweep_hot (eigenvector sweep, hot iterate in LDS)"),
            "layout": {k: lay_e[k] for k in ("block_tiles", "block_groups", "block_long_rows")},
            "traffic": profiled_traffic(workload, ke)[0],
        }
        x_eig = xs.cpu().numpy().copy()
        if not ev_blocks:
            nh_e = hot_len(kb, 8, n)
            deg_s = np.bincount(S, minlength=n)
            cold_e = int(deg_s.sum() - np.sort(deg_s)[::-1][:nh_e].sum())
            floors("eigenvector_sweep", cold_e, nh_e, avge, "ev_sweep_hot<false>")
        log(f"[bench] eigenvector: {dte/args.steps*1e3:.3f} ms/solve, sweeps {swe[0]}, "
            f"lambda {eig.value:.6f}")
        devs.close()

    # ---- CPU baseline and FP64 oracle answers
    if not args.no_cpu:
        import oracle

        threads = cpu_share()
        hinfo = host_info()
        t0 = time.perf_counter()
        o = oracle.Adj(n, src, tgt)
        log(f"[bench] oracle adjacency built in {time.perf_counter()-t0:.1f}s")
        # Five whole cold-start solves (graph already built), the median one.
        solve_s = []
        for _ in range(5):
            t0 = time.perf_counter()
            want, it_o = o.pagerank(epsilon=args.epsilon, max_iter=0, grain=4096,
                                    threads=threads)
            solve_s.append(time.perf_counter() - t0)
        dtc = float(np.median(solve_s))
        # The same path on one host thread: a solve capped at 6 sweeps (its
        # setup is spread over fewer sweeps, so this rate is slightly low).
        t0 = time.perf_counter()
        _, it1 = o.pagerank(epsilon=args.epsilon, max_iter=6, grain=4096, threads=1)
        dt1 = time.perf_counter() - t0
        del o
        cpu_gteps = m * it_o / dtc / 1e9
        cpu1 = m * it1 / dt1 / 1e9
        log(f"[bench] oracle pagerank: median of 5 solves {dtc:.2f}s, {it_o} sweeps each, "
            f"{cpu_gteps:.3f} GTEPS on {threads} threads; {cpu1:.3f} GTEPS on 1 thread; "
            f"host {hinfo}")
        nz = want != 0

        def errs(x):
            return {"l1": float(np.abs(x - want).sum()),
                    "max_rel": float((np.abs(x - want)[nz] / np.abs(want[nz])).max())}

        result["cpu_baseline"] = {
            "value": round(cpu_gteps, 4), "unit": "GTEPS", "cores": threads, "kind": "port",
            "stat": "median of 5",
            "sample": f"5 cold-start pagerank_with solves (oracle C restatement of the "
                      f"reference's rayon path, Plan grain 4096, {it_o} sweeps each) on the same "
                      f"R-MAT-{args.scale} graph, graph built beforehand; median {dtc:.2f}s "
                      f"(all: {', '.join(f'{x:.2f}' for x in solve_s)} s)",
            "host": hinfo,
            "one_thread": {"value": round(cpu1, 4), "unit": "GTEPS", "cores": 1,
                           "sample": f"one solve capped at {it1} sweeps, {dt1:.1f}s"},
        }
        result["accuracy_vs_fp64_oracle"] = {
            "mixed": errs(x32), "fp64_dropin": errs(x64), "fp64_dropin_wave_engine": errs(x64b),
            "oracle_sweeps": it_o, "fp64_sweeps": sw64[0],
        }
        if not args.no_eigen:
            o2 = oracle.Adj(n, S, T)
            t0 = time.perf_counter()
            xo, eo, ite = o2.eigenvector(epsilon=args.epsilon, max_iter=0, grain=4096,
                                         threads=threads)
            dte_c = time.perf_counter() - t0
            del o2
            result["eigenvector"]["vs_oracle"] = {
                "l1": float(np.abs(x_eig - xo).sum()), "oracle_sweeps": ite,
                "eigenvalue_rel_err": abs(result["eigenvector"]["eigenvalue"] - eo) / abs(eo),
                "cpu_gteps": round(ms_ * ite / dte_c / 1e9, 4),
            }
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
