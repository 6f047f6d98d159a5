"""Compile libgt_hip.so for gfx950 with hipcc, in-tree.

Every HIP translation unit is built with packed-FP32 instruction selection
disabled (``-Xclang -target-feature -Xclang -packed-fp32-ops``); host-only C++
goes through the same compiler driver without offload.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent  # graph-tool-rs_amd/
CSRC = ROOT / "csrc"
INCLUDE = ROOT.parent / "include"
BUILD = ROOT / "_build"
LIB = PKG / "libgt_hip.so"

ARCH = os.environ.get("GT_HIP_ARCH", "gfx950")

HIP_SOURCES = ["graph.hip", "props.hip", "pagerank.hip", "eigenvector.hip", "builder.hip",
               "views.hip", "batched.hip", "adjoint.hip", "relabel.hip", "uneven.hip", "degree.hip",
               "betweenness.hip", "sliced.hip", "blocks.hip"]
CXX_SOURCES = ["host.cpp", "io.cpp"]
PROBE_SOURCE = "gather_probe.hip"
PROBE_LIB = PKG / "libgt_probe.so"

COMMON = ["-O3", "-std=c++17", "-fPIC", f"-I{INCLUDE}", f"-I{CSRC}", "-Wall",
          "-Wno-unused-result"]
HIP_FLAGS = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-munsafe-fp-atomics"]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ... (exit {r.returncode})")


def build(verbose: bool = False, jobs: int = 8) -> Path:
    """Compile every source and link libgt_hip.so; returns its path."""
    hipcc = _hipcc()
    BUILD.mkdir(exist_ok=True)
    cmds = []
    objs = []
    for src in HIP_SOURCES:
        obj = BUILD / (src + ".o")
        objs.append(obj)
        cmds.append([hipcc, *COMMON, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
    for src in CXX_SOURCES:
        obj = BUILD / (src + ".o")
        objs.append(obj)
        cmds.append([hipcc, *COMMON, "-x", "c++", "-c", str(CSRC / src), "-o", str(obj)])
    procs = []
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                                            text=True)))
        if len(procs) >= jobs:
            _drain(procs.pop(0))
    while procs:
        _drain(procs.pop(0))
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", str(tmp),
          *map(str, objs), "-lpthread"])
    os.replace(tmp, LIB)
    # The measurement library (the random-gather rate behind the roofline).
    tmp = PROBE_LIB.with_suffix(".so.tmp")
    _run([hipcc, *COMMON, *HIP_FLAGS, "-shared", str(CSRC / PROBE_SOURCE), "-o", str(tmp)])
    os.replace(tmp, PROBE_LIB)
    return LIB


def _drain(item) -> None:
    cmd, p = item
    out, err = p.communicate()
    if p.returncode != 0:
        sys.stderr.write(out + err)
        raise RuntimeError(f"compile failed: {cmd[-3]}")


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
