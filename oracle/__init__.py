"""The oracle: a CPU restatement of graph-tool-rs' PageRank / eigenvector path.

TEST INFRASTRUCTURE ONLY.  It is the checker the GPU library is compared
against (tests/, __graft_entry__.smoke) and the CPU baseline bench.py times;
the product never imports it.  The arithmetic lives in gt_oracle.c (each
function there cites the reference file:line it restates); this module is a
ctypes wrapper plus the numpy references the reference's own tests use
(dense_pagerank: crates/gt-algo/tests/u16_centrality.rs:138-175).
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path
from typing import Optional, Tuple

import numpy as np

HERE = Path(__file__).resolve().parent
_SO = HERE / "libgt_oracle.so"

_u32p = C.POINTER(C.c_uint32)
_u8p = C.POINTER(C.c_uint8)
_f64p = C.POINTER(C.c_double)

UNITY, CONSTANT, DENSE = 0, 1, 2

_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return _SO


def lib():
    global _lib
    if _lib is None:
        if not _SO.exists():
            build()
        L = C.CDLL(str(_SO))
        L.oracle_gen_er.argtypes = [C.c_uint32, C.c_uint64, C.c_uint64, _u32p, _u32p]
        L.oracle_gen_rmat.argtypes = [C.c_uint32, C.c_uint64, C.c_double, C.c_double,
                                      C.c_double, C.c_uint64, _u32p, _u32p]
        L.oracle_adj_build.argtypes = [C.c_uint32, C.c_uint64, _u32p, _u32p]
        L.oracle_adj_build.restype = C.c_void_p
        L.oracle_adj_free.argtypes = [C.c_void_p]
        L.oracle_adj_export.argtypes = [C.c_void_p, _u32p, _u32p, _u32p, _u32p, _u32p, _u32p]
        L.oracle_adj_block.argtypes = [C.c_void_p, C.c_uint32, _u32p, _u32p, _u32p]
        L.oracle_adj_block.restype = C.c_uint32
        L.oracle_pagerank.argtypes = [C.c_void_p, C.c_int, _u8p, _u8p, C.c_int, C.c_double,
                                      _f64p, C.c_int, C.c_double, _f64p, C.c_double,
                                      C.c_double, C.c_uint64, C.c_uint64, C.c_uint64, _f64p,
                                      C.c_uint64, C.c_int]
        L.oracle_pagerank.restype = C.c_int64
        L.oracle_eigenvector.argtypes = [C.c_void_p, C.c_int, _u8p, _u8p, C.c_int, C.c_double,
                                         _f64p, C.c_double, C.c_uint64, C.c_uint64, C.c_uint64,
                                         _f64p, C.c_uint64, C.c_int, _f64p]
        L.oracle_eigenvector.restype = C.c_int64
        L.oracle_pagerank_grad.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, _f64p,
                                           C.c_int, C.c_double, _f64p, C.c_double, _f64p, _f64p,
                                           C.c_double, C.c_uint64, _f64p, _f64p, _f64p, _f64p]
        L.oracle_pagerank_grad.restype = C.c_int64
        L.oracle_betweenness.argtypes = [C.c_void_p, C.c_int, C.c_int, _u8p, _u8p, _u32p,
                                         C.c_uint64, C.c_int, _f64p, _f64p]
        L.oracle_betweenness.restype = C.c_int
        _lib = L
    return _lib


def _p(a, t=_u32p):
    return None if a is None else a.ctypes.data_as(t)


def splitmix_stream(seed: int, count: int) -> np.ndarray:
    """``count`` successive outputs of SplitMix64 seeded with ``seed``
    (u16_centrality.rs:58-72), in pure Python integer arithmetic."""
    out = np.zeros(count, dtype=np.uint64)
    s = seed & (2**64 - 1)
    M = 2**64 - 1
    for i in range(count):
        s = (s + 0x9E3779B97F4A7C15) & M
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        out[i] = z ^ (z >> 31)
    return out


def gen_er(n: int, m: int, seed: int) -> Tuple[np.ndarray, np.ndarray]:
    src = np.zeros(max(m, 1), np.uint32)
    tgt = np.zeros(max(m, 1), np.uint32)
    lib().oracle_gen_er(n, m, seed & (2**64 - 1), _p(src), _p(tgt))
    return src[:m], tgt[:m]


def gen_rmat(scale: int, m: int, a: float, b: float, c: float, seed: int):
    src = np.zeros(max(m, 1), np.uint32)
    tgt = np.zeros(max(m, 1), np.uint32)
    lib().oracle_gen_rmat(scale, m, a, b, c, seed & (2**64 - 1), _p(src), _p(tgt))
    return src[:m], tgt[:m]


def _map(m):
    """(kind, value, array) of a weight / vertex map given as None (unity),
    a float (constant) or an array (dense, float64)."""
    if m is None:
        return UNITY, 1.0, None
    if np.isscalar(m):
        return CONSTANT, float(m), None
    a = np.ascontiguousarray(m, dtype=np.float64)
    return DENSE, 0.0, a


class Adj:
    """The reference's AdjList after add_edge(src[i], tgt[i]) for i = 0..m."""

    def __init__(self, n: int, src, tgt):
        self.n = int(n)
        self.src = np.ascontiguousarray(src, dtype=np.uint32)
        self.tgt = np.ascontiguousarray(tgt, dtype=np.uint32)
        self.m = int(self.src.shape[0])
        h = lib().oracle_adj_build(self.n, self.m, _p(self.src), _p(self.tgt))
        if not h:
            raise ValueError("no such vertex")
        self._h = C.c_void_p(h)

    def __del__(self):
        try:
            if self._h:
                lib().oracle_adj_free(self._h)
        except Exception:
            pass

    def export(self):
        """(in_off, in_other, in_eid, out_off, out_other, out_eid) in storage order."""
        n, m = self.n, self.m
        a = [np.zeros(n + 1, np.uint32), np.zeros(max(m, 1), np.uint32),
             np.zeros(max(m, 1), np.uint32), np.zeros(n + 1, np.uint32),
             np.zeros(max(m, 1), np.uint32), np.zeros(max(m, 1), np.uint32)]
        lib().oracle_adj_export(self._h, *[_p(x) for x in a])
        return a[0], a[1][:m], a[2][:m], a[3], a[4][:m], a[5][:m]

    def block(self, v: int):
        """(entries [(other, idx)...], out_len) of vertex v's block."""
        cap = 2 * self.m + 1
        o = np.zeros(cap, np.uint32)
        i = np.zeros(cap, np.uint32)
        ol = np.zeros(1, np.uint32)
        k = lib().oracle_adj_block(self._h, v, _p(o), _p(i), _p(ol))
        return list(zip(o[:k].tolist(), i[:k].tolist())), int(ol[0])

    def pagerank(self, weight=None, pers=None, damping=0.85, epsilon=1e-6, max_iter=0,
                 grain=64, rank: Optional[np.ndarray] = None, had: Optional[int] = None,
                 reversed: bool = False, vmask=None, emask=None, threads: int = 1,
                 plan_len: Optional[int] = None, plan_chunks: Optional[int] = None):
        """pagerank_with on this graph; pers None = uniform 1/num_vertices.
        Returns (rank, sweeps); raises ValueError on a plan mismatch."""
        n = self.n
        wk, wv, wa = _map(weight)
        vm = None if vmask is None else np.ascontiguousarray(vmask, dtype=np.uint8)
        em = None if emask is None else np.ascontiguousarray(emask, dtype=np.uint8)
        if pers is None:
            nv = n if vm is None else int(np.count_nonzero(vm[:n]))
            pk, pv, pa = CONSTANT, (0.0 if nv == 0 else 1.0 / nv), None
        else:
            pk, pv, pa = _map(pers)
        r = np.zeros(n, np.float64)
        if rank is not None:
            k = min(len(rank), n)
            r[:k] = rank[:k]
            h = len(rank) if had is None else had
        else:
            h = 0
        pl = n if plan_len is None else plan_len
        pc = (max(1, -(-pl // grain))) if plan_chunks is None else plan_chunks
        it = lib().oracle_pagerank(self._h, int(reversed), _p(vm, _u8p), _p(em, _u8p), wk, wv,
                                   _p(wa, _f64p), pk, pv, _p(pa, _f64p), damping, epsilon,
                                   max_iter, pl, pc, _p(r, _f64p), h, threads)
        if it < 0:
            raise ValueError("sweep: the plan must cover the output exactly")
        return r, int(it)

    def pagerank_grad(self, x, cvec, weight=None, pers=None, damping=0.85, epsilon=1e-14,
                      max_iter=0, reversed: bool = False):
        """Gradient of L = cvec . x through the PageRank fixed point x.
        Returns (grad_w [edge_bound], grad_pers [n], grad_damping, lambda, sweeps)."""
        n = self.n
        wk, wv, wa = _map(weight)
        if pers is None:
            pk, pv, pa = CONSTANT, (0.0 if n == 0 else 1.0 / n), None
        else:
            pk, pv, pa = _map(pers)
        x = np.ascontiguousarray(x, dtype=np.float64)
        c = np.ascontiguousarray(cvec, dtype=np.float64)
        lam = np.zeros(max(n, 1))
        gw = np.zeros(max(self.m, 1))
        gp = np.zeros(max(n, 1))
        gd = np.zeros(1)
        it = lib().oracle_pagerank_grad(self._h, int(reversed), wk, wv, _p(wa, _f64p), pk, pv,
                                        _p(pa, _f64p), damping, _p(x, _f64p), _p(c, _f64p),
                                        epsilon, max_iter, _p(lam, _f64p), _p(gw, _f64p),
                                        _p(gp, _f64p), _p(gd, _f64p))
        return gw[: self.m], gp[:n], float(gd[0]), lam[:n], int(it)

    def eigenvector(self, weight=None, epsilon=1e-6, max_iter=0, grain=64,
                    x: Optional[np.ndarray] = None, reversed: bool = False, vmask=None,
                    emask=None, threads: int = 1):
        """Returns (x, eigenvalue, sweeps)."""
        n = self.n
        wk, wv, wa = _map(weight)
        vm = None if vmask is None else np.ascontiguousarray(vmask, dtype=np.uint8)
        em = None if emask is None else np.ascontiguousarray(emask, dtype=np.uint8)
        r = np.zeros(n, np.float64)
        h = 0
        if x is not None:
            k = min(len(x), n)
            r[:k] = x[:k]
            h = len(x)
        eig = np.zeros(1, np.float64)
        pc = max(1, -(-n // grain))
        it = lib().oracle_eigenvector(self._h, int(reversed), _p(vm, _u8p), _p(em, _u8p), wk,
                                      wv, _p(wa, _f64p), epsilon, max_iter, n, pc,
                                      _p(r, _f64p), h, threads, _p(eig, _f64p))
        if it < 0:
            raise ValueError("sweep: the plan must cover the output exactly")
        return r, float(eig[0]), int(it)


def dense_pagerank(n: int, edges, pers, d: float, sweeps: int) -> np.ndarray:
    """The dense reference of u16_centrality.rs:138-175: an explicit n x n
    transition matrix, a fixed sweep count, zero-out-weight columns dropped."""
    deg = np.zeros(n)
    for s, _, w in edges:
        deg[s] += w
    M = np.zeros((n, n))
    for s, t, w in edges:
        if deg[s] != 0.0:
            M[t, s] += w / deg[s]
    pers = np.asarray(pers, dtype=np.float64)
    r = pers.copy()
    sink = deg == 0.0
    for _ in range(sweeps):
        p_sink = r[sink].sum()
        r = (1.0 - d) * pers + d * (p_sink * pers + M @ r)
    return r


def betweenness(adj: "Adj", reversed: bool = False, undirected: bool = False, vmask=None,
                emask=None, pivots=None, threads: int = 1):
    """betweenness (centrality.rs:630-689) on a view of ``adj``: returns
    (vertex_bc [n], edge_bc [edge_bound]).  ``pivots`` None means every vertex
    of the view (the reference's behaviour); a list restricts the sources."""
    L = lib()
    vm = None if vmask is None else np.ascontiguousarray(vmask, dtype=np.uint8)
    em = None if emask is None else np.ascontiguousarray(emask, dtype=np.uint8)
    pv = None if pivots is None else np.ascontiguousarray(pivots, dtype=np.uint32)
    vb = np.zeros(max(adj.n, 1))
    eb = np.zeros(max(adj.m, 1))
    rc = L.oracle_betweenness(adj._h, int(reversed), int(undirected), _p(vm, _u8p), _p(em, _u8p),
                              _p(pv), 0 if pv is None else len(pv), threads, _p(vb, _f64p),
                              _p(eb, _f64p))
    if rc != 0:
        raise ValueError("pivot out of range")
    return vb[: adj.n], eb[: adj.m]
