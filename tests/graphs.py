"""Fixture graphs of the reference's own tests, regenerated bit for bit.

thousand()  -- crates/gt-algo/tests/u16_centrality.rs:93-107
sixty()     -- crates/gt-algo/tests/u16_centrality.rs:495-503 (eigenvector residual graph)
bench()     -- crates/gt-algo/benches/kernels.rs:49-73 (ring + SplitMix64 chords)
LAYOUT      -- crates/gt-core/tests/u34_diff_adjacency.rs:74-83
brandes_cases() -- crates/gt-algo/tests/u16_centrality.rs:707-770 (100 small
                   multigraphs with self-loops and parallel edges)
"""
import numpy as np

import oracle


class Rng:
    """SplitMix64 with the reference's below(n) = next() % n."""

    def __init__(self, seed):
        self.s = seed & (2**64 - 1)

    def next(self):
        M = 2**64 - 1
        self.s = (self.s + 0x9E3779B97F4A7C15) & M
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        return z ^ (z >> 31)

    def below(self, n):
        return self.next() % n


def thousand():
    N = 1000
    r = Rng(0x243F6A8885A308D3)
    edges = [(i, (i + 1) % N) for i in range(N)]
    for _ in range(4000):
        s = r.below(N)
        t = r.below(N)
        edges.append((s, t))

    def sink(i):
        return i >= 17 and (i - 17) % 20 == 0 and i < 17 + 20 * 50

    edges = [(s, t) for (s, t) in edges if not sink(s)]
    src = np.array([e[0] for e in edges], np.uint32)
    tgt = np.array([e[1] for e in edges], np.uint32)
    return N, src, tgt


def sixty():
    n = 60
    r = Rng(0xFEEDFACECAFEBEEF)
    edges = [(i, (i + 1) % n) for i in range(n)]
    for _ in range(240):
        s = r.below(n)
        t = r.below(n)
        edges.append((s, t))
    src = np.array([e[0] for e in edges], np.uint32)
    tgt = np.array([e[1] for e in edges], np.uint32)
    w = np.array([1.0 + (i % 5) * 0.25 for i in range(len(edges))])
    return n, src, tgt, w


def bench(n=100_000, m=500_000):
    """Ring first, then m - n SplitMix64 edges (source draw before target)."""
    ring_s = np.arange(n, dtype=np.uint32)
    ring_t = ((np.arange(n) + 1) % n).astype(np.uint32)
    s, t = oracle.gen_er(n, m - n, 0x243F6A8885A308D3)
    return n, np.concatenate([ring_s, s]), np.concatenate([ring_t, t])


EDGES = [(0, 1), (1, 2), (2, 0), (0, 0), (0, 1)]
LAYOUT = [
    ([(1, 0), (0, 3), (1, 4), (0, 3), (2, 2)], 3),
    ([(2, 1), (0, 0), (0, 4)], 1),
    ([(0, 2), (1, 1)], 1),
    ([], 0),
]


def c1():
    """Plain PageRank workload: directed G(n=10,000, m=100,000), SplitMix64
    seed 0x243f6a8885a308d3, source then target per edge."""
    s, t = oracle.gen_er(10_000, 100_000, 0x243F6A8885A308D3)
    return 10_000, s, t


def brandes_cases():
    """The 100 (n, edges) of u16_centrality.rs:707-770, drawn in its order:
    seed 0x0123456789abcdef, n = 3 + below(10), m = below(3n + 1), each edge
    (below(n), below(n))."""
    r = Rng(0x0123456789ABCDEF)
    cases = []
    for _ in range(100):
        n = 3 + r.below(10)
        m = r.below(3 * n + 1)
        edges = []
        for _ in range(m):
            s = r.below(n)
            t = r.below(n)
            edges.append((s, t))
        cases.append((n, edges))
    return cases
