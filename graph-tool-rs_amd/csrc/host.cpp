// Host-side pieces of the library: seeded workload generators and the host
// adjacency an edge list produces under AdjList::add_edge.
//
// Generators are counter-based SplitMix64 (crates/gt-algo/tests/u16_centrality.rs:58-72):
// draw k of the stream seeded with `seed` is mix(seed + (k+1) * 0x9e3779b97f4a7c15),
// exactly what k+1 calls of the sequential `next()` return, so a generator can
// run on many threads and still produce the sequential stream.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "gt_hip.h"

namespace gt {
gt_status fail(gt_status code, const std::string &msg);
void clear_error();
}  // namespace gt

namespace {

constexpr uint64_t kGamma = 0x9e3779b97f4a7c15ull;

inline uint64_t mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
// Draw k (0-based) of the stream.
inline uint64_t draw(uint64_t seed, uint64_t k) { return mix(seed + (k + 1) * kGamma); }
inline double unit(uint64_t seed, uint64_t k) { return (double)(draw(seed, k) >> 11) * 0x1.0p-53; }

template <typename F>
void parallel_for(uint64_t n, F f) {
    unsigned t = std::thread::hardware_concurrency();
    if (t == 0) t = 1;
    if (t > 16) t = 16;
    if (n < (1u << 16)) t = 1;
    std::vector<std::thread> th;
    for (unsigned k = 0; k < t; ++k) {
        uint64_t lo = n * k / t, hi = n * (k + 1) / t;
        th.emplace_back([=] {
            for (uint64_t i = lo; i < hi; ++i) f(i);
        });
    }
    for (auto &x : th) x.join();
}

}  // namespace

extern "C" {

// A batch of directed Chung-Lu graphs (the uneven-batch workload; no
// counterpart in the reference).  Graph g draws from the stream seeded with
// mix(seed + g):
//   draw 0      n_g = round(exp(U[ln n_min, ln n_max]))
//   draw 1      damping_g = d_lo + (d_hi - d_lo) U
//   per edge i  draws 2 + 3i, 3 + 3i, 4 + 3i: source, target, weight
// with m_g = round(mean_degree * n_g).  Vertex v of graph g has weight
// theta_v = (v + 1)^(-1 / (exponent - 1)) (expected degrees of a power law with
// that exponent); the source and the target of an edge are drawn
// independently with probability proportional to theta (inverse CDF over the
// prefix sums), so the mean out-degree is mean_degree.  Weights are
// 1 - U[0,1), in (0, 1].  Vertices and edges of graph g follow those of
// graphs 0..g-1 (ids are global).
//
// Two calls: with src == NULL only n_g, m_g and damping_g are written; then,
// with arrays of sum(m_g) entries, the edges.
gt_status gt_gen_chung_lu_batch(uint32_t num_graphs, uint64_t seed, uint32_t n_min,
                                uint32_t n_max, double mean_degree, double exponent,
                                double d_lo, double d_hi, uint32_t *n_g, uint64_t *m_g,
                                double *damping, uint32_t *src, uint32_t *tgt, float *weight) {
    gt::clear_error();
    if (!n_g || !m_g || !damping) return gt::fail(GT_ERR_INVALID_ARG, "null size output");
    if (n_min == 0 || n_max < n_min) return gt::fail(GT_ERR_INVALID_ARG, "bad vertex range");
    if (!(exponent > 1.0) || !(mean_degree >= 0.0))
        return gt::fail(GT_ERR_INVALID_ARG, "exponent must exceed 1, mean degree be >= 0");
    uint64_t tn = 0;
    for (uint32_t g = 0; g < num_graphs; ++g) {
        const uint64_t gs = mix(seed + g);
        const double l0 = std::log((double)n_min), l1 = std::log((double)n_max);
        double nv = std::round(std::exp(l0 + (l1 - l0) * unit(gs, 0)));
        nv = std::min<double>(std::max<double>(nv, n_min), n_max);
        n_g[g] = (uint32_t)nv;
        m_g[g] = (uint64_t)std::llround(mean_degree * nv);
        damping[g] = d_lo + (d_hi - d_lo) * unit(gs, 1);
        tn += n_g[g];
        if (tn > 0xffffffffull) return gt::fail(GT_ERR_INVALID_ARG, "batch exceeds 2^32 vertices");
    }
    if (!src) return GT_OK;
    if (!tgt || !weight) return gt::fail(GT_ERR_INVALID_ARG, "null edge output");
    std::vector<uint64_t> eoff(num_graphs + 1, 0);
    std::vector<uint32_t> voff(num_graphs + 1, 0);
    for (uint32_t g = 0; g < num_graphs; ++g) {
        eoff[g + 1] = eoff[g] + m_g[g];
        voff[g + 1] = voff[g] + n_g[g];
    }
    const double beta = 1.0 / (exponent - 1.0);
    unsigned t = std::thread::hardware_concurrency();
    if (t == 0) t = 1;
    if (t > 16) t = 16;
    std::vector<std::thread> th;
    for (unsigned k = 0; k < t; ++k) {
        th.emplace_back([&, k] {
            std::vector<double> cdf;
            for (uint32_t g = k; g < num_graphs; g += t) {
                const uint64_t gs = mix(seed + g);
                const uint32_t n = n_g[g];
                cdf.resize(n);
                double acc = 0.0;
                for (uint32_t v = 0; v < n; ++v) {
                    acc += std::pow((double)(v + 1), -beta);
                    cdf[v] = acc;
                }
                auto pick = [&](double u) {
                    const double x = u * acc;
                    uint32_t v = (uint32_t)(std::upper_bound(cdf.begin(), cdf.end(), x) - cdf.begin());
                    return v < n ? v : n - 1;
                };
                for (uint64_t i = 0; i < m_g[g]; ++i) {
                    const uint64_t e = eoff[g] + i;
                    src[e] = voff[g] + pick(unit(gs, 2 + 3 * i));
                    tgt[e] = voff[g] + pick(unit(gs, 3 + 3 * i));
                    weight[e] = (float)(1.0 - unit(gs, 4 + 3 * i));
                }
            }
        });
    }
    for (auto &x : th) x.join();
    return GT_OK;
}

// Directed G(n, m) as crates/gt-algo/benches/kernels.rs:49-73 draws it: per
// edge, source = next % n then target = next % n.
gt_status gt_gen_er(uint32_t n, uint64_t m, uint64_t seed, uint32_t *src, uint32_t *tgt) {
    gt::clear_error();
    if (n == 0 && m) return gt::fail(GT_ERR_INVALID_ARG, "no vertices to draw from");
    if (m && (!src || !tgt)) return gt::fail(GT_ERR_INVALID_ARG, "null output");
    parallel_for(m, [&](uint64_t i) {
        src[i] = (uint32_t)(draw(seed, 2 * i) % n);
        tgt[i] = (uint32_t)(draw(seed, 2 * i + 1) % n);
    });
    return GT_OK;
}

// R-MAT: edge i uses draws i*scale .. i*scale+scale-1, one 53-bit uniform per
// bit level, most significant level first; quadrant thresholds a, a+b,
// a+b+c over (source bit, target bit) = (0,0), (0,1), (1,0), (1,1).
gt_status gt_gen_rmat(uint32_t scale, uint64_t m, double a, double b, double c, uint64_t seed,
                      uint32_t *src, uint32_t *tgt) {
    gt::clear_error();
    if (scale == 0 || scale > 31) return gt::fail(GT_ERR_INVALID_ARG, "scale must be in 1..31");
    if (m && (!src || !tgt)) return gt::fail(GT_ERR_INVALID_ARG, "null output");
    const double ab = a + b, abc = a + b + c;
    parallel_for(m, [&](uint64_t i) {
        uint32_t u = 0, v = 0;
        for (uint32_t l = 0; l < scale; ++l) {
            const double r = unit(seed, i * scale + l);
            const uint32_t bit = 1u << (scale - 1 - l);
            if (r < a) {
            } else if (r < ab) {
                v |= bit;
            } else if (r < abc) {
                u |= bit;
            } else {
                u |= bit;
                v |= bit;
            }
        }
        src[i] = u;
        tgt[i] = v;
    });
    return GT_OK;
}

// The adjacency add_edge(src[i], tgt[i]), i = 0..m, builds
// (crates/gt-core/src/adj/list.rs:258-280, 471-500; adj/block.rs:171-209),
// flattened: in-halves and out-halves in storage order.  Each vertex's block
// is simulated in place: an out-insert while the in-half is non-empty moves the
// in-half's first entry to the back (the push-swap), an in-insert appends.
gt_status gt_host_adjacency(uint32_t n, uint64_t m, const uint32_t *src, const uint32_t *tgt,
                            uint32_t *in_off, uint32_t *in_other, uint32_t *in_eid,
                            uint32_t *out_off, uint32_t *out_other, uint32_t *out_eid) {
    gt::clear_error();
    if (m > 0xffffffffull) return gt::fail(GT_ERR_INVALID_ARG, "edge count exceeds the index width");
    if (!in_off || !out_off) return gt::fail(GT_ERR_INVALID_ARG, "null output");
    if (m && (!src || !tgt || !in_other || !in_eid || !out_other || !out_eid))
        return gt::fail(GT_ERR_INVALID_ARG, "null argument");
    for (uint64_t i = 0; i < m; ++i)
        if (src[i] >= n || tgt[i] >= n) return gt::fail(GT_ERR_NO_SUCH_VERTEX, "no such vertex");
    std::vector<uint32_t> outd(n, 0), ind(n, 0);
    for (uint64_t i = 0; i < m; ++i) {
        ++outd[src[i]];
        ++ind[tgt[i]];
    }
    in_off[0] = out_off[0] = 0;
    for (uint32_t v = 0; v < n; ++v) {
        in_off[v + 1] = in_off[v] + ind[v];
        out_off[v + 1] = out_off[v] + outd[v];
    }
    // Out-halves: insertion order.  In-halves: a ring buffer per vertex of
    // capacity in_degree, head[v] = position of the in-half's first entry.
    std::vector<uint32_t> ocur(out_off, out_off + n);
    std::vector<uint32_t> head(n, 0), size(n, 0);
    for (uint64_t i = 0; i < m; ++i) {
        const uint32_t s = src[i], t = tgt[i];
        out_other[ocur[s]] = t;
        out_eid[ocur[s]] = (uint32_t)i;
        ++ocur[s];
        // Push-swap on s's in-half: rotate by one when non-empty.
        if (size[s] > 0) {
            const uint32_t cap = ind[s], base = in_off[s];
            const uint32_t front = base + head[s];
            if (size[s] < cap) {
                uint32_t b = head[s] + size[s];
                if (b >= cap) b -= cap;
                in_other[base + b] = in_other[front];
                in_eid[base + b] = in_eid[front];
            }
            head[s] = (head[s] + 1 == cap) ? 0 : head[s] + 1;
        }
        {
            const uint32_t cap = ind[t], base = in_off[t];
            uint32_t b = head[t] + size[t];
            if (b >= cap) b -= cap;
            in_other[base + b] = s;
            in_eid[base + b] = (uint32_t)i;
            ++size[t];
        }
    }
    // Unroll each ring so the in-half starts at its first entry.
    std::vector<uint32_t> tmp_o, tmp_e;
    for (uint32_t v = 0; v < n; ++v) {
        const uint32_t cap = ind[v];
        if (cap < 2 || head[v] == 0) continue;
        const uint32_t base = in_off[v];
        tmp_o.assign(in_other + base, in_other + base + cap);
        tmp_e.assign(in_eid + base, in_eid + base + cap);
        for (uint32_t k = 0; k < cap; ++k) {
            in_other[base + k] = tmp_o[(head[v] + k) % cap];
            in_eid[base + k] = tmp_e[(head[v] + k) % cap];
        }
    }
    return GT_OK;
}

}  // extern "C"
