// Betweenness centrality (Brandes) on the device: crates/gt-algo/src/centrality.rs:500-689.
//
// Semantics (the reference's): unweighted, every vertex of the view a pivot,
// self-loops never on a shortest path, parallel edges distinct shortest paths,
// dependency factor (sigma[v] / sigma[u]) * (1 + dep[u]) per DAG edge v -> u,
// the source excluded from its own vertex sum, both maps zero-filled first and
// halved on an undirected view.  Path counts are f64, as in the reference.
//
// Layout: 64 sources run together, one per lane of a wavefront.  Per vertex
// the device holds rows of 64 values -- dist (u32), sigma and dep (f64) -- so
// that every neighbour a wavefront touches is one contiguous 256- or 512-byte
// row, and 64-bit lane masks: seen (lanes that reached the vertex) and the
// level masks of the current and the next BFS level.  Both phases are
// level-synchronous pulls:
//   forward, level L -> L + 1: a vertex w with unreached lanes U gathers the
//     level-L masks of its predecessors (the view's in-entries); each entry
//     whose mask meets U adds that predecessor's sigma row on those lanes.
//     Path counts are integers, so the sum is exact in any order.
//   backward, level L: a vertex v with lanes at level L gathers the level-(L+1)
//     masks of its successors (the view's out-entries); each DAG entry adds
//     its factor to dep[v] on those lanes, and the 64-lane sum of the factors
//     to the entry's edge accumulator.  Every entry and every vertex has one
//     owner per launch, so no atomics and a fixed summation order.
// Rows longer than `piece` entries are cut into pieces that separate
// wavefronts work on; a fix-up kernel sums their partial rows in piece order
// and finishes the row.  Edge sums are kept per entry position of the
// orientation they came from and scattered to edge ids at the end.
#include <algorithm>
#include <string>
#include <vector>

#include "device_common.hpp"
#include "gt_internal.hpp"
#include "props.hpp"

namespace gt {

constexpr uint32_t kBcUnreached = 0xffffffffu;
// Entries per piece of a long row (default).
constexpr uint32_t kBcPiece = 512;
// Forward levels launched between two reads of the level counter.
constexpr uint32_t kBcLevelsPerCheck = 4;

struct BcSeg {
    const uint32_t *off;
    const uint32_t *idx;
};

struct BcArgs {
    BcSeg seg[2];
    int nseg;
    uint32_t n;
    uint32_t piece;
    const uint4 *pieces;    // {row, begin, end, 0}: pieces of the long rows, row-major
    uint32_t npieces;
    const uint32_t *srow;   // long rows
    const uint32_t *sfirst; // first piece of each long row, nsplit + 1 entries
    uint32_t nsplit;
    uint64_t act;           // lanes holding a source in this batch
    uint64_t *seen;         // [n]
    uint64_t *mask[2];      // [n] each: level masks, level L in mask[L & 1]
    uint32_t *lo, *hi;      // [n]: first and last level of the vertex over the lanes
    uint32_t *dist;         // [n][64]
    double *sigma, *dep;    // [n][64]
    double *part;           // [npieces][64]
    uint64_t *pmask;        // [npieces]
    double *vacc;           // [n]
    double *eacc[2];        // per entry position of each segment
    uint32_t *level;        // deepest level reached so far (device counter)
    unsigned long long *work;  // [2]: entries scanned, lane hits (all launches)
};

__device__ __forceinline__ uint64_t readlane64(uint64_t x, int k) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, k);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), k);
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint32_t row_len(const BcArgs &a, uint32_t r) {
    uint32_t d = 0;
    for (int s = 0; s < a.nseg; ++s) d += a.seg[s].off[r + 1] - a.seg[s].off[r];
    return d;
}

// Calls f(segment, first entry, end entry) for the part of row r's virtual
// range [vb, ve) (segment 0's entries, then segment 1's) in each segment.
template <typename F>
__device__ __forceinline__ void piece_entries(const BcArgs &a, uint32_t r, uint32_t vb,
                                              uint32_t ve, F &&f) {
    uint32_t base = 0;
    for (int s = 0; s < a.nseg; ++s) {
        const uint32_t o = a.seg[s].off[r];
        const uint32_t len = a.seg[s].off[r + 1] - o;
        const uint32_t lo = max(vb, base), hi = min(ve, base + len);
        if (lo < hi) f(s, o + (lo - base), o + (hi - base));
        base += len;
    }
}

// Work item i of a level launch: rows [0, n) whole, then the pieces of the
// long rows.  Returns false for a long row in the first range (its pieces
// cover it).
__device__ __forceinline__ bool work_item(const BcArgs &a, uint32_t i, uint32_t *r, uint32_t *vb,
                                          uint32_t *ve) {
    if (i < a.n) {
        *r = i;
        *vb = 0;
        *ve = row_len(a, i);
        return *ve <= a.piece;
    }
    const uint4 p = a.pieces[i - a.n];
    *r = p.x;
    *vb = p.y;
    *ve = p.z;
    return true;
}

// ------------------------------------------------------------------ forward

// Entries [jb, je) of a row w whose unreached lanes are U: lanes add sigma of
// every predecessor at level L on the lanes its level mask shares with U.
__device__ __forceinline__ void fwd_entries(const BcArgs &a, const uint32_t *idx, uint32_t jb,
                                            uint32_t je, uint32_t w, uint64_t U,
                                            const uint64_t *cur, double &acc, uint64_t &H,
                                            uint32_t &hits_seen, int lane) {
    for (uint32_t base = jb; base < je; base += kWave) {
        const uint32_t j = base + lane;
        uint32_t u = w;
        uint64_t F = 0;
        if (j < je) {
            u = idx[j];
            if (u != w) F = cur[u] & U;
        }
        uint64_t hits = __ballot(F != 0);
        while (hits) {
            const int k = __builtin_ctzll(hits);
            hits &= hits - 1;
            const uint32_t uk = (uint32_t)__builtin_amdgcn_readlane((int)u, k);
            const uint64_t Fk = readlane64(F, k);
            H |= Fk;
            if ((Fk >> lane) & 1) {
                acc += a.sigma[(size_t)uk * kWave + lane];
                ++hits_seen;
            }
        }
    }
}

__device__ __forceinline__ bool fwd_finish(const BcArgs &a, uint64_t *nxt, uint32_t w,
                                           uint64_t seen, uint64_t H, double acc, uint32_t L,
                                           int lane) {
    if (lane == 0) nxt[w] = H;
    if (!H) return false;
    if (lane == 0) {
        a.seen[w] = seen | H;
        if (a.lo[w] == kBcUnreached) a.lo[w] = L + 1;
        a.hi[w] = L + 1;
    }
    if ((H >> lane) & 1) {
        a.dist[(size_t)w * kWave + lane] = L + 1;
        a.sigma[(size_t)w * kWave + lane] = acc;
    }
    return true;
}

__device__ __forceinline__ void add_work(const BcArgs &a, uint32_t entries, uint32_t hits,
                                         int lane) {
    const uint32_t h = (uint32_t)wave_sum((double)hits);
    if (lane == 0 && (entries | h)) {
        atomicAdd(&a.work[0], (unsigned long long)entries);
        atomicAdd(&a.work[1], (unsigned long long)h);
    }
}

// Level L -> L + 1.  Does nothing once a level came out empty.
__global__ __launch_bounds__(kBlock) void bc_forward(BcArgs a, uint32_t L) {
    if (__hip_atomic_load(a.level, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < L) return;
    const int lane = threadIdx.x & (kWave - 1);
    const uint32_t wave = (blockIdx.x * kBlock + threadIdx.x) / kWave;
    const uint32_t nwaves = gridDim.x * (kBlock / kWave);
    const uint64_t *cur = a.mask[L & 1];
    uint64_t *nxt = a.mask[(L + 1) & 1];
    bool found = false;
    uint32_t entries = 0, hits = 0;
    const uint32_t items = a.n + a.npieces;
    for (uint32_t i = wave; i < items; i += nwaves) {
        uint32_t w, vb, ve;
        if (!work_item(a, i, &w, &vb, &ve)) continue;
        const uint64_t seen = a.seen[w];
        const uint64_t U = ~seen & a.act;
        double acc = 0.0;
        uint64_t H = 0;
        if (U) {
            entries += ve - vb;
            piece_entries(a, w, vb, ve, [&](int s, uint32_t jb, uint32_t je) {
                fwd_entries(a, a.seg[s].idx, jb, je, w, U, cur, acc, H, hits, lane);
            });
        }
        if (i >= a.n) {
            const uint32_t p = i - a.n;
            a.part[(size_t)p * kWave + lane] = acc;
            if (lane == 0) a.pmask[p] = H;
        } else {
            found |= fwd_finish(a, nxt, w, seen, H, acc, L, lane);
        }
    }
    add_work(a, entries, hits, lane);
    if (found && lane == 0) atomicMax(a.level, L + 1);
}

// Finishes the long rows of level L -> L + 1 from their pieces, in piece order.
__global__ __launch_bounds__(kBlock) void bc_forward_fix(BcArgs a, uint32_t L) {
    if (__hip_atomic_load(a.level, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < L) return;
    const int lane = threadIdx.x & (kWave - 1);
    const uint32_t wave = (blockIdx.x * kBlock + threadIdx.x) / kWave;
    const uint32_t nwaves = gridDim.x * (kBlock / kWave);
    uint64_t *nxt = a.mask[(L + 1) & 1];
    bool found = false;
    for (uint32_t k = wave; k < a.nsplit; k += nwaves) {
        const uint32_t w = a.srow[k];
        double acc = 0.0;
        uint64_t H = 0;
        for (uint32_t p = a.sfirst[k]; p < a.sfirst[k + 1]; ++p) {
            acc += a.part[(size_t)p * kWave + lane];
            H |= a.pmask[p];
        }
        found |= fwd_finish(a, nxt, w, a.seen[w], H, acc, L, lane);
    }
    if (found && lane == 0) atomicMax(a.level, L + 1);
}

// ----------------------------------------------------------------- backward

// Lanes of row v at level L.
__device__ __forceinline__ uint64_t level_lanes(const BcArgs &a, uint32_t v, uint32_t L,
                                                int lane) {
    const uint32_t lo = a.lo[v], hi = a.hi[v];
    if (lo == kBcUnreached || L < lo || L > hi) return 0;
    return __ballot(a.dist[(size_t)v * kWave + lane] == L) & a.seen[v];
}

// Entries [jb, je) of a row v with lanes m at level L: every successor at
// level L + 1 on a lane of m contributes (sigma_v / sigma_u) * (1 + dep_u).
__device__ __forceinline__ void bwd_entries(const BcArgs &a, const uint32_t *idx, double *eacc,
                                            uint32_t jb, uint32_t je, uint32_t v, uint64_t m,
                                            const uint64_t *succ, double sv, double &dacc,
                                            uint32_t &hits_seen, int lane) {
#pragma clang fp contract(off)
    for (uint32_t base = jb; base < je; base += kWave) {
        const uint32_t j = base + lane;
        uint32_t u = v;
        uint64_t S = 0;
        if (j < je) {
            u = idx[j];
            if (u != v) S = succ[u] & m;
        }
        const uint64_t hit0 = __ballot(S != 0);
        uint64_t hits = hit0;
        double mine = 0.0;
        while (hits) {
            const int k = __builtin_ctzll(hits);
            hits &= hits - 1;
            const uint32_t uk = (uint32_t)__builtin_amdgcn_readlane((int)u, k);
            const uint64_t Sk = readlane64(S, k);
            double f = 0.0;
            if ((Sk >> lane) & 1) {
                const size_t q = (size_t)uk * kWave + lane;
                f = (sv / a.sigma[q]) * (1.0 + a.dep[q]);
                dacc += f;
                ++hits_seen;
            }
            const double t = wave_sum(f);
            if (lane == k) mine = t;
        }
        if ((hit0 >> lane) & 1) eacc[j] += mine;
    }
}

__device__ __forceinline__ void bwd_finish(const BcArgs &a, uint64_t *curm, uint32_t v,
                                           uint64_t m, double dacc, uint32_t L, int lane) {
    if (lane == 0) curm[v] = m;
    if (!m) return;
    const bool on = (m >> lane) & 1;
    if (on) a.dep[(size_t)v * kWave + lane] = dacc;
    if (L > 0) {  // level 0 is the source itself
        const double t = wave_sum(on ? dacc : 0.0);
        if (lane == 0) a.vacc[v] += t;
    }
}

// Level L (top = the deepest level, which has no successors).
__global__ __launch_bounds__(kBlock) void bc_backward(BcArgs a, uint32_t L, uint32_t top) {
    const int lane = threadIdx.x & (kWave - 1);
    const uint32_t wave = (blockIdx.x * kBlock + threadIdx.x) / kWave;
    const uint32_t nwaves = gridDim.x * (kBlock / kWave);
    const uint64_t *succ = a.mask[(L + 1) & 1];
    uint64_t *curm = a.mask[L & 1];
    uint32_t entries = 0, hits = 0;
    const uint32_t items = a.n + a.npieces;
    for (uint32_t i = wave; i < items; i += nwaves) {
        uint32_t v, vb, ve;
        if (!work_item(a, i, &v, &vb, &ve)) continue;
        const uint64_t m = level_lanes(a, v, L, lane);
        double dacc = 0.0;
        if (m && L < top) {
            const double sv = ((m >> lane) & 1) ? a.sigma[(size_t)v * kWave + lane] : 1.0;
            entries += ve - vb;
            piece_entries(a, v, vb, ve, [&](int s, uint32_t jb, uint32_t je) {
                bwd_entries(a, a.seg[s].idx, a.eacc[s], jb, je, v, m, succ, sv, dacc, hits,
                            lane);
            });
        }
        if (i >= a.n)
            a.part[(size_t)(i - a.n) * kWave + lane] = dacc;
        else
            bwd_finish(a, curm, v, m, dacc, L, lane);
    }
    add_work(a, entries, hits, lane);
}

__global__ __launch_bounds__(kBlock) void bc_backward_fix(BcArgs a, uint32_t L) {
    const int lane = threadIdx.x & (kWave - 1);
    const uint32_t wave = (blockIdx.x * kBlock + threadIdx.x) / kWave;
    const uint32_t nwaves = gridDim.x * (kBlock / kWave);
    uint64_t *curm = a.mask[L & 1];
    for (uint32_t k = wave; k < a.nsplit; k += nwaves) {
        const uint32_t v = a.srow[k];
        const uint64_t m = level_lanes(a, v, L, lane);
        double dacc = 0.0;
        for (uint32_t p = a.sfirst[k]; p < a.sfirst[k + 1]; ++p)
            dacc += a.part[(size_t)p * kWave + lane];
        bwd_finish(a, curm, v, m, dacc, L, lane);
    }
}

// ----------------------------------------------------------- batch set-up

__global__ __launch_bounds__(kBlock) void bc_reset(BcArgs a) {
    for (uint32_t v = blockIdx.x * kBlock + threadIdx.x; v < a.n; v += gridDim.x * kBlock) {
        a.seen[v] = 0;
        a.mask[0][v] = 0;
        a.lo[v] = kBcUnreached;
        a.hi[v] = 0;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *a.level = 0;
}

// Lane b starts at src[b] (level 0, one path).  A source listed twice gets
// two lanes, hence the atomic ORs.
__global__ void bc_seed(BcArgs a, const uint32_t *src, uint32_t count) {
    const uint32_t b = threadIdx.x;
    if (b >= count) return;
    const uint32_t s = src[b];
    a.dist[(size_t)s * kWave + b] = 0;
    a.sigma[(size_t)s * kWave + b] = 1.0;
    atomicOr((unsigned long long *)&a.seen[s], 1ull << b);
    atomicOr((unsigned long long *)&a.mask[0][s], 1ull << b);
    a.lo[s] = 0;
    a.hi[s] = 0;
}

// ------------------------------------------------------------------ output

__global__ __launch_bounds__(kBlock) void bc_emit_vertices(const double *vacc, uint32_t n,
                                                           bool halve, double *out) {
    for (uint32_t v = blockIdx.x * kBlock + threadIdx.x; v < n; v += gridDim.x * kBlock)
        out[v] = halve ? vacc[v] / 2.0 : vacc[v];
}

__global__ __launch_bounds__(kBlock) void bc_fill(double *out, uint64_t len) {
    for (uint64_t e = blockIdx.x * (uint64_t)kBlock + threadIdx.x; e < len;
         e += (uint64_t)gridDim.x * kBlock)
        out[e] = 0.0;
}

// out[eid[j]] (+)= acc[j] for the entries of one segment; each edge id
// appears once per segment.
__global__ __launch_bounds__(kBlock) void bc_emit_edges(const uint32_t *eid, const double *acc,
                                                        uint64_t entries, bool add, bool halve,
                                                        double *out) {
    for (uint64_t j = blockIdx.x * (uint64_t)kBlock + threadIdx.x; j < entries;
         j += (uint64_t)gridDim.x * kBlock) {
        double x = add ? out[eid[j]] + acc[j] : acc[j];
        out[eid[j]] = halve ? x / 2.0 : x;
    }
}

namespace {

unsigned lgrid(uint64_t n) {
    uint64_t b = (n + kBlock - 1) / kBlock;
    if (b > 4096) b = 4096;
    return b ? (unsigned)b : 1u;
}

struct BcView {
    Orient *fwd[2] = {nullptr, nullptr};  // rows list the predecessors
    Orient *bwd[2] = {nullptr, nullptr};  // rows list the successors (out_edges)
    int nseg = 1;
    bool undirected = false;
    const uint8_t *host_vmask = nullptr;
};

// out_edges of each view kind (GraphRef for AdjList, Rev, Und and their
// filtered forms): the out-half, the in-half when reversed, the whole block
// (out-half first) when undirected (view/undirected.rs:110-112).
gt_status bc_view(gt_graph *g, const gt_view *view, BcView *bv) {
    const int kind = view ? view->kind : GT_VIEW_DIRECTED;
    Orient *pull = nullptr, *push = nullptr;
    const uint8_t *act = nullptr;
    uint32_t nact = 0;
    switch (kind) {
    case GT_VIEW_DIRECTED:
        pull = &g->in;
        push = &g->out;
        break;
    case GT_VIEW_REVERSED:
        pull = &g->out;
        push = &g->in;
        break;
    case GT_VIEW_DIRECTED_FILTERED:
    case GT_VIEW_REVERSED_FILTERED:
    case GT_VIEW_UNDIRECTED_FILTERED:
        GT_TRY(filtered_view(g, view, kind == GT_VIEW_REVERSED_FILTERED, &pull, &push, &act,
                             &nact));
        bv->host_vmask = view->vmask;
        break;
    case GT_VIEW_UNDIRECTED:
        pull = &g->in;
        push = &g->out;
        break;
    default:
        return fail(GT_ERR_INVALID_ARG, "unknown view kind");
    }
    if (kind == GT_VIEW_UNDIRECTED || kind == GT_VIEW_UNDIRECTED_FILTERED) {
        bv->undirected = true;
        bv->nseg = 2;
        bv->fwd[0] = bv->bwd[0] = push;  // out-half first
        bv->fwd[1] = bv->bwd[1] = pull;
    } else {
        bv->fwd[0] = pull;
        bv->bwd[0] = push;
    }
    return GT_OK;
}

gt_status check_map(gt_graph *g, const double *data, uint64_t len, uint64_t need, int32_t loc,
                    uint64_t graph_id, const char *what) {
    if (graph_id && graph_id != g->id)
        return fail(GT_ERR_WRONG_GRAPH, std::string("betweenness: ") + what +
                                            " map does not belong to this graph");
    if (!data && need) return fail(GT_ERR_INVALID_ARG, std::string(what) + " map is null");
    if (len < need)
        return fail(GT_ERR_INVALID_ARG, std::string(what) + " map is shorter than its bound");
    if (loc != GT_HOST && loc != GT_DEVICE) return fail(GT_ERR_INVALID_ARG, "bad location");
    return GT_OK;
}

// The pieces of the long rows over `segs` (host offsets).
struct Pieces {
    std::vector<uint32_t> q;  // 4 per piece: row, begin, end, 0
    std::vector<uint32_t> srow, sfirst;
};

void cut_pieces(const BcView &bv, const Orient *const *segs, uint32_t n, uint32_t piece,
                Pieces &P) {
    P.sfirst.push_back(0);
    for (uint32_t r = 0; r < n; ++r) {
        uint32_t len = 0;
        for (int s = 0; s < bv.nseg; ++s) len += segs[s]->h_off[r + 1] - segs[s]->h_off[r];
        if (len <= piece) continue;
        P.srow.push_back(r);
        for (uint32_t b = 0; b < len; b += piece) {
            P.q.insert(P.q.end(), {r, b, std::min(len, b + piece), 0u});
        }
        P.sfirst.push_back((uint32_t)(P.q.size() / 4));
    }
}

}  // namespace

gt_status betweenness_impl(gt_graph *g, const gt_view *view, const uint32_t *pivots,
                           uint64_t npivots, const gt_bc_out *out, gt_bc_stats *stats) {
    clear_error();
    if (!g || !out) return fail(GT_ERR_INVALID_ARG, "null argument");
    const uint32_t n = g->n;
    const uint64_t eb = g->edge_bound;
    GT_TRY(check_map(g, out->vertex, out->vertex_len, n, out->loc, out->vertex_graph_id,
                     "vertex"));
    if (out->edge) GT_TRY(check_map(g, out->edge, out->edge_len, eb, out->loc,
                                    out->edge_graph_id, "edge"));
    else if (out->edge_graph_id && out->edge_graph_id != g->id)
        return fail(GT_ERR_WRONG_GRAPH, "betweenness: edge map does not belong to this graph");
    if (pivots == nullptr && npivots != 0) return fail(GT_ERR_INVALID_ARG, "null pivot list");
    BcView bv;
    GT_TRY(bc_view(g, view, &bv));
    hipStream_t s = g->stream;

    // Sources: the pivots (every vertex of the view by default) whose
    // out_edges are not empty; a source with none reaches nothing and adds
    // nothing.
    std::vector<uint32_t> src;
    auto out_len = [&](uint32_t v) {
        uint64_t d = 0;
        for (int q = 0; q < bv.nseg; ++q) d += bv.bwd[q]->h_off[v + 1] - bv.bwd[q]->h_off[v];
        return d;
    };
    if (pivots) {
        for (uint64_t i = 0; i < npivots; ++i) {
            const uint32_t v = pivots[i];
            if (v >= n) return fail(GT_ERR_NO_SUCH_VERTEX, "betweenness: pivot out of range");
            if (bv.host_vmask && !bv.host_vmask[v])
                return fail(GT_ERR_INVALID_ARG, "betweenness: pivot is not a vertex of the view");
            if (out_len(v)) src.push_back(v);
        }
    } else {
        for (uint32_t v = 0; v < n; ++v)
            if ((!bv.host_vmask || bv.host_vmask[v]) && out_len(v)) src.push_back(v);
    }

    const uint32_t piece = kBcPiece;
    Pieces fp, bp;
    cut_pieces(bv, bv.fwd, n, piece, fp);
    cut_pieces(bv, bv.bwd, n, piece, bp);

    Workspace &ws = g->ws;
    const size_t rows = (size_t)std::max<uint32_t>(n, 1);
    uint64_t seg_entries[2] = {0, 0};
    for (int q = 0; q < bv.nseg; ++q) seg_entries[q] = bv.bwd[q]->h_off[n];
    const size_t np = std::max<size_t>(std::max(fp.q.size(), bp.q.size()) / 4, 1);
    const size_t nsplit = std::max<size_t>(std::max(fp.srow.size(), bp.srow.size()), 1);
    DevBuf *B = ws.btw;
    GT_TRY(B[0].reserve(rows * kWave * sizeof(double)));           // sigma
    GT_TRY(B[1].reserve(rows * kWave * sizeof(double)));           // dep
    GT_TRY(B[2].reserve(rows * kWave * sizeof(uint32_t)));         // dist
    GT_TRY(B[3].reserve(rows * 3 * sizeof(uint64_t) + 64));        // seen, mask[2], level, work
    GT_TRY(B[4].reserve(rows * 2 * sizeof(uint32_t)));             // lo, hi
    GT_TRY(B[5].reserve(rows * sizeof(double)));                   // vacc
    GT_TRY(B[6].reserve(std::max<uint64_t>(seg_entries[0] + seg_entries[1], 1) * sizeof(double)));
    GT_TRY(B[7].reserve(np * kWave * sizeof(double) + np * sizeof(uint64_t)));
    GT_TRY(B[8].reserve((np * 4 + 2 * (nsplit + 1)) * sizeof(uint32_t) * 2));
    GT_TRY(B[9].reserve(std::max<size_t>(src.size(), 1) * sizeof(uint32_t)));

    BcArgs base{};
    base.n = n;
    base.piece = piece;
    base.nseg = bv.nseg;
    base.sigma = B[0].as<double>();
    base.dep = B[1].as<double>();
    base.dist = B[2].as<uint32_t>();
    base.seen = B[3].as<uint64_t>();
    base.mask[0] = base.seen + rows;
    base.mask[1] = base.mask[0] + rows;
    base.level = reinterpret_cast<uint32_t *>(base.mask[1] + rows);
    base.work = reinterpret_cast<unsigned long long *>(base.mask[1] + rows + 1);
    base.lo = B[4].as<uint32_t>();
    base.hi = base.lo + rows;
    base.vacc = B[5].as<double>();
    base.eacc[0] = B[6].as<double>();
    base.eacc[1] = base.eacc[0] + seg_entries[0];
    base.part = B[7].as<double>();
    base.pmask = reinterpret_cast<uint64_t *>(base.part + np * kWave);
    // Piece tables: forward then backward.
    uint32_t *tab = B[8].as<uint32_t>();
    auto upload = [&](const std::vector<uint32_t> &h, uint32_t *d) -> gt_status {
        if (!h.empty())
            GT_HIP(hipMemcpyAsync(d, h.data(), h.size() * sizeof(uint32_t),
                                  hipMemcpyHostToDevice, s));
        return GT_OK;
    };
    uint32_t *fq = tab, *bq = fq + np * 4;  // 16-byte aligned piece records first
    uint32_t *fs = bq + np * 4, *ff = fs + nsplit;
    uint32_t *bs = ff + nsplit + 1, *bf = bs + nsplit;
    GT_TRY(upload(fp.q, fq));
    GT_TRY(upload(fp.srow, fs));
    GT_TRY(upload(fp.sfirst, ff));
    GT_TRY(upload(bp.q, bq));
    GT_TRY(upload(bp.srow, bs));
    GT_TRY(upload(bp.sfirst, bf));
    uint32_t *dsrc = B[9].as<uint32_t>();
    GT_TRY(upload(src, dsrc));

    BcArgs fa = base, ba = base;
    for (int q = 0; q < bv.nseg; ++q) {
        fa.seg[q] = {bv.fwd[q]->off.as<uint32_t>(), bv.fwd[q]->idx.as<uint32_t>()};
        ba.seg[q] = {bv.bwd[q]->off.as<uint32_t>(), bv.bwd[q]->idx.as<uint32_t>()};
    }
    fa.pieces = reinterpret_cast<const uint4 *>(fq);
    fa.npieces = (uint32_t)(fp.q.size() / 4);
    fa.srow = fs;
    fa.sfirst = ff;
    fa.nsplit = (uint32_t)fp.srow.size();
    ba.pieces = reinterpret_cast<const uint4 *>(bq);
    ba.npieces = (uint32_t)(bp.q.size() / 4);
    ba.srow = bs;
    ba.sfirst = bf;
    ba.nsplit = (uint32_t)bp.srow.size();

    GT_HIP(hipMemsetAsync(base.vacc, 0, rows * sizeof(double), s));
    GT_HIP(hipMemsetAsync(base.eacc[0], 0,
                          std::max<uint64_t>(seg_entries[0] + seg_entries[1], 1) * sizeof(double),
                          s));
    GT_HIP(hipMemsetAsync(base.work, 0, 2 * sizeof(unsigned long long), s));

    const int fcap = occupancy_cap((const void *)bc_forward, 8);
    const int bcap = occupancy_cap((const void *)bc_backward, 8);
    const uint64_t fitems = (uint64_t)n + fa.npieces, bitems = (uint64_t)n + ba.npieces;
    const unsigned fgrid = (unsigned)std::max<uint64_t>(
        1, std::min<uint64_t>(fcap, (fitems + kWavesPerBlock - 1) / kWavesPerBlock));
    const unsigned bgrid = (unsigned)std::max<uint64_t>(
        1, std::min<uint64_t>(bcap, (bitems + kWavesPerBlock - 1) / kWavesPerBlock));
    const unsigned fixf = (unsigned)std::max<uint32_t>(1, (fa.nsplit + kWavesPerBlock - 1) / kWavesPerBlock);
    const unsigned fixb = (unsigned)std::max<uint32_t>(1, (ba.nsplit + kWavesPerBlock - 1) / kWavesPerBlock);

    hipEvent_t ev[4] = {};
    const bool prof = g->profiling;
    if (prof)
        for (auto &e : ev) GT_HIP(hipEventCreate(&e));
    uint64_t levels_total = 0, launches_f = 0, launches_b = 0;
    double fms = 0.0, bms = 0.0;
    uint32_t probe = kBcLevelsPerCheck;
    for (size_t first = 0; first < src.size(); first += kWave) {
        const uint32_t cnt = (uint32_t)std::min<size_t>(kWave, src.size() - first);
        const uint64_t act = cnt == kWave ? ~0ull : ((1ull << cnt) - 1);
        fa.act = ba.act = act;
        hipLaunchKernelGGL(bc_reset, dim3(lgrid(n)), dim3(kBlock), 0, s, fa);
        hipLaunchKernelGGL(bc_seed, dim3(1), dim3(kWave), 0, s, fa, dsrc + first, cnt);
        GT_HIP(hipGetLastError());
        if (prof) GT_HIP(hipEventRecord(ev[0], s));
        uint32_t L = 0, top = 0;
        for (;;) {
            for (uint32_t k = 0; k < probe; ++k, ++L) {
                hipLaunchKernelGGL(bc_forward, dim3(fgrid), dim3(kBlock), 0, s, fa, L);
                if (fa.nsplit)
                    hipLaunchKernelGGL(bc_forward_fix, dim3(fixf), dim3(kBlock), 0, s, fa, L);
                ++launches_f;
            }
            GT_HIP(hipGetLastError());
            GT_HIP(hipMemcpyAsync(&g->h_state->counter, base.level, sizeof(uint32_t),
                                  hipMemcpyDeviceToHost, s));
            GT_HIP(hipStreamSynchronize(s));
            top = g->h_state->counter;
            if (top < L) break;  // a level came out empty
        }
        // Launch enough levels in one go next time.
        probe = std::max<uint32_t>(kBcLevelsPerCheck, top + 2);
        if (prof) {
            GT_HIP(hipEventRecord(ev[1], s));
            GT_HIP(hipEventRecord(ev[2], s));
        }
        for (uint32_t l = top + 1; l-- > 0;) {
            hipLaunchKernelGGL(bc_backward, dim3(bgrid), dim3(kBlock), 0, s, ba, l, top);
            if (ba.nsplit)
                hipLaunchKernelGGL(bc_backward_fix, dim3(fixb), dim3(kBlock), 0, s, ba, l);
            ++launches_b;
        }
        GT_HIP(hipGetLastError());
        levels_total += top + 1;
        if (prof) {
            GT_HIP(hipEventRecord(ev[3], s));
            GT_HIP(hipEventSynchronize(ev[3]));
            float a = 0.f, b = 0.f;
            GT_HIP(hipEventElapsedTime(&a, ev[0], ev[1]));
            GT_HIP(hipEventElapsedTime(&b, ev[2], ev[3]));
            fms += a;
            bms += b;
        }
    }
    if (prof) {
        for (auto &e : ev) (void)hipEventDestroy(e);
        g->prof_ms[GT_SWEEP_BC_FORWARD] += fms;
        g->prof_launches[GT_SWEEP_BC_FORWARD] += launches_f;
        g->prof_ms[GT_SWEEP_BC_BACKWARD] += bms;
        g->prof_launches[GT_SWEEP_BC_BACKWARD] += launches_b;
    }

    // Outputs: vertex sums, edge sums scattered from entry positions.
    double *vdst = out->vertex;
    if (out->loc == GT_HOST) {
        GT_TRY(ws.y3.reserve(rows * sizeof(double)));
        vdst = ws.y3.as<double>();
    }
    if (n)
        hipLaunchKernelGGL(bc_emit_vertices, dim3(lgrid(n)), dim3(kBlock), 0, s, base.vacc, n,
                           bv.undirected, vdst);
    if (out->edge && eb) {
        double *edst = out->edge;
        if (out->loc == GT_HOST) {
            GT_TRY(ws.gw.reserve(eb * sizeof(double)));
            edst = ws.gw.as<double>();
        }
        hipLaunchKernelGGL(bc_fill, dim3(lgrid(eb)), dim3(kBlock), 0, s, edst, eb);
        // Every edge of the view has one entry in each segment, so the halving
        // rides on the last segment's scatter.
        for (int q = 0; q < bv.nseg; ++q)
            if (seg_entries[q])
                hipLaunchKernelGGL(bc_emit_edges, dim3(lgrid(seg_entries[q])), dim3(kBlock), 0,
                                   s, bv.bwd[q]->eid.as<uint32_t>(), base.eacc[q],
                                   seg_entries[q], q > 0, bv.undirected && q == bv.nseg - 1,
                                   edst);
        GT_HIP(hipGetLastError());
        if (out->loc == GT_HOST)
            GT_HIP(hipMemcpyAsync(out->edge, edst, eb * sizeof(double), hipMemcpyDeviceToHost, s));
    }
    GT_HIP(hipGetLastError());
    if (out->loc == GT_HOST && n)
        GT_HIP(hipMemcpyAsync(out->vertex, vdst, n * sizeof(double), hipMemcpyDeviceToHost, s));
    unsigned long long work[2] = {0, 0};
    GT_HIP(hipMemcpyAsync(work, base.work, sizeof(work), hipMemcpyDeviceToHost, s));
    GT_HIP(hipStreamSynchronize(s));
    if (stats) {
        stats->sources = src.size();
        stats->batches = (src.size() + kWave - 1) / kWave;
        stats->levels = levels_total;
        stats->entries_scanned = work[0];
        stats->lane_hits = work[1];
        stats->forward_ms = fms;
        stats->backward_ms = bms;
    }
    return GT_OK;
}

}  // namespace gt

using namespace gt;

extern "C" {

gt_status gt_betweenness(gt_graph *g, const gt_view *view, const uint32_t *pivots,
                         uint64_t npivots, const gt_bc_out *out, gt_bc_stats *stats) {
    return betweenness_impl(g, view, pivots, npivots, out, stats);
}

}  // extern "C"
