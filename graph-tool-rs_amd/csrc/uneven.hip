// Uneven batch (extension, no counterpart in the reference): thousands of
// independent directed graphs of very different sizes, solved together.
//
// The graphs are the vertex ranges [voff[g], voff[g+1]) of one block-diagonal
// graph (no edge joins two ranges).  Each graph g runs pagerank
// (crates/gt-algo/src/centrality.rs:234-384 with pers = 1 / n_g, its own
// damping d_g) in the FP32 form -- FP32 values, weights and SpMV accumulation,
// FP64 dangling mass, teleport update and L1 test -- stopped on its own L1
// change and frozen from then on; its backward solves the transposed system
// of the single-graph adjoint (adjoint.hip) with d_g.
//
// One launch pair per sweep serves every graph: the wave engine's tiles are
// cut at graph bounds, so each tile belongs to one graph; a tile whose graph
// has converged is skipped before any load, so the work of a sweep shrinks as
// graphs finish.  Each tile writes its own (delta, dangling) partials; the
// fold kernel gives every graph one wavefront, which sums that graph's tile
// partials in tile order, finishes its split rows and decides its
// convergence -- no cross-graph reduction and no ticket.  A graph's latest
// iterate sits in the buffer of its own sweep parity.
//
// The sweep/fold pairs of kUnevenCapture sweeps are captured once per call in
// a HIP graph and the graph is replayed until no graph is active (a count the
// fold keeps on the device), so the host issues one launch per
// kUnevenCapture sweeps.  No memset node is recorded: every buffer a replay
// needs cleared is cleared by a kernel.
//
// A sweep walks a list of the tiles of the graphs still iterating, which the
// fold before it writes (one atomic add per graph for its place in the list),
// so converged graphs cost nothing.  Each tile is still summed by one
// wavefront in a fixed order and the fold reads the partials in tile order:
// the list's order changes no result.
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <vector>

#include "blocks.hpp"
#include "device_common.hpp"
#include "gt_internal.hpp"
#include "props.hpp"
#include "sweep_common.hpp"

namespace gt {

constexpr int kUnevenCapture = 8;

// Per-graph state, structure of arrays.
struct UState {
    double *p_sink;    // dangling mass (forward) / p . lambda (adjoint)
    double *delta;     // last L1 change
    double *scale;     // adjoint: ||dL/dx_g||_1; forward: unused
    uint64_t *iter;    // sweeps done
    uint32_t *done;    // converged or capped
    uint32_t *nsink;   // sinks of the graph (forward init)
};

struct UArgs {
    const Tile *tiles;
    uint32_t ntiles;
    uint32_t ngraphs;
    const uint32_t *off;        // pull (forward) / push (adjoint) row offsets
    const uint32_t *packed;
    uint32_t bits;
    const uint32_t *tbase;      // source base of each tile
    uint32_t short_row;        // rows with fewer entries are summed by one lane
    const float *w;             // packed-order weights (weighted)
    const uint32_t *deg_off;    // unity out-degree: push orientation offsets
    const double *deg;          // weighted out-degree
    const uint32_t *tile_graph; // [ntiles]
    const uint32_t *voff;       // [G+1]
    const uint32_t *tb;         // [G+1]: first tile of each graph
    const uint32_t *lb;         // [G+1]: first split row of each graph
    const uint32_t *long_rows;
    const uint32_t *long_slot;
    double *long_part;
    const double *damp;         // [G]
    const double *cvec;         // adjoint: dL/dx (f64)
    double eps;
    double eps_rel;             // adjoint FP32 floor (relative to ||dL/dx_g||_1)
    uint64_t max_iter;
    const float *src;           // gathered: c (forward) / lambda (adjoint)
    const float *x_cur;         // forward x / adjoint lambda
    float *x_next;
    float *c_next;              // forward only
    double *tpart;              // [2 * ntiles]
    UState st;
    uint32_t *active;           // graphs still iterating
    uint32_t *list;             // [ntiles]: tiles of the graphs still iterating
    uint32_t *list_count;       // [2]: list lengths, by sweep parity
    uint32_t *list_n;           // this sweep's list length (an entry of list_count)
    uint32_t *list_n_next;      // the next sweep's (the other entry)
};

__device__ __forceinline__ uint32_t uniform_u32(uint32_t v) {
    return __builtin_amdgcn_readfirstlane(v);
}

// Forward finisher: the PageRank update of row i with its graph's d_g,
// p_sink_g and pers 1/n_g.
template <bool WEIGHTED>
struct UFwdFin {
    const UArgs &a;
    double d = 0.0, ps = 0.0, pv = 0.0;
    double dsum = 0.0, ssum = 0.0;
    __device__ explicit UFwdFin(const UArgs &args) : a(args) {}
    __device__ __forceinline__ bool set_graph(uint32_t g) {
        if (a.st.done[g]) return false;
        d = a.damp[g];
        ps = a.st.p_sink[g];
        pv = 1.0 / (double)(a.voff[g + 1] - a.voff[g]);
        dsum = 0.0;
        ssum = 0.0;
        return true;
    }
    __device__ __forceinline__ bool tile(uint32_t ti) {
        return set_graph(uniform_u32(a.tile_graph[ti]));
    }
    __device__ __forceinline__ void operator()(uint32_t i, double acc) {
        const double nv = (1.0 - d) * pv + d * (ps * pv + acc);
        const float nf = static_cast<float>(nv);
        dsum += fabs(static_cast<double>(nf) - static_cast<double>(a.x_cur[i]));
        a.x_next[i] = nf;
        const double dg =
            WEIGHTED ? a.deg[i] : static_cast<double>(a.deg_off[i + 1] - a.deg_off[i]);
        if (dg != 0.0) {
            a.c_next[i] = static_cast<float>(static_cast<double>(nf) / dg);
        } else {
            a.c_next[i] = 0.f;
            ssum += static_cast<double>(nf);
        }
    }
    __device__ __forceinline__ double second() const { return ssum; }
    __device__ __forceinline__ void tile_end(uint32_t ti) {
        const double dt = wave_sum(dsum), st = wave_sum(ssum);
        if ((threadIdx.x & (kWave - 1)) == 0) {
            a.tpart[2 * (size_t)ti] = dt;
            a.tpart[2 * (size_t)ti + 1] = st;
        }
    }
};

// Adjoint finisher: lambda_s = c_s + d_g mu_s + (deg_s == 0 ? d_g (p.lambda)_g : 0).
template <bool WEIGHTED>
struct UAdjFin {
    const UArgs &a;
    double d = 0.0, pl = 0.0, pv = 0.0;
    double dsum = 0.0, psum = 0.0;
    __device__ explicit UAdjFin(const UArgs &args) : a(args) {}
    __device__ __forceinline__ bool set_graph(uint32_t g) {
        if (a.st.done[g]) return false;
        d = a.damp[g];
        pl = a.st.p_sink[g];
        pv = 1.0 / (double)(a.voff[g + 1] - a.voff[g]);
        dsum = 0.0;
        psum = 0.0;
        return true;
    }
    __device__ __forceinline__ bool tile(uint32_t ti) {
        return set_graph(uniform_u32(a.tile_graph[ti]));
    }
    __device__ __forceinline__ void operator()(uint32_t s, double acc) {
        const double dg =
            WEIGHTED ? a.deg[s] : static_cast<double>(a.deg_off[s + 1] - a.deg_off[s]);
        const double mu = dg != 0.0 ? acc / dg : 0.0;
        const double v = a.cvec[s] + d * mu + (dg == 0.0 ? d * pl : 0.0);
        const float vf = static_cast<float>(v);
        dsum += fabs(static_cast<double>(vf) - static_cast<double>(a.x_cur[s]));
        a.x_next[s] = vf;
        psum += pv * static_cast<double>(vf);
    }
    __device__ __forceinline__ double second() const { return psum; }
    __device__ __forceinline__ void tile_end(uint32_t ti) {
        const double dt = wave_sum(dsum), pt = wave_sum(psum);
        if ((threadIdx.x & (kWave - 1)) == 0) {
            a.tpart[2 * (size_t)ti] = dt;
            a.tpart[2 * (size_t)ti + 1] = pt;
        }
    }
};

template <typename Fin, bool WEIGHTED>
__global__ __launch_bounds__(kBlock) void u_sweep(const UArgs a) {
    if (*a.active == 0) return;
    // The listed tiles; the fold after this sweep counts the next list from 0.
    const uint32_t nt = uniform_u32(*a.list_n);
    if (blockIdx.x == 0 && threadIdx.x == 0) *a.list_n_next = 0u;
    __shared__ WaveLds<float> L;
    Fin fin(a);
    // Workgroups are dealt to the XCDs in turn (workgroup b on XCD b % 8), so
    // XCD x walks the x-th eighth of the list: its L2 then holds the vectors
    // of the graphs in that eighth, not of every graph at once.
    const XcdSpan sp = xcd_span(nt);
    spmv_waves<float, WEIGHTED>(a.tiles, sp.count, a.off, a.packed, a.bits, a.tbase,
                                a.src, a.w, a.long_part, a.short_row, L, fin,
                                static_cast<const float *>(nullptr), 0u, a.list + sp.begin,
                                sp.first(kWavesPerBlock), sp.stride(kWavesPerBlock));
}

// Before the first sweep of a call: every tile listed.
__global__ __launch_bounds__(kBlock) void u_list_init(UArgs a) {
    const uint32_t i0 = blockIdx.x * kBlock + threadIdx.x;
    for (uint32_t i = i0; i < a.ntiles; i += gridDim.x * kBlock) a.list[i] = i;
    if (i0 == 0) {
        a.list_count[0] = a.ntiles;
        a.list_count[1] = 0u;
    }
}

// One wavefront per graph: the graph's tile partials in tile order, its split
// rows, then its convergence decision.  Forward: stop when delta < eps;
// adjoint: when delta < max(eps, eps_rel) * ||dL/dx_g||_1.  Both stop at
// max_iter sweeps when it is non-zero.
template <typename Fin, bool ADJ>
__global__ __launch_bounds__(kBlock) void u_fold(const UArgs a) {
    if (*a.active == 0) return;
    const int lane = threadIdx.x & (kWave - 1);
    const uint32_t g = uniform_u32(blockIdx.x * kWavesPerBlock + threadIdx.x / kWave);
    if (g >= a.ngraphs) return;
    Fin fin(a);
    if (!fin.set_graph(g)) return;
    double d = 0.0, s = 0.0;
    for (uint32_t t = a.tb[g] + lane; t < a.tb[g + 1]; t += kWave) {
        d += a.tpart[2 * (size_t)t];
        s += a.tpart[2 * (size_t)t + 1];
    }
    for (uint32_t k = a.lb[g] + lane; k < a.lb[g + 1]; k += kWave)
        fin(a.long_rows[k], long_row_sum(a.long_part, a.long_slot, k));
    d = wave_sum(d + fin.dsum);
    s = wave_sum(s + fin.second());
    bool more = false;
    if (lane == 0) {
        const uint64_t it = a.st.iter[g] + 1;
        const double thr = ADJ ? fmax(a.eps, a.eps_rel) * (a.st.scale[g] > 0.0 ? a.st.scale[g] : 1.0)
                               : a.eps;
        a.st.iter[g] = it;
        a.st.delta[g] = d;
        a.st.p_sink[g] = s;
        if (!(d >= thr) || (a.max_iter > 0 && it >= a.max_iter)) {
            a.st.done[g] = 1u;
            atomicSub(a.active, 1u);
        } else {
            more = true;
        }
    }
    // A graph still iterating lists its tiles for the next sweep.
    if (__ballot(more)) {
        const uint32_t t0 = a.tb[g], cnt = a.tb[g + 1] - t0;
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(a.list_n_next, cnt);
        base = __shfl(base, 0, kWave);
        for (uint32_t i = lane; i < cnt; i += kWave) a.list[base + i] = t0 + i;
    }
}

// vgraph[v] = the graph whose range holds v (binary search over voff).
__global__ __launch_bounds__(kBlock) void u_vgraph(const uint32_t *__restrict__ voff, uint32_t G,
                                                   uint32_t n, uint32_t *__restrict__ vgraph) {
    for (uint32_t v = blockIdx.x * kBlock + threadIdx.x; v < n; v += gridDim.x * kBlock) {
        uint32_t lo = 0, hi = G;  // voff[lo] <= v < voff[hi]
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (voff[mid] <= v) lo = mid;
            else hi = mid;
        }
        vgraph[v] = lo;
    }
}

// Counts edges whose endpoints lie in different graphs (out-halves: row s,
// entry t).
__global__ __launch_bounds__(kBlock) void u_cross(const uint32_t *__restrict__ off,
                                                  const uint32_t *__restrict__ idx,
                                                  const uint32_t *__restrict__ vgraph, uint32_t n,
                                                  uint32_t *__restrict__ bad) {
    uint32_t cnt = 0;
    const int lane = threadIdx.x & (kWave - 1);
    const uint32_t waves = gridDim.x * kWavesPerBlock;
    for (uint32_t s = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave; s < n; s += waves) {
        const uint32_t gs = vgraph[s];
        for (uint32_t j = off[s] + lane; j < off[s + 1]; j += kWave) cnt += vgraph[idx[j]] != gs;
    }
    if (cnt) atomicAdd(bad, cnt);
}

// Per-graph state reset; active graphs are the non-empty ones.
__global__ __launch_bounds__(kBlock) void u_reset(UArgs a) {
    for (uint32_t g = blockIdx.x * kBlock + threadIdx.x; g < a.ngraphs; g += gridDim.x * kBlock) {
        a.st.iter[g] = 0;
        a.st.delta[g] = 0.0;
        a.st.p_sink[g] = 0.0;
        a.st.scale[g] = 0.0;
        a.st.nsink[g] = 0;
        a.st.done[g] = a.voff[g + 1] == a.voff[g] ? 1u : 0u;
    }
}

// Forward start: x0 = pers = 1/n_g, c0 = x0 / deg.
template <bool WEIGHTED>
__global__ __launch_bounds__(kBlock) void u_fwd_init(UArgs a, const uint32_t *vgraph, uint32_t n) {
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const uint32_t g = vgraph[i];
        const float x0 = static_cast<float>(1.0 / (double)(a.voff[g + 1] - a.voff[g]));
        a.x_next[i] = x0;
        const double dg =
            WEIGHTED ? a.deg[i] : static_cast<double>(a.deg_off[i + 1] - a.deg_off[i]);
        a.c_next[i] = dg != 0.0 ? static_cast<float>(static_cast<double>(x0) / dg) : 0.f;
    }
}

// One workgroup per graph: out0[g] = sum f0(v), out1[g] = sum f1(v) over the
// graph's vertices (fixed order: thread-strided, then the workgroup tree).
template <typename F>
__global__ __launch_bounds__(kBlock) void u_graph_sums(const uint32_t *__restrict__ voff,
                                                       uint32_t G, F f, double *out0,
                                                       double *out1) {
    __shared__ double red[8];
    for (uint32_t g = blockIdx.x; g < G; g += gridDim.x) {
        double s0 = 0.0, s1 = 0.0;
        for (uint32_t v = voff[g] + threadIdx.x; v < voff[g + 1]; v += kBlock) f(g, v, s0, s1);
        block_sum2(s0, s1, red);
        if (threadIdx.x == 0) {
            out0[g] = s0;
            out1[g] = s1;
        }
    }
}

// The first dangling mass of each graph: (number of sinks) x x0, the sinks
// counted by the graph's own workgroup (no atomics on a shared counter).
struct FwdStartSinks {
    const double *deg;          // weighted out-degree, or null
    const uint32_t *deg_off;    // unity: push offsets
    const uint32_t *voff;
    __device__ void operator()(uint32_t g, uint32_t v, double &ps, double &) const {
        const double dg = deg ? deg[v] : static_cast<double>(deg_off[v + 1] - deg_off[v]);
        if (dg == 0.0)
            ps += static_cast<double>(static_cast<float>(1.0 / (double)(voff[g + 1] - voff[g])));
    }
};

// Adjoint start: lambda_0 = c, p . lambda_0 and ||c||_1 per graph.
struct AdjStartSums {
    const double *c;
    const uint32_t *voff;
    __device__ void operator()(uint32_t g, uint32_t v, double &pl, double &c1) const {
        const double pv = 1.0 / (double)(voff[g + 1] - voff[g]);
        pl += pv * static_cast<double>(static_cast<float>(c[v]));
        c1 += fabs(c[v]);
    }
};

__global__ __launch_bounds__(kBlock) void u_adj_init(UArgs a, uint32_t n, float *lam0) {
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
        lam0[i] = static_cast<float>(a.cvec[i]);
}

// out[v] = x[iter_g & 1][v]: every graph's latest iterate in one buffer.
template <typename O>
__global__ __launch_bounds__(kBlock) void u_latest(const float *x0, const float *x1,
                                                   const uint32_t *vgraph, const uint64_t *iter,
                                                   uint32_t n, O *out) {
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
        out[i] = static_cast<O>((iter[vgraph[i]] & 1) ? x1[i] : x0[i]);
}

// mu_s = (M^T lambda)_s over the push orientation (no per-graph state).
template <bool WEIGHTED>
struct UMuFin : TileAll {
    const UArgs &a;
    double *mu;
    __device__ UMuFin(const UArgs &args, double *m) : a(args), mu(m) {}
    __device__ __forceinline__ void operator()(uint32_t s, double acc) {
        const double dg =
            WEIGHTED ? a.deg[s] : static_cast<double>(a.deg_off[s + 1] - a.deg_off[s]);
        mu[s] = dg != 0.0 ? acc / dg : 0.0;
    }
};

template <bool WEIGHTED>
__global__ __launch_bounds__(kBlock) void u_mu(const UArgs a, double *mu) {
    __shared__ WaveLds<float> L;
    UMuFin<WEIGHTED> fin(a, mu);
    spmv_waves<float, WEIGHTED>(a.tiles, a.ntiles, a.off, a.packed, a.bits, a.tbase, a.src, a.w,
                                a.long_part, a.short_row, L, fin);
}

template <bool WEIGHTED>
__global__ __launch_bounds__(kBlock) void u_mu_long(const UArgs a, uint32_t nlong, double *mu) {
    UMuFin<WEIGHTED> fin(a, mu);
    for (uint32_t k = blockIdx.x * kBlock + threadIdx.x; k < nlong; k += gridDim.x * kBlock)
        fin(a.long_rows[k], long_row_sum(a.long_part, a.long_slot, k));
}

// dL/dw_e = d_g x_s / deg_s (lambda_t - mu_s) in edge-id order.  The per-
// source factors come packed (u_grad_coef): pair[s] = ((d_g x_s) / deg_s, mu_s),
// 0 and 0 for a source with deg_s = 0, so an edge gathers one 16-byte pair
// and lambda_t.  Each XCD takes a contiguous eighth of the edge ids, so the
// pairs of the graphs it covers stay in its L2.
__global__ __launch_bounds__(kBlock) void u_grad_coef(const UArgs a, const uint32_t *vgraph,
                                                      const double *mu, const double *x,
                                                      uint32_t n, double2 *pair) {
    for (uint32_t v = blockIdx.x * kBlock + threadIdx.x; v < n; v += gridDim.x * kBlock) {
        const double dg = a.deg ? a.deg[v] : static_cast<double>(a.deg_off[v + 1] - a.deg_off[v]);
        pair[v] = make_double2(dg != 0.0 ? a.damp[vgraph[v]] * x[v] / dg : 0.0, mu[v]);
    }
}

__global__ __launch_bounds__(kBlock) void u_grad_edges(const uint32_t *es, const uint32_t *et,
                                                       const double2 *pair, const float *lam,
                                                       uint64_t bound, float *grad) {
    uint64_t begin = 0, end = bound, first = (uint64_t)blockIdx.x * kBlock + threadIdx.x,
             stride = (uint64_t)gridDim.x * kBlock;
    if (gridDim.x % kXcds == 0) {
        const uint32_t x = blockIdx.x % kXcds;
        begin = bound * x / kXcds;
        end = bound * (x + 1) / kXcds;
        first = begin + (uint64_t)(blockIdx.x / kXcds) * kBlock + threadIdx.x;
        stride = (uint64_t)(gridDim.x / kXcds) * kBlock;
    }
    // The edge list is read once and the gradients written once, past L2, so
    // they do not evict the pairs and iterates the edges gather.
    for (uint64_t e = first; e < end; e += stride) {
        const uint32_t sv = __builtin_nontemporal_load(es + e);
        const uint32_t tv = __builtin_nontemporal_load(et + e);
        float out = 0.f;
        if (sv != kNoEdge) {
            const double2 p = pair[sv];
            out = static_cast<float>(p.x * (static_cast<double>(lam[tv]) - p.y));
        }
        __builtin_nontemporal_store(out, grad + e);
    }
}

// Per graph: the dangling mass of x*, and lambda . (x* - p).
template <bool WEIGHTED>
struct GradSums {
    const double *x;
    const float *lam;
    const double *deg;
    const uint32_t *deg_off;
    const uint32_t *voff;
    __device__ void operator()(uint32_t g, uint32_t v, double &sk, double &ld) const {
        const double dg = WEIGHTED ? deg[v] : static_cast<double>(deg_off[v + 1] - deg_off[v]);
        if (dg == 0.0) sk += x[v];
        const double pv = 1.0 / (double)(voff[g + 1] - voff[g]);
        ld += static_cast<double>(lam[v]) * (x[v] - pv);
    }
};

// dL/dp_v = ((1 - d_g) + d_g sk_g) lambda_v.
__global__ __launch_bounds__(kBlock) void u_grad_pers(const float *lam, const uint32_t *vgraph,
                                                      const double *damp, const double *sk,
                                                      uint32_t n, double *gp) {
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const uint32_t g = vgraph[i];
        gp[i] = ((1.0 - damp[g]) + damp[g] * sk[g]) * static_cast<double>(lam[i]);
    }
}

// ------------------------------------------------------------------ host

namespace {

unsigned ugrid(uint64_t work) {
    uint64_t b = (work + kBlock - 1) / kBlock;
    if (b > 8192) b = 8192;
    return b ? (unsigned)b : 1u;
}

// Workspace slots of the batch arrays.
enum { kUVoff, kUDamp, kUVg, kUState, kUTpart, kUBad, kUSums, kUList, kUListN, kUBMass,
       kUCount };
static_assert(kUCount <= (int)(sizeof(Workspace::ub) / sizeof(DevBuf)), "workspace slots");

// Per-graph loop state back to the start: every non-empty graph active, the
// bail word (right after the active count) clear.  Waits for the stream.
static_assert(offsetof(SweepState, launches) == offsetof(SweepState, counter) + 4,
              "the active count and the bail word are read into adjacent fields");
gt_status u_restart(gt_graph *g, uint32_t G, const uint32_t *voff, const UArgs &a) {
    hipStream_t s = g->stream;
    hipLaunchKernelGGL(u_reset, dim3(ugrid(G)), dim3(kBlock), 0, s, a);
    uint32_t live = 0;
    for (uint32_t k = 0; k < G; ++k) live += voff[k + 1] > voff[k];
    g->h_state->counter = live;
    g->h_state->launches = 0;
    GT_HIP(hipMemcpyAsync(a.active, &g->h_state->counter, 2 * sizeof(uint32_t),
                          hipMemcpyHostToDevice, s));
    GT_HIP(hipStreamSynchronize(s));
    return GT_OK;
}

// The device description of a batch over one orientation `o` (pull for the
// forward, push for the adjoint): validation, per-vertex graph ids, the
// schedule cut at graph bounds, per-tile graph and per-graph tile / split-row
// ranges, the per-graph state arrays.
gt_status u_setup(gt_graph *g, uint32_t G, const uint32_t *voff, const double *damping,
                  Orient &o, UArgs &a) {
    const uint32_t n = g->n;
    if (G == 0) return fail(GT_ERR_INVALID_ARG, "empty batch");
    if (!voff || !damping) return fail(GT_ERR_INVALID_ARG, "null batch description");
    if (voff[0] != 0 || voff[G] != n)
        return fail(GT_ERR_INVALID_ARG, "vertex offsets must run from 0 to the vertex bound");
    for (uint32_t k = 0; k < G; ++k) {
        if (voff[k + 1] < voff[k])
            return fail(GT_ERR_INVALID_ARG, "vertex offsets must not decrease");
        if (!(damping[k] >= 0.0 && damping[k] <= 1.0))
            return fail(GT_ERR_INVALID_ARG, "damping must lie in [0, 1]");
    }
    Workspace &ws = g->ws;
    hipStream_t s = g->stream;
    std::vector<uint32_t> bounds(voff, voff + G + 1);
    GT_TRY(ensure_schedule(g, o, n, &bounds));
    note_layout(g, o);
    const std::vector<Tile> &T = o.h_tiles;
    // Per-tile graph and per-graph tile / split-row ranges, computed and
    // uploaded once per schedule and set of bounds.
    if (o.batch_bounds != bounds) {
        const std::vector<uint32_t> &off = o.h_off;
        std::vector<uint32_t> tg(std::max<size_t>(1, T.size())), tb(G + 1, 0), lb(G + 1, 0);
        uint32_t gi = 0;
        for (size_t t = 0; t < T.size(); ++t) {
            uint32_t row = T[t].row_begin;
            if (T[t].row_end == kSplit)  // the row whose entries hold this chunk
                row = (uint32_t)(std::upper_bound(off.begin(), off.end(), T[t].nnz_begin) -
                                 off.begin()) - 1;
            while (gi + 1 < G && voff[gi + 1] <= row) ++gi;
            tg[t] = gi;
        }
        size_t t = 0;
        for (uint32_t k = 0; k <= G; ++k) {
            while (t < T.size() && tg[t] < k) ++t;
            tb[k] = (uint32_t)t;
        }
        tb[G] = (uint32_t)T.size();
        size_t l = 0;
        const std::vector<uint32_t> &LR = o.h_long_rows;
        for (uint32_t k = 0; k < G; ++k) {
            while (l < LR.size() && LR[l] < voff[k]) ++l;
            lb[k] = (uint32_t)l;
        }
        lb[G] = (uint32_t)LR.size();
        GT_TRY(o.batch[0].reserve(sizeof(uint32_t) * tg.size()));
        GT_TRY(o.batch[1].reserve(sizeof(uint32_t) * (G + 1)));
        GT_TRY(o.batch[2].reserve(sizeof(uint32_t) * (G + 1)));
        GT_HIP(hipMemcpy(o.batch[0].ptr, tg.data(), sizeof(uint32_t) * tg.size(),
                         hipMemcpyHostToDevice));
        GT_HIP(hipMemcpy(o.batch[1].ptr, tb.data(), sizeof(uint32_t) * (G + 1),
                         hipMemcpyHostToDevice));
        GT_HIP(hipMemcpy(o.batch[2].ptr, lb.data(), sizeof(uint32_t) * (G + 1),
                         hipMemcpyHostToDevice));
        o.batch_bounds = bounds;
    }
    DevBuf *U = ws.ub;
    GT_TRY(U[kUVoff].reserve(sizeof(uint32_t) * (G + 1)));
    GT_TRY(U[kUDamp].reserve(sizeof(double) * G));
    GT_TRY(U[kUVg].reserve(sizeof(uint32_t) * std::max<uint32_t>(1, n)));
    // state: p_sink, delta, scale (f64), iter (u64), done, nsink (u32), + active
    GT_TRY(U[kUState].reserve((3 * sizeof(double) + sizeof(uint64_t) + 2 * sizeof(uint32_t)) *
                                  (size_t)G + 64));
    GT_TRY(U[kUTpart].reserve(sizeof(double) * 2 * std::max<size_t>(1, T.size())));
    GT_TRY(U[kUBad].reserve(sizeof(uint32_t)));
    GT_TRY(U[kUSums].reserve(sizeof(double) * 2 * (size_t)G));
    GT_TRY(U[kUList].reserve(sizeof(uint32_t) * std::max<size_t>(1, T.size())));
    GT_TRY(U[kUListN].reserve(sizeof(uint32_t) * 2));
    GT_HIP(hipMemcpyAsync(U[kUVoff].ptr, voff, sizeof(uint32_t) * (G + 1), hipMemcpyHostToDevice, s));
    GT_HIP(hipMemcpyAsync(U[kUDamp].ptr, damping, sizeof(double) * G, hipMemcpyHostToDevice, s));
    char *st = U[kUState].as<char>();
    a.st.p_sink = reinterpret_cast<double *>(st);
    a.st.delta = a.st.p_sink + G;
    a.st.scale = a.st.delta + G;
    a.st.iter = reinterpret_cast<uint64_t *>(a.st.scale + G);
    a.st.done = reinterpret_cast<uint32_t *>(a.st.iter + G);
    a.st.nsink = a.st.done + G;
    a.active = a.st.nsink + G;
    a.ngraphs = G;
    a.voff = U[kUVoff].as<uint32_t>();
    a.damp = U[kUDamp].as<double>();
    a.tile_graph = o.batch[0].as<uint32_t>();
    a.tb = o.batch[1].as<uint32_t>();
    a.lb = o.batch[2].as<uint32_t>();
    a.tpart = U[kUTpart].as<double>();
    a.list = U[kUList].as<uint32_t>();
    a.list_count = U[kUListN].as<uint32_t>();
    a.tiles = o.tiles.as<Tile>();
    a.ntiles = o.ntiles;
    a.off = o.off.as<uint32_t>();
    a.packed = o.packed.as<uint32_t>();
    a.bits = o.bits;
    a.tbase = o.tbase.as<uint32_t>();
    a.short_row = (uint32_t)g->tune.short_row;
    a.long_rows = o.long_rows.as<uint32_t>();
    a.long_slot = o.long_slot.as<uint32_t>();
    GT_TRY(ws.long_part.reserve(sizeof(double) * std::max<uint32_t>(1, o.nslots)));
    a.long_part = ws.long_part.as<double>();
    hipLaunchKernelGGL(u_vgraph, dim3(ugrid(n)), dim3(kBlock), 0, s, a.voff, G, n,
                       U[kUVg].as<uint32_t>());
    // Reject an edge that joins two graphs (checked once per set of offsets:
    // the graph's edges do not change).
    if (g->batch_checked != bounds) {
        GT_HIP(hipMemsetAsync(U[kUBad].ptr, 0, sizeof(uint32_t), s));
        hipLaunchKernelGGL(u_cross, dim3(ugrid((uint64_t)n * kWave)), dim3(kBlock), 0, s,
                           g->out.off.as<uint32_t>(), g->out.idx.as<uint32_t>(),
                           U[kUVg].as<uint32_t>(), n, U[kUBad].as<uint32_t>());
        uint32_t bad = 0;
        GT_HIP(hipMemcpyAsync(&g->h_state->counter, U[kUBad].ptr, sizeof(uint32_t),
                              hipMemcpyDeviceToHost, s));
        GT_HIP(hipStreamSynchronize(s));
        bad = g->h_state->counter;
        if (bad) return fail(GT_ERR_INVALID_ARG, "an edge joins two graphs of the batch");
        g->batch_checked = bounds;
    }
    GT_TRY(u_restart(g, G, voff, a));
    GT_HIP(hipGetLastError());
    return GT_OK;
}

// Sweep/fold pairs until no graph is active: kUnevenCapture pairs captured in
// a HIP graph on the graph's own stream (never the legacy default stream),
// replayed.  x[par] / x[par ^ 1] alternate, so a graph's latest iterate is
// x[iter_g & 1].  `sweep(args, k)` and `fold(args)` launch one pair on `cs`.
// Ends early, with *bailed set, when a block sweep's fold raised the bail
// word (the word after the active count).
template <typename Sweep, typename Fold>
gt_status u_iterate(gt_graph *g, UArgs a, const float *xs[2], float *xn[2], float *cn[2],
                    const float *srcs[2], uint64_t max_iter, Sweep sweep, Fold fold,
                    uint64_t *replays_out, bool *bailed = nullptr) {
    hipStream_t cs = g->own_stream;
    if (g->stream != cs) GT_HIP(hipStreamSynchronize(g->stream));
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipLaunchKernelGGL(u_list_init, dim3(ugrid(std::max<uint32_t>(1, a.ntiles))), dim3(kBlock), 0,
                       cs, a);
    GT_HIP(hipGetLastError());
    GT_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    static_assert(kUnevenCapture % 2 == 0, "a replay starts at parity 0");
    for (int k = 0; k < kUnevenCapture; ++k) {
        const int par = k & 1;
        UArgs b = a;
        b.list_n = a.list_count + par;
        b.list_n_next = a.list_count + (par ^ 1);
        b.x_cur = xs[par];
        b.x_next = xn[par ^ 1];
        b.c_next = cn ? cn[par ^ 1] : nullptr;
        b.src = srcs[par];
        sweep(b, cs);
        fold(b, cs);
    }
    hipError_t e = hipStreamEndCapture(cs, &graph);
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(graph);
        return hip_fail(e, "hipGraphInstantiate", __FILE__, __LINE__);
    }
    // Each replay advances every active graph by kUnevenCapture sweeps; with a
    // cap, every graph is done after ceil(max_iter / kUnevenCapture) replays.
    const uint64_t limit = max_iter ? (max_iter + kUnevenCapture - 1) / kUnevenCapture
                                    : (uint64_t)1 << 20;
    uint64_t r = 0;
    gt_status st = GT_OK;
    for (;;) {
        e = hipGraphLaunch(exec, cs);
        if (e == hipSuccess)
            e = hipMemcpyAsync(&g->h_state->counter, a.active, 2 * sizeof(uint32_t),
                               hipMemcpyDeviceToHost, cs);
        if (e == hipSuccess) e = hipStreamSynchronize(cs);
        if (e != hipSuccess) {
            st = hip_fail(e, "uneven replay", __FILE__, __LINE__);
            break;
        }
        ++r;
        if (g->h_state->launches != 0) {
            if (bailed) *bailed = true;
            break;
        }
        if (g->h_state->counter == 0) break;
        if (r >= limit) {
            st = fail(GT_ERR_INVALID_ARG, "uneven batch: no convergence within 2^23 sweeps");
            break;
        }
    }
    (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
    *replays_out = r;
    return st;
}

// The weight map's device source and the weighted out-degree (push: the
// out-halves); the sweep's weights are staged for the layout that runs.
struct UWeights {
    bool on = false;
    const void *src = nullptr;
    int dtype = GT_F64;
};

gt_status u_weights(gt_graph *g, const gt_prop *weight, Orient &push, UArgs &a, UWeights *w) {
    w->on = weight && weight->kind != GT_MAP_UNITY;
    a.deg_off = push.off.as<uint32_t>();
    if (!w->on) return GT_OK;
    GT_TRY(stage_edge_source(g, weight, &w->src, &w->dtype));
    GT_TRY(stage_degree(g, weight, push, w->src, w->dtype, g->ws.deg));
    a.deg = g->ws.deg.as<double>();
    return GT_OK;
}

// The wave engine's weights: orientation o's in its packed order.
gt_status u_wave_weights(gt_graph *g, const gt_prop *weight, Orient &o, const UWeights &w,
                         UArgs &a, DevBuf &wbuf) {
    if (!w.on || a.w) return GT_OK;
    GT_TRY(stage_pull_weights<float>(g, weight, o, w.src, w.dtype, wbuf));
    a.w = wbuf.as<float>();
    return GT_OK;
}

// The batch on block tiles of orientation o cut at the graphs' bounds
// (`gathered`: device offsets whose row lengths count how often each vertex
// is gathered), unless the tuning turns them off: the per-tile graph, each
// graph's first tile and first long row, the weights in block order, and
// the sweep arguments.  *ok is false when the batch runs on the wave engine:
// block tiles off, or a weight negative or not finite.
gt_status u_blk_setup(gt_graph *g, uint32_t G, const uint32_t *voff, Orient &o,
                      const uint32_t *gathered, const gt_prop *weight, const UWeights &w,
                      UArgs &a, BlkArgs &bk, bool *ok) {
    *ok = false;
    if (g->tune.block_tile_nnz <= 0) return GT_OK;
    Workspace &ws = g->ws;
    const std::vector<uint32_t> bounds(voff, voff + G + 1);
    Blocks *B = nullptr;
    GT_TRY(ensure_blocks_cut(g, o, gathered, (uint32_t)g->tune.block_tile_nnz,
                             (uint32_t)g->tune.block_long_min, bounds, &B));
    if (w.on) {
        bool usable = false;
        GT_TRY(blk_stage_weights(g, *B, weight, w.src, w.dtype, false, ws.blk_w, &usable));
        if (!usable) return GT_OK;
    }
    GT_TRY(blk_reserve(g, *B));
    const uint32_t nt = B->ntiles;
    if (B->batch_bounds != bounds) {
        std::vector<uint32_t> tg(std::max<uint32_t>(1, nt), 0), tb(G + 1), lb(G + 1);
        for (uint32_t t = 0; t < nt; ++t)
            tg[t] = (uint32_t)(std::upper_bound(voff, voff + G + 1, B->h_tile_row[t]) - voff) - 1;
        const std::vector<uint32_t> &LR = B->h_long_rows;
        for (uint32_t k = 0; k <= G; ++k) {
            tb[k] = (uint32_t)(std::lower_bound(tg.begin(), tg.begin() + nt, k) - tg.begin());
            lb[k] = (uint32_t)(std::lower_bound(LR.begin(), LR.end(), voff[k]) - LR.begin());
        }
        GT_TRY(B->batch[0].reserve(sizeof(uint32_t) * tg.size()));
        GT_TRY(B->batch[1].reserve(sizeof(uint32_t) * (G + 1)));
        GT_TRY(B->batch[2].reserve(sizeof(uint32_t) * (G + 1)));
        GT_HIP(hipMemcpy(B->batch[0].ptr, tg.data(), sizeof(uint32_t) * tg.size(),
                         hipMemcpyHostToDevice));
        GT_HIP(hipMemcpy(B->batch[1].ptr, tb.data(), sizeof(uint32_t) * (G + 1),
                         hipMemcpyHostToDevice));
        GT_HIP(hipMemcpy(B->batch[2].ptr, lb.data(), sizeof(uint32_t) * (G + 1),
                         hipMemcpyHostToDevice));
        B->batch_bounds = bounds;
    }
    DevBuf *U = ws.ub;
    GT_TRY(U[kUBMass].reserve(sizeof(double) * G + 2 * sizeof(uint32_t)));
    if (U[kUList].bytes < sizeof(uint32_t) * std::max<uint32_t>(1, nt)) {
        GT_TRY(U[kUList].reserve(sizeof(uint32_t) * std::max<uint32_t>(1, nt)));
        a.list = U[kUList].as<uint32_t>();
    }
    bk = blk_layout_args(g, *B);
    BlkUneven &u = bk.u;
    u.tile_graph = B->batch[0].as<uint32_t>();
    u.voff = a.voff;
    u.damp = a.damp;
    u.mass = U[kUBMass].as<double>();
    u.ps = a.st.p_sink;
    u.delta = a.st.delta;
    u.scale = a.st.scale;
    u.iter = a.st.iter;
    u.done = a.st.done;
    u.active = a.active;
    u.tb = B->batch[1].as<uint32_t>();
    u.lb = B->batch[2].as<uint32_t>();
    u.list = a.list;
    u.ngraphs = G;
    // The two tile counters, one per sweep parity, after the masses.
    u.ctr = reinterpret_cast<uint32_t *>(u.mass + G);
    bk.eps_abs = a.eps;
    bk.eps_rel = a.eps_rel;
    bk.max_iter = a.max_iter;
    bk.cvec = a.cvec;
    bk.row_off = a.deg_off;
    if (w.on) {
        bk.wblk = ws.blk_w.ptr;
        bk.wdeg = a.deg;
    }
    *ok = true;
    return GT_OK;
}

// Runs the batch's loop on block tiles (`mode` kBlkUFwd / kBlkUAdj) from the
// start iterate the caller wrote into xs[0]; *bailed when a graph's mass
// left the fixed point's range.
gt_status u_blk_iterate(gt_graph *g, const UArgs &a, BlkArgs bk, int mode, const float *xs[2],
                        float *xn[2], float *cn[2], const float *srcs[2], uint64_t max_iter,
                        bool gathered, bool *bailed) {
    hipStream_t s = g->stream;
    blk_launch_ustart(bk, xs[0], gathered, s);
    GT_HIP(hipMemsetAsync(bk.u.ctr, 0, 2 * sizeof(uint32_t), s));
    GT_HIP(hipGetLastError());
    UArgs ab = a;
    ab.ntiles = bk.ntiles;
    uint32_t *ctr = bk.u.ctr;
    const int per_cu = g->tune.block_per_cu;
    auto at = [&](const UArgs &b) {
        BlkArgs k = bk;
        const int par = (int)(b.list_n - a.list_count);
        k.x_cur = b.x_cur;
        k.x_next = b.x_next;
        k.c_cur = b.src;
        k.c_next = b.c_next;
        k.u.list_n = b.list_n;
        k.u.list_n_next = b.list_n_next;
        k.u.ctr = ctr + par;
        k.u.ctr_next = ctr + (par ^ 1);
        return k;
    };
    uint64_t replays = 0;
    *bailed = false;
    return u_iterate(
        g, ab, xs, xn, cn, srcs, max_iter,
        [&](const UArgs &b, hipStream_t cs) { blk_launch_usweep(at(b), per_cu, mode, cs); },
        [&](const UArgs &b, hipStream_t cs) { blk_launch_ufold(at(b), mode, cs); }, &replays,
        bailed);
}

gt_status u_params(const gt_power_iteration *params, uint32_t n) {
    if (!params) return fail(GT_ERR_INVALID_ARG, "params is null");
    if (params->plan.len != n)
        return fail(GT_ERR_PLAN_COVER, "sweep: the plan must cover the output exactly");
    if (params->max_iter == 0 && !(params->epsilon > 0.0))
        return fail(GT_ERR_INVALID_ARG,
                    "epsilon must be positive when max_iter is 0: the loop would never end");
    return GT_OK;
}

}  // namespace

gt_status pagerank_uneven_impl(gt_graph *g, uint32_t G, const uint32_t *voff,
                               const double *damping, const gt_prop *weight,
                               const gt_power_iteration *params, double *rank, int32_t rank_loc,
                               uint64_t *iters) {
    clear_error();
    if (!g) return fail(GT_ERR_INVALID_ARG, "null graph");
    if (!iters || !rank) return fail(GT_ERR_INVALID_ARG, "null output");
    if (rank_loc != GT_HOST && rank_loc != GT_DEVICE)
        return fail(GT_ERR_INVALID_ARG, "bad location");
    GT_TRY(u_params(params, g->n));
    const uint32_t n = g->n;
    Workspace &ws = g->ws;
    hipStream_t s = g->stream;
    UArgs a{};
    GT_TRY(u_setup(g, G, voff, damping, g->in, a));
    UWeights wt;
    GT_TRY(u_weights(g, weight, g->out, a, &wt));
    const bool weighted = wt.on;
    // The FP32 change of a unit-mass iterate bottoms out near 2^-22.
    a.eps = std::max(params->epsilon, std::ldexp(1.0, -22));
    a.max_iter = params->max_iter;
    for (int k = 0; k < 2; ++k) {
        GT_TRY(ws.x32[k].reserve(sizeof(float) * std::max<uint32_t>(1, n)));
        GT_TRY(ws.c32[k].reserve(sizeof(float) * std::max<uint32_t>(1, n)));
    }
    float *x[2] = {ws.x32[0].as<float>(), ws.x32[1].as<float>()};
    float *c[2] = {ws.c32[0].as<float>(), ws.c32[1].as<float>()};
    const uint32_t *vg = ws.ub[kUVg].as<uint32_t>();
    // x0 = 1/n_g, c0 = x0 / deg, and each graph's first dangling mass.
    auto init = [&]() -> gt_status {
        UArgs b = a;
        b.x_next = x[0];
        b.c_next = c[0];
        if (weighted)
            hipLaunchKernelGGL(u_fwd_init<true>, dim3(ugrid(n)), dim3(kBlock), 0, s, b, vg, n);
        else
            hipLaunchKernelGGL(u_fwd_init<false>, dim3(ugrid(n)), dim3(kBlock), 0, s, b, vg, n);
        const unsigned gg = std::min<uint32_t>(G, 65535);
        hipLaunchKernelGGL(u_graph_sums<FwdStartSinks>, dim3(gg), dim3(kBlock), 0, s, a.voff, G,
                           FwdStartSinks{weighted ? a.deg : nullptr, a.deg_off, a.voff},
                           a.st.p_sink, a.st.scale);
        GT_HIP(hipGetLastError());
        return GT_OK;
    };
    const float *xs[2] = {x[0], x[1]};
    const float *srcs[2] = {c[0], c[1]};
    // Block tiles of the in-halves, cut at the graphs' bounds; the wave
    // engine when they are off or cannot hold the batch.
    bool on_blocks = false;
    {
        BlkArgs bk{};
        bool ok = false;
        GT_TRY(u_blk_setup(g, G, voff, g->in, g->out.off.as<uint32_t>(), weight, wt, a, bk, &ok));
        if (ok) {
            GT_TRY(init());
            bool bailed = false;
            GT_TRY(u_blk_iterate(g, a, bk, kBlkUFwd, xs, x, c, srcs, params->max_iter, false,
                                 &bailed));
            on_blocks = !bailed;
            if (bailed) GT_TRY(u_restart(g, G, voff, a));
        }
    }
    if (!on_blocks) {
        note_layout(g, g->in);
        GT_TRY(u_wave_weights(g, weight, g->in, wt, a, ws.wcsc));
        GT_TRY(init());
        const int cap_w = occupancy_cap(
            (const void *)u_sweep<UFwdFin<true>, true>, g->tune.sweep_blocks_per_cu);
        const int cap_u = occupancy_cap(
            (const void *)u_sweep<UFwdFin<false>, false>, g->tune.sweep_blocks_per_cu);
        const dim3 sg(wave_grid(weighted ? cap_w : cap_u, a.ntiles)),
            fg((unsigned)std::max<uint32_t>(1, (G + kWavesPerBlock - 1) / kWavesPerBlock)),
            bd(kBlock);
        uint64_t replays = 0;
        GT_TRY(u_iterate(
            g, a, xs, x, c, srcs, params->max_iter,
            [&](const UArgs &b, hipStream_t cs) {
                if (weighted) hipLaunchKernelGGL((u_sweep<UFwdFin<true>, true>), sg, bd, 0, cs, b);
                else hipLaunchKernelGGL((u_sweep<UFwdFin<false>, false>), sg, bd, 0, cs, b);
            },
            [&](const UArgs &b, hipStream_t cs) {
                if (weighted) hipLaunchKernelGGL((u_fold<UFwdFin<true>, false>), fg, bd, 0, cs, b);
                else hipLaunchKernelGGL((u_fold<UFwdFin<false>, false>), fg, bd, 0, cs, b);
            },
            &replays));
    }
    hipStream_t cs = g->own_stream;
    double *out = rank;
    DevBuf host;
    if (rank_loc == GT_HOST) {
        GT_TRY(ws.y3.reserve(sizeof(double) * std::max<uint32_t>(1, n)));
        out = ws.y3.as<double>();
    }
    hipLaunchKernelGGL(u_latest<double>, dim3(ugrid(n)), dim3(kBlock), 0, cs, x[0], x[1], vg,
                       a.st.iter, n, out);
    GT_HIP(hipGetLastError());
    if (rank_loc == GT_HOST)
        GT_HIP(hipMemcpyAsync(rank, out, sizeof(double) * n, hipMemcpyDeviceToHost, cs));
    GT_HIP(hipMemcpyAsync(iters, a.st.iter, sizeof(uint64_t) * G, hipMemcpyDeviceToHost, cs));
    GT_HIP(hipStreamSynchronize(cs));
    uint64_t total = 0;
    for (uint32_t k = 0; k < G; ++k) total += iters[k];
    g->last_sweeps = total;
    return GT_OK;
}

gt_status pagerank_uneven_backward_impl(gt_graph *g, uint32_t G, const uint32_t *voff,
                                        const double *damping, const gt_prop *weight,
                                        const gt_power_iteration *params, const double *x,
                                        const double *dl_dx, int32_t loc, float *grad_w,
                                        double *grad_pers, double *grad_damping,
                                        uint64_t *iters) {
    clear_error();
    if (!g) return fail(GT_ERR_INVALID_ARG, "null graph");
    if (!x || !dl_dx || !iters) return fail(GT_ERR_INVALID_ARG, "null argument");
    if (loc != GT_HOST && loc != GT_DEVICE) return fail(GT_ERR_INVALID_ARG, "bad location");
    GT_TRY(u_params(params, g->n));
    const uint32_t n = g->n;
    Workspace &ws = g->ws;
    hipStream_t s = g->stream;
    UArgs a{};
    GT_TRY(u_setup(g, G, voff, damping, g->out, a));
    UWeights wt;
    GT_TRY(u_weights(g, weight, g->out, a, &wt));
    const bool weighted = wt.on;
    a.eps = params->epsilon;
    a.eps_rel = std::ldexp(1.0, -22);
    a.max_iter = params->max_iter;
    // x and dL/dx as device f64.
    const double *xd = x, *cd = dl_dx;
    if (loc == GT_HOST) {
        GT_TRY(ws.ax.reserve(sizeof(double) * std::max<uint32_t>(1, n)));
        GT_TRY(ws.ac.reserve(sizeof(double) * std::max<uint32_t>(1, n)));
        GT_HIP(hipMemcpyAsync(ws.ax.ptr, x, sizeof(double) * n, hipMemcpyHostToDevice, s));
        GT_HIP(hipMemcpyAsync(ws.ac.ptr, dl_dx, sizeof(double) * n, hipMemcpyHostToDevice, s));
        xd = ws.ax.as<double>();
        cd = ws.ac.as<double>();
    }
    a.cvec = cd;
    for (int k = 0; k < 2; ++k) GT_TRY(ws.lam[k].reserve(sizeof(float) * std::max<uint32_t>(1, n)));
    float *lam[2] = {ws.lam[0].as<float>(), ws.lam[1].as<float>()};
    const uint32_t *vg = ws.ub[kUVg].as<uint32_t>();
    double *sums = ws.ub[kUSums].as<double>();
    const unsigned gg = std::min<uint32_t>(G, 65535);
    // lambda_0 = dL/dx, and per graph p . lambda_0 and ||dL/dx_g||_1.
    auto init = [&]() -> gt_status {
        hipLaunchKernelGGL(u_adj_init, dim3(ugrid(n)), dim3(kBlock), 0, s, a, n, lam[0]);
        hipLaunchKernelGGL(u_graph_sums<AdjStartSums>, dim3(gg), dim3(kBlock), 0, s, a.voff, G,
                           AdjStartSums{cd, a.voff}, a.st.p_sink, a.st.scale);
        GT_HIP(hipGetLastError());
        return GT_OK;
    };
    const int cap_w = occupancy_cap(
        (const void *)u_sweep<UAdjFin<true>, true>, g->tune.sweep_blocks_per_cu);
    const int cap_u = occupancy_cap(
        (const void *)u_sweep<UAdjFin<false>, false>, g->tune.sweep_blocks_per_cu);
    const dim3 sg(wave_grid(weighted ? cap_w : cap_u, a.ntiles)),
        fg((unsigned)std::max<uint32_t>(1, (G + kWavesPerBlock - 1) / kWavesPerBlock)),
        bd(kBlock);
    const float *ls[2] = {lam[0], lam[1]};
    // Block tiles of the out-halves, cut at the graphs' bounds (a vertex is
    // gathered once per in-edge); the wave engine when they are off or
    // cannot hold the batch.
    bool on_blocks = false;
    BlkArgs bk{};
    {
        bool ok = false;
        GT_TRY(u_blk_setup(g, G, voff, g->out, g->in.off.as<uint32_t>(), weight, wt, a, bk, &ok));
        if (ok) {
            GT_TRY(init());
            bool bailed = false;
            GT_TRY(u_blk_iterate(g, a, bk, kBlkUAdj, ls, lam, nullptr, ls, params->max_iter, true,
                                 &bailed));
            on_blocks = !bailed;
            if (bailed) GT_TRY(u_restart(g, G, voff, a));
        }
    }
    if (!on_blocks) {
        note_layout(g, g->out);
        GT_TRY(u_wave_weights(g, weight, g->out, wt, a, ws.wcsr));
        GT_TRY(init());
        uint64_t replays = 0;
        GT_TRY(u_iterate(
            g, a, ls, lam, nullptr, ls, params->max_iter,
            [&](const UArgs &b, hipStream_t cs) {
                if (weighted) hipLaunchKernelGGL((u_sweep<UAdjFin<true>, true>), sg, bd, 0, cs, b);
                else hipLaunchKernelGGL((u_sweep<UAdjFin<false>, false>), sg, bd, 0, cs, b);
            },
            [&](const UArgs &b, hipStream_t cs) {
                if (weighted) hipLaunchKernelGGL((u_fold<UAdjFin<true>, true>), fg, bd, 0, cs, b);
                else hipLaunchKernelGGL((u_fold<UAdjFin<false>, true>), fg, bd, 0, cs, b);
            },
            &replays));
    }
    hipStream_t cs = g->own_stream;
    // Every graph's final lambda in one buffer (ws.y3 as f32 storage).
    GT_TRY(ws.y3.reserve(sizeof(double) * std::max<uint32_t>(1, n)));
    float *lf = reinterpret_cast<float *>(ws.y3.ptr);
    hipLaunchKernelGGL(u_latest<float>, dim3(ugrid(n)), dim3(kBlock), 0, cs, lam[0], lam[1], vg,
                       a.st.iter, n, lf);
    if (grad_w) {
        GT_TRY(ensure_edge_list(g));
        GT_TRY(ws.mu.reserve(sizeof(double) * std::max<uint32_t>(1, n)));
        GT_TRY(ws.arl[3].reserve(sizeof(double2) * std::max<uint32_t>(1, n)));
        double *mu = ws.mu.as<double>();
        double2 *pair = ws.arl[3].as<double2>();
        if (on_blocks) {
            // mu = M^T lambda on the loop's block tiles: one more sweep over
            // every tile at each graph's last scale, rows written as acc / deg.
            BlkArgs k = bk;
            k.mu_out = mu;
            k.x_cur = lf;
            k.c_cur = lf;
            GT_HIP(hipMemsetAsync(k.u.ctr, 0, sizeof(uint32_t), cs));
            blk_launch_usweep(k, g->tune.block_per_cu, kBlkUAdj, cs);
            blk_launch_ufold(k, kBlkUAdj, cs);
        } else {
            UArgs b = a;
            b.src = lf;
            const dim3 lg(fold_grid(g->out.nlong));
            if (weighted) {
                hipLaunchKernelGGL(u_mu<true>, sg, bd, 0, cs, b, mu);
                hipLaunchKernelGGL(u_mu_long<true>, lg, bd, 0, cs, b, g->out.nlong, mu);
            } else {
                hipLaunchKernelGGL(u_mu<false>, sg, bd, 0, cs, b, mu);
                hipLaunchKernelGGL(u_mu_long<false>, lg, bd, 0, cs, b, g->out.nlong, mu);
            }
        }
        const uint64_t eb = g->edge_bound;
        float *gdst = grad_w;
        if (loc == GT_HOST) {
            GT_TRY(ws.gw.reserve(sizeof(float) * std::max<uint64_t>(1, eb)));
            gdst = ws.gw.as<float>();
        }
        const dim3 eg((unsigned)std::max<uint64_t>(
            1, std::min<uint64_t>(16384, (eb + kBlock - 1) / kBlock)));
        UArgs c = a;
        c.deg = weighted ? a.deg : nullptr;
        hipLaunchKernelGGL(u_grad_coef, dim3(ugrid(n)), bd, 0, cs, c, vg, mu, xd, n, pair);
        hipLaunchKernelGGL(u_grad_edges, eg, bd, 0, cs, g->esrc.as<uint32_t>(),
                           g->etgt.as<uint32_t>(), pair, lf, eb, gdst);
        if (loc == GT_HOST && eb)
            GT_HIP(hipMemcpyAsync(grad_w, gdst, sizeof(float) * eb, hipMemcpyDeviceToHost, cs));
    }
    // Per graph: sk = dangling mass of x*, ld = lambda . (x* - p).
    double *sk = sums, *ld = sums + G;
    if (weighted)
        hipLaunchKernelGGL(u_graph_sums<GradSums<true>>, dim3(gg), dim3(kBlock), 0, cs, a.voff, G,
                           GradSums<true>{xd, lf, a.deg, a.deg_off, a.voff}, sk, ld);
    else
        hipLaunchKernelGGL(u_graph_sums<GradSums<false>>, dim3(gg), dim3(kBlock), 0, cs, a.voff, G,
                           GradSums<false>{xd, lf, a.deg, a.deg_off, a.voff}, sk, ld);
    if (grad_pers) {
        double *gp = grad_pers;
        if (loc == GT_HOST) {
            GT_TRY(ws.gp.reserve(sizeof(double) * std::max<uint32_t>(1, n)));
            gp = ws.gp.as<double>();
        }
        hipLaunchKernelGGL(u_grad_pers, dim3(ugrid(n)), dim3(kBlock), 0, cs, lf, vg, a.damp, sk, n,
                           gp);
        if (loc == GT_HOST)
            GT_HIP(hipMemcpyAsync(grad_pers, gp, sizeof(double) * n, hipMemcpyDeviceToHost, cs));
    }
    GT_HIP(hipGetLastError());
    std::vector<double> hld(G);
    GT_HIP(hipMemcpyAsync(hld.data(), ld, sizeof(double) * G, hipMemcpyDeviceToHost, cs));
    GT_HIP(hipMemcpyAsync(iters, a.st.iter, sizeof(uint64_t) * G, hipMemcpyDeviceToHost, cs));
    GT_HIP(hipStreamSynchronize(cs));
    if (grad_damping)
        for (uint32_t k = 0; k < G; ++k) grad_damping[k] = damping[k] != 0.0 ? hld[k] / damping[k] : 0.0;
    return GT_OK;
}

}  // namespace gt

using namespace gt;

extern "C" {

gt_status gt_pagerank_uneven(gt_graph *g, uint32_t num_graphs, const uint32_t *vertex_offsets,
                             const double *damping, const gt_prop *weight,
                             const gt_power_iteration *params, double *rank, int32_t rank_loc,
                             uint64_t *iters) {
    return pagerank_uneven_impl(g, num_graphs, vertex_offsets, damping, weight, params, rank,
                                rank_loc, iters);
}

gt_status gt_pagerank_uneven_backward(gt_graph *g, uint32_t num_graphs,
                                      const uint32_t *vertex_offsets, const double *damping,
                                      const gt_prop *weight, const gt_power_iteration *params,
                                      const double *x, const double *dl_dx, int32_t loc,
                                      float *grad_w, double *grad_pers, double *grad_damping,
                                      uint64_t *iters) {
    return pagerank_uneven_backward_impl(g, num_graphs, vertex_offsets, damping, weight, params,
                                         x, dl_dx, loc, grad_w, grad_pers, grad_damping, iters);
}

}  // extern "C"
