// PageRank power iteration on gfx950.
//
// Semantics follow crates/gt-algo/src/centrality.rs:261-384 (pagerank_with):
// per active vertex i, r = p_sink * pers(i) + sum over in-entries (s, e) with
// deg(s) != 0 of prev(s) * w(e) / deg(s); next(i) = (1-d) pers(i) + d r; the
// sweep's L1 change decides convergence; inactive slots are carried; sinks
// (deg == 0) feed p_sink; an odd sweep count leaves the latest iterate in the
// caller's map (:378-382).
//
// Data layout per sweep (pull form, one launch per sweep):
//   in_off  u32[n+1]                  row offsets (LDS positions in a tile)
//   packed  u32[m]                    tile-sorted (source << bits | slot), streamed once
//   c_cur   V[n]                      gathered: c(s) = x(s)/deg(s), 0 for sinks
//   x_cur   V[n] read, x_next V[n] written, c_next V[n] written
//   out_off u32[n+1]                  unity out-degree (weighted: deg f64[n])
// The wave engine (sweep_common.hpp) does the gather and row sums.  Rows longer
// than a tile are split; a second, small kernel (pr_fold) combines their
// partial sums in chunk order, folds the per-workgroup delta and
// dangling-mass partials in index order and decides convergence on the
// device.  Launches after convergence exit at once.
#include <cmath>
#include <mutex>
#include <type_traits>

#include "blocks.hpp"
#include "device_common.hpp"
#include "gt_internal.hpp"
#include "props.hpp"
#include "driver.hpp"
#include "sweep_common.hpp"

namespace gt {

struct PrArgs {
    const Tile *tiles;
    uint32_t ntiles;
    uint32_t nlong;
    const uint32_t *in_off;
    const uint32_t *packed;   // packed entry stream of the wave tiles
    uint32_t bits;            // slot bits of `packed`
    const uint32_t *tbase;    // source base of each tile
    uint32_t short_row;        // rows with fewer entries are summed by one lane
    const void *w;            // weights in packed order (value type), weighted only
    const uint32_t *out_off;  // unity out-degree source
    const double *deg;        // weighted out-degree
    const uint8_t *active;    // masked views only
    const uint32_t *iperm;    // sweep label -> caller's label (null: identity)
    const double *pers;       // dense personalisation, in sweep labels
    const double *pers_c;     // the same in the caller's labels (seeding the map)
    double pval;              // constant personalisation
    double damping;
    double eps_abs;           // stop when delta < max(eps_abs, eps_rel * mass)
    double eps_rel;
    uint64_t max_iter;        // cap on st->iter; 0 = none
    const void *x_cur;
    void *x_next;
    const void *c_cur;
    void *c_next;
    double *partials;         // kPartials: sweep partials, then fold partials
    double *long_part;        // [nslots]
    const uint32_t *long_rows;
    const uint32_t *long_slot;
    const uint32_t *long_idx;  // null: a split row's slots are its partials
    SweepState *st;
};

template <typename V, bool WEIGHTED, bool DPERS, bool MASK>
struct PrFin : TileAll {
    static constexpr bool kWeighted = WEIGHTED;
    const PrArgs &a;
    double p_sink;
    double dsum = 0.0;
    double ssum = 0.0;
    __device__ PrFin(const PrArgs &args, double ps) : a(args), p_sink(ps) {}
    // What row i reads: its iterate, out-degree and personalisation.
    static constexpr bool kLoadsFirst = false;
    struct In {
        V xc;
        double dg;
        double pv;
        bool act;
    };
    __device__ __forceinline__ In load(uint32_t i) const {
        In in;
        in.xc = static_cast<const V *>(a.x_cur)[i];
        in.dg = WEIGHTED ? a.deg[i] : static_cast<double>(a.out_off[i + 1] - a.out_off[i]);
        in.pv = DPERS ? a.pers[i] : a.pval;
        in.act = MASK ? a.active[i] != 0 : true;
        return in;
    }
    __device__ __forceinline__ void finish(uint32_t i, double acc, const In &in) {
        V *xn = static_cast<V *>(a.x_next);
        V *cn = static_cast<V *>(a.c_next);
        if (MASK && !in.act) {
            xn[i] = in.xc;
            cn[i] = V(0);
            return;
        }
        const double pv = in.pv;
        const double nv = (1.0 - a.damping) * pv + a.damping * (p_sink * pv + acc);
        const V nvv = static_cast<V>(nv);
        dsum += fabs(static_cast<double>(nvv) - static_cast<double>(in.xc));
        xn[i] = nvv;
        if (in.dg != 0.0) {
            cn[i] = static_cast<V>(static_cast<double>(nvv) / in.dg);
        } else {
            cn[i] = V(0);
            ssum += static_cast<double>(nvv);
        }
    }
    __device__ __forceinline__ void operator()(uint32_t i, double acc) { finish(i, acc, load(i)); }
};

// The last sweep of the FP32 form, in FP64: the FP32 phase's final
// contributions c = x / deg (FP32) are gathered, widened and summed in FP64,
// the teleport update is FP64 and the new iterate FP64; delta is taken
// against the FP32 iterate.  No sweep follows, so no contribution or dangling
// mass is written.
template <bool WEIGHTED, bool DPERS, bool MASK>
struct RefineFin : TileAll {
    static constexpr bool kWeighted = WEIGHTED;
    const PrArgs &a;
    double p_sink;
    double dsum = 0.0;
    double ssum = 0.0;
    __device__ RefineFin(const PrArgs &args, double ps) : a(args), p_sink(ps) {}
    static constexpr bool kLoadsFirst = false;
    struct In {
        float xc;
        double pv;
        bool act;
    };
    __device__ __forceinline__ In load(uint32_t i) const {
        In in;
        in.xc = static_cast<const float *>(a.x_cur)[i];
        in.pv = DPERS ? a.pers[i] : a.pval;
        in.act = MASK ? a.active[i] != 0 : true;
        return in;
    }
    __device__ __forceinline__ void finish(uint32_t i, double acc, const In &in) {
        double *xn = static_cast<double *>(a.x_next);
        if (MASK && !in.act) {
            xn[i] = static_cast<double>(in.xc);
            return;
        }
        const double nv = (1.0 - a.damping) * in.pv + a.damping * (p_sink * in.pv + acc);
        dsum += fabs(nv - static_cast<double>(in.xc));
        xn[i] = nv;
    }
    __device__ __forceinline__ void operator()(uint32_t i, double acc) { finish(i, acc, load(i)); }
};

// The SpMV of one sweep: every row but the split ones, with per-workgroup
// delta and dangling-mass partials.  No fence and no ticket: the fold, the
// next kernel on the stream, completes the sweep.  S is the type of the
// gathered contributions, V that of the sums.
template <typename V, typename S, typename Fin>
__device__ __forceinline__ void sweep_body(const PrArgs &a) {
    __shared__ WaveLds<V> L;
    Fin fin(a, a.st->p_sink);
    spmv_waves<V, Fin::kWeighted>(a.tiles, a.ntiles, a.in_off, a.packed, a.bits, a.tbase,
                                  static_cast<const S *>(a.c_cur), static_cast<const V *>(a.w),
                                  a.long_part, a.short_row, L, fin);
    double dsum = fin.dsum, ssum = fin.ssum;
    block_sum2(dsum, ssum, L.red);
    if (threadIdx.x == 0) {
        a.partials[blockIdx.x] = dsum;
        a.partials[gridDim.x + blockIdx.x] = ssum;
    }
}

// Completes the sweep of the preceding sweep launch (`sgrid` workgroups):
// the split rows, one thread each, then -- in the last workgroup to arrive --
// the fold of the sweep's and this kernel's partials in index order and the
// convergence decision.
template <typename Fin>
__device__ __forceinline__ void fold_body(const PrArgs &a, uint32_t sgrid) {
    __shared__ double red[8];
    __shared__ uint32_t flag;
    const int tid = threadIdx.x;
    Fin lf(a, a.st->p_sink);
    for (uint32_t k = blockIdx.x * kBlock + tid; k < a.nlong; k += gridDim.x * kBlock)
        lf(a.long_rows[k], a.long_idx ? long_row_sum_idx(a.long_part, a.long_slot, a.long_idx, k)
                                      : long_row_sum(a.long_part, a.long_slot, k));
    double d = lf.dsum, s = lf.ssum;
    block_sum2(d, s, red);
    double *fp = a.partials + 2 * kMaxSweepGrid;
    if (tid == 0) {
        fp[blockIdx.x] = d;
        fp[gridDim.x + blockIdx.x] = s;
    }
    if (!last_to_arrive(&a.st->counter, &flag)) return;
    d = 0.0;
    s = 0.0;
    for (uint32_t k = tid; k < sgrid; k += kBlock) {
        d += a.partials[k];
        s += a.partials[sgrid + k];
    }
    for (uint32_t k = tid; k < gridDim.x; k += kBlock) {
        d += fp[k];
        s += fp[gridDim.x + k];
    }
    block_sum2(d, s, red);
    if (tid == 0) {
        SweepState *st = a.st;
        const uint64_t it = st->iter + 1;
        const double thr = fmax(a.eps_abs, a.eps_rel * st->aux);
        st->delta = d;
        st->p_sink = s;
        st->iter = it;
        st->launches += 1;
        st->done = (!(d >= thr) || (a.max_iter > 0 && it >= a.max_iter)) ? 1u : 0u;
        st->counter = 0;
    }
}

template <typename V, bool WEIGHTED, bool DPERS, bool MASK>
__global__ __launch_bounds__(kBlock) void pr_sweep(const PrArgs a) {
    if (a.st->done) return;
    sweep_body<V, V, PrFin<V, WEIGHTED, DPERS, MASK>>(a);
}

// The sweep over a relabelled graph with the hot slice (kHotWaves wavefronts,
// one workgroup per compute unit; the first `nhot` contributions -- the most
// gathered vertices -- in LDS).
template <typename V, bool WEIGHTED, bool DPERS>
__global__ __launch_bounds__(kHotBlock) void pr_sweep_hot(const PrArgs a, uint32_t nhot) {
    if (a.st->done) return;
    __shared__ WaveLds<V, kHotWaves, false> L;
    __shared__ V hot[hot_max<V>()];
    const V *c = static_cast<const V *>(a.c_cur);
    fill_hot(c, nhot, hot);
    PrFin<V, WEIGHTED, DPERS, false> fin(a, a.st->p_sink);
    spmv_waves<V, WEIGHTED, true>(a.tiles, a.ntiles, a.in_off, a.packed, a.bits, a.tbase, c,
                                  static_cast<const V *>(a.w), a.long_part, a.short_row, L, fin,
                                  static_cast<const V *>(hot), nhot);
    double dsum = fin.dsum, ssum = fin.ssum;
    block_sum2_waves<kHotWaves>(dsum, ssum, L.red);
    if (threadIdx.x == 0) {
        a.partials[blockIdx.x] = dsum;
        a.partials[gridDim.x + blockIdx.x] = ssum;
    }
}

// The sliced sweep (relabelled, unweighted): workgroups below
// sa.heavy_wgs sum the long rows' pieces from LDS slices of the
// contributions (their partials are zero: those rows are finished by the
// fold), the others sweep the light rows with the hot slice.
template <typename V, bool DPERS>
__global__ __launch_bounds__(kHotBlock) void pr_sweep_sliced(const PrArgs a, const SlicedArgs sa,
                                                             uint32_t nhot) {
    if (a.st->done) return;
    __shared__ WaveLds<V, kHotWaves, false> L;
    __shared__ V hot[hot_max<V>()];
    const V *c = static_cast<const V *>(a.c_cur);
    if (blockIdx.x < sa.heavy_wgs) {
        sliced_pieces<V, V>(sa, c, a.long_part, a.short_row, L, hot);
        if (threadIdx.x == 0) {
            a.partials[blockIdx.x] = 0.0;
            a.partials[gridDim.x + blockIdx.x] = 0.0;
        }
        return;
    }
    fill_hot(c, nhot, hot);
    PrFin<V, false, DPERS, false> fin(a, a.st->p_sink);
    sliced_light<V, V>(sa, a.in_off, c, a.long_part, a.short_row, L, fin,
                       static_cast<const V *>(hot), nhot, blockIdx.x - sa.heavy_wgs,
                       gridDim.x - sa.heavy_wgs);
    double dsum = fin.dsum, ssum = fin.ssum;
    block_sum2_waves<kHotWaves>(dsum, ssum, L.red);
    if (threadIdx.x == 0) {
        a.partials[blockIdx.x] = dsum;
        a.partials[gridDim.x + blockIdx.x] = ssum;
    }
}

template <typename V, bool WEIGHTED, bool DPERS, bool MASK>
__global__ __launch_bounds__(kBlock) void pr_fold(const PrArgs a, uint32_t sgrid) {
    if (a.st->done) return;
    fold_body<PrFin<V, WEIGHTED, DPERS, MASK>>(a, sgrid);
}

// The refinement runs once, after the FP32 loop finished (its state says
// done), so neither kernel looks at the done flag.
template <bool WEIGHTED, bool DPERS, bool MASK>
__global__ __launch_bounds__(kBlock) void pr_refine(const PrArgs a) {
    sweep_body<double, float, RefineFin<WEIGHTED, DPERS, MASK>>(a);
}

// The refinement over a relabelled, unweighted graph: FP32 tiles and hot
// slice (the gathered values are FP32), FP64 row sums in registers.
template <bool DPERS>
__global__ __launch_bounds__(kHotBlock) void pr_refine_hot(const PrArgs a, uint32_t nhot) {
    __shared__ WaveLds<float, kHotWaves, false> L;
    __shared__ float hot[kHotMax];
    const float *c = static_cast<const float *>(a.c_cur);
    fill_hot(c, nhot, hot);
    RefineFin<false, DPERS, false> fin(a, a.st->p_sink);
    spmv_waves<double, false, true>(a.tiles, a.ntiles, a.in_off, a.packed, a.bits, a.tbase, c,
                                    static_cast<const double *>(a.w), a.long_part, a.short_row,
                                    L, fin, static_cast<const float *>(hot), nhot);
    double dsum = fin.dsum, ssum = fin.ssum;
    block_sum2_waves<kHotWaves>(dsum, ssum, L.red);
    if (threadIdx.x == 0) {
        a.partials[blockIdx.x] = dsum;
        a.partials[gridDim.x + blockIdx.x] = ssum;
    }
}

// The refinement with the sliced split of the FP32 sweeps (FP32 tiles and
// slices, FP64 sums).
template <bool DPERS>
__global__ __launch_bounds__(kHotBlock) void pr_refine_sliced(const PrArgs a, const SlicedArgs sa,
                                                              uint32_t nhot) {
    __shared__ WaveLds<float, kHotWaves, false> L;
    __shared__ float hot[kHotMax];
    const float *c = static_cast<const float *>(a.c_cur);
    if (blockIdx.x < sa.heavy_wgs) {
        sliced_pieces<double, float>(sa, c, a.long_part, a.short_row, L, hot);
        if (threadIdx.x == 0) {
            a.partials[blockIdx.x] = 0.0;
            a.partials[gridDim.x + blockIdx.x] = 0.0;
        }
        return;
    }
    fill_hot(c, nhot, hot);
    RefineFin<false, DPERS, false> fin(a, a.st->p_sink);
    sliced_light<double, float>(sa, a.in_off, c, a.long_part, a.short_row, L, fin,
                                static_cast<const float *>(hot), nhot, blockIdx.x - sa.heavy_wgs,
                                gridDim.x - sa.heavy_wgs);
    double dsum = fin.dsum, ssum = fin.ssum;
    block_sum2_waves<kHotWaves>(dsum, ssum, L.red);
    if (threadIdx.x == 0) {
        a.partials[blockIdx.x] = dsum;
        a.partials[gridDim.x + blockIdx.x] = ssum;
    }
}

template <bool WEIGHTED, bool DPERS, bool MASK>
__global__ __launch_bounds__(kBlock) void pr_refine_fold(const PrArgs a, uint32_t sgrid) {
    fold_body<RefineFin<WEIGHTED, DPERS, MASK>>(a, sgrid);
}

// Seed slots [had, n) of the f64 map with pers, load the initial iterate into
// the value type, compute c0 = x0/deg, the first dangling mass and the total
// mass; resets the loop state (iter = `iter0`).
template <typename V, bool WEIGHTED, bool DPERS, bool MASK>
__global__ __launch_bounds__(kBlock) void pr_init(const PrArgs a, double *rank, uint64_t had,
                                                  uint32_t n, uint64_t iter0) {
    __shared__ double red[8];
    __shared__ uint32_t flag;
    V *x0 = static_cast<V *>(a.x_next);
    V *c0 = static_cast<V *>(a.c_next);
    double ssum = 0.0, mass = 0.0;
    const uint32_t stride = gridDim.x * kBlock;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
        // Row i of the sweep is slot iperm[i] of the caller's map; a cold
        // start (had == 0, the usual case) reads no map and needs no label.
        double xi;
        if (had == 0) {
            xi = DPERS ? a.pers[i] : a.pval;
        } else {
            const uint32_t o = a.iperm ? a.iperm[i] : i;
            xi = o >= had ? (DPERS ? a.pers[i] : a.pval) : rank[o];
        }
        const V xv = static_cast<V>(xi);
        x0[i] = xv;
        if (MASK && !a.active[i]) {
            c0[i] = V(0);
            continue;
        }
        mass += fabs(static_cast<double>(xv));
        const double dg =
            WEIGHTED ? a.deg[i] : static_cast<double>(a.out_off[i + 1] - a.out_off[i]);
        if (dg != 0.0) {
            c0[i] = static_cast<V>(static_cast<double>(xv) / dg);
        } else {
            c0[i] = V(0);
            ssum += static_cast<double>(xv);
        }
    }
    // Seed the caller's map: slots [had, n) take the personalisation.
    for (uint64_t o = had + blockIdx.x * kBlock + threadIdx.x; o < n; o += stride)
        rank[o] = DPERS ? a.pers_c[o] : a.pval;
    block_sum2(ssum, mass, red);
    if (threadIdx.x == 0) {
        a.partials[blockIdx.x] = ssum;
        a.partials[gridDim.x + blockIdx.x] = mass;
    }
    if (!last_to_arrive(&a.st->counter, &flag)) return;
    double s = 0.0, ms = 0.0;
    for (uint32_t k = threadIdx.x; k < gridDim.x; k += kBlock) {
        s += a.partials[k];
        ms += a.partials[gridDim.x + k];
    }
    block_sum2(s, ms, red);
    if (threadIdx.x == 0) {
        SweepState *st = a.st;
        st->p_sink = s;
        st->aux = ms;
        st->delta = 0.0;
        st->iter = iter0;
        st->done = 0;
        st->launches = 0;
        st->counter = 0;
        __threadfence();
    }
}

// out[o] = x[perm[o]] (perm null: identity), widened to f64: an iterate in
// sweep labels back to the caller's labels.
template <typename V>
__global__ __launch_bounds__(kBlock) void unpermute(const V *__restrict__ x,
                                                    const uint32_t *__restrict__ perm, uint32_t n,
                                                    double *__restrict__ out) {
    for (uint32_t o = blockIdx.x * kBlock + threadIdx.x; o < n; o += gridDim.x * kBlock)
        out[o] = static_cast<double>(x[perm ? perm[o] : o]);
}

// out[k] = in[iperm[k]]: a vertex map into sweep labels.
__global__ __launch_bounds__(kBlock) void permute_f64(const double *__restrict__ in,
                                                      const uint32_t *__restrict__ iperm,
                                                      uint32_t n, double *__restrict__ out) {
    for (uint32_t k = blockIdx.x * kBlock + threadIdx.x; k < n; k += gridDim.x * kBlock)
        out[k] = in[iperm[k]];
}

// ------------------------------------------------------------------ host

namespace {

struct PrPlan {
    bool weighted = false;
    bool dpers = false;
    bool masked = false;
};

// The sweep kernel with at most `per_cu` resident workgroups per compute
// unit; returns its workgroup count.
template <typename V, bool W, bool P, bool M>
int launch_sweep(const PrArgs &a, int per_cu, hipStream_t s) {
    const int sg = wave_grid(occupancy_cap((const void *)pr_sweep<V, W, P, M>, per_cu), a.ntiles);
    hipLaunchKernelGGL((pr_sweep<V, W, P, M>), dim3(sg), dim3(kBlock), 0, s, a);
    return sg;
}
// The hot-slice sweep over `nhot` contributions; returns its workgroup count
// (one per compute unit, fewer for a small graph).
template <typename V, bool W, bool P>
int launch_sweep_hot(const PrArgs &a, uint32_t nhot, hipStream_t s) {
    const int sg = fit_grid(occupancy_cap((const void *)pr_sweep_hot<V, W, P>, 1),
                            (a.ntiles + kHotWaves - 1) / kHotWaves);
    hipLaunchKernelGGL((pr_sweep_hot<V, W, P>), dim3(sg), dim3(kHotBlock), 0, s, a, nhot);
    return sg;
}
SlicedArgs sliced_args(const Sliced &S) {
    SlicedArgs sa;
    sa.htiles = S.htiles.as<Tile>();
    sa.htbase = S.htbase.as<uint32_t>();
    sa.hoff = S.hoff.as<uint32_t>();
    sa.hpacked = S.hpacked.as<uint32_t>();
    sa.seg = S.seg.as<SliceSeg>();
    sa.wseg = S.wseg.as<uint32_t>();
    sa.heavy_wgs = S.heavy_wgs;
    sa.nltiles = S.nltiles;
    sa.ltiles = S.ltiles.as<Tile>();
    sa.ltbase = S.ltbase.as<uint32_t>();
    sa.lpacked = S.lpacked.as<uint32_t>();
    sa.lbits = S.lbits;
    sa.nmtiles = S.nmtiles;
    sa.mtiles = S.mtiles.as<Tile>();
    sa.mtbase = S.mtbase.as<uint32_t>();
    sa.moff = S.moff.as<uint32_t>();
    sa.mpacked = S.mpacked.as<uint32_t>();
    sa.mbits = S.mbits;
    sa.mpart = S.mpart;
    return sa;
}
// The sliced sweep on `grid` workgroups (the grid the split was built for).
template <typename V>
int dispatch_sweep_sliced(const PrPlan &p, const PrArgs &a, const SlicedArgs &sa, uint32_t nhot,
                          int grid, hipStream_t s) {
    if (p.dpers)
        hipLaunchKernelGGL((pr_sweep_sliced<V, true>), dim3(grid), dim3(kHotBlock), 0, s, a, sa,
                           nhot);
    else
        hipLaunchKernelGGL((pr_sweep_sliced<V, false>), dim3(grid), dim3(kHotBlock), 0, s, a, sa,
                           nhot);
    return grid;
}
template <typename V>
int dispatch_sweep_hot(const PrPlan &p, const PrArgs &a, uint32_t nhot, hipStream_t s) {
    switch ((p.weighted ? 2 : 0) | (p.dpers ? 1 : 0)) {
        case 0: return launch_sweep_hot<V, false, false>(a, nhot, s);
        case 1: return launch_sweep_hot<V, false, true>(a, nhot, s);
        case 2: return launch_sweep_hot<V, true, false>(a, nhot, s);
        default: return launch_sweep_hot<V, true, true>(a, nhot, s);
    }
}
template <typename V, bool W, bool P, bool M>
int launch_fold(const PrArgs &a, int sg, hipStream_t s) {
    hipLaunchKernelGGL((pr_fold<V, W, P, M>), dim3(fold_grid(a.nlong)), dim3(kBlock), 0, s,
                       a, (uint32_t)sg);
    return sg;
}
template <typename V, bool W, bool P, bool M>
void launch_init(const PrArgs &a, double *rank, uint64_t had, uint32_t n, uint64_t iter0,
                 hipStream_t s) {
    uint32_t blocks = (n + kBlock - 1) / kBlock;
    if (blocks > kTicketGrid) blocks = kTicketGrid;
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL((pr_init<V, W, P, M>), dim3(blocks), dim3(kBlock), 0, s, a, rank, had, n,
                       iter0);
}

#define GT_PR_DISPATCH(FN)                                                      \
    template <typename V>                                                       \
    int dispatch_##FN(const PrPlan &p, const PrArgs &a, int grid, hipStream_t s) { \
        switch ((p.weighted ? 4 : 0) | (p.dpers ? 2 : 0) | (p.masked ? 1 : 0)) { \
            case 0: return launch_##FN<V, false, false, false>(a, grid, s);      \
            case 1: return launch_##FN<V, false, false, true>(a, grid, s);       \
            case 2: return launch_##FN<V, false, true, false>(a, grid, s);       \
            case 3: return launch_##FN<V, false, true, true>(a, grid, s);        \
            case 4: return launch_##FN<V, true, false, false>(a, grid, s);       \
            case 5: return launch_##FN<V, true, false, true>(a, grid, s);        \
            case 6: return launch_##FN<V, true, true, false>(a, grid, s);        \
            default: return launch_##FN<V, true, true, true>(a, grid, s);        \
        }                                                                       \
    }
GT_PR_DISPATCH(sweep)
GT_PR_DISPATCH(fold)
#undef GT_PR_DISPATCH

// The FP64 refinement sweep and its fold (one sweep, then the loop is done).
template <bool W, bool P, bool M>
void launch_refine(const PrArgs &a, int per_cu, hipStream_t s) {
    const int sg = wave_grid(occupancy_cap((const void *)pr_refine<W, P, M>, per_cu), a.ntiles);
    hipLaunchKernelGGL((pr_refine<W, P, M>), dim3(sg), dim3(kBlock), 0, s, a);
    hipLaunchKernelGGL((pr_refine_fold<W, P, M>), dim3(fold_grid(a.nlong)), dim3(kBlock), 0, s,
                       a, (uint32_t)sg);
}
template <bool P>
void launch_refine_hot(const PrArgs &a, uint32_t nhot, hipStream_t s) {
    const int sg = fit_grid(occupancy_cap((const void *)pr_refine_hot<P>, 1),
                            (a.ntiles + kHotWaves - 1) / kHotWaves);
    hipLaunchKernelGGL((pr_refine_hot<P>), dim3(sg), dim3(kHotBlock), 0, s, a, nhot);
    hipLaunchKernelGGL((pr_refine_fold<false, P, false>), dim3(fold_grid(a.nlong)), dim3(kBlock),
                       0, s, a, (uint32_t)sg);
}
// `nhot` > 0: the hot-slice refinement (unweighted, unfiltered, relabelled);
// with `sl` (the FP32 sweeps' sliced split, active), the sliced one.
void dispatch_refine(const PrPlan &p, const PrArgs &a, int per_cu, uint32_t nhot,
                     const Sliced *sl, hipStream_t s) {
    if (nhot && sl) {
        PrArgs b = a;
        b.nlong = sl->nlong;
        b.long_rows = sl->long_rows.as<uint32_t>();
        b.long_slot = sl->long_slot.as<uint32_t>();
        b.long_idx = sl->long_idx.as<uint32_t>();
        const SlicedArgs sa = sliced_args(*sl);
        const int grid = (int)sl->grid;
        if (p.dpers) {
            hipLaunchKernelGGL((pr_refine_sliced<true>), dim3(grid), dim3(kHotBlock), 0, s, b, sa,
                               nhot);
            hipLaunchKernelGGL((pr_refine_fold<false, true, false>), dim3(fold_grid(b.nlong)),
                               dim3(kBlock), 0, s, b, (uint32_t)grid);
        } else {
            hipLaunchKernelGGL((pr_refine_sliced<false>), dim3(grid), dim3(kHotBlock), 0, s, b,
                               sa, nhot);
            hipLaunchKernelGGL((pr_refine_fold<false, false, false>), dim3(fold_grid(b.nlong)),
                               dim3(kBlock), 0, s, b, (uint32_t)grid);
        }
        return;
    }
    if (nhot) {
        if (p.dpers)
            launch_refine_hot<true>(a, nhot, s);
        else
            launch_refine_hot<false>(a, nhot, s);
        return;
    }
    switch ((p.weighted ? 4 : 0) | (p.dpers ? 2 : 0) | (p.masked ? 1 : 0)) {
        case 0: launch_refine<false, false, false>(a, per_cu, s); break;
        case 1: launch_refine<false, false, true>(a, per_cu, s); break;
        case 2: launch_refine<false, true, false>(a, per_cu, s); break;
        case 3: launch_refine<false, true, true>(a, per_cu, s); break;
        case 4: launch_refine<true, false, false>(a, per_cu, s); break;
        case 5: launch_refine<true, false, true>(a, per_cu, s); break;
        case 6: launch_refine<true, true, false>(a, per_cu, s); break;
        default: launch_refine<true, true, true>(a, per_cu, s); break;
    }
}

template <typename V>
void dispatch_init(const PrPlan &p, const PrArgs &a, double *rank, uint64_t had, uint32_t n,
                   uint64_t iter0, hipStream_t s) {
    const int key = (p.weighted ? 4 : 0) | (p.dpers ? 2 : 0) | (p.masked ? 1 : 0);
    switch (key) {
        case 0: launch_init<V, false, false, false>(a, rank, had, n, iter0, s); break;
        case 1: launch_init<V, false, false, true>(a, rank, had, n, iter0, s); break;
        case 2: launch_init<V, false, true, false>(a, rank, had, n, iter0, s); break;
        case 3: launch_init<V, false, true, true>(a, rank, had, n, iter0, s); break;
        case 4: launch_init<V, true, false, false>(a, rank, had, n, iter0, s); break;
        case 5: launch_init<V, true, false, true>(a, rank, had, n, iter0, s); break;
        case 6: launch_init<V, true, true, false>(a, rank, had, n, iter0, s); break;
        default: launch_init<V, true, true, true>(a, rank, had, n, iter0, s); break;
    }
}

}  // namespace

gt_status resolve_view(gt_graph *g, const gt_view *view, Orient **pull, Orient **push,
                       bool *reversed, const uint8_t **active, uint32_t *nactive) {
    int kind = view ? view->kind : GT_VIEW_DIRECTED;
    *active = nullptr;
    *nactive = g->n;
    // The message of ViewError::NotDirected (crates/gt-py/src/dispatch.rs:176-179).
    if (kind == GT_VIEW_UNDIRECTED || kind == GT_VIEW_UNDIRECTED_FILTERED)
        return fail(GT_ERR_NOT_DIRECTED,
                    std::string("view: ") +
                        (kind == GT_VIEW_UNDIRECTED ? "Undirected" : "UndirectedFiltered") +
                        " has no in-edges; this kernel accepts [Directed, Reversed, "
                        "DirectedFiltered, ReversedFiltered]");
    if (kind == GT_VIEW_DIRECTED_FILTERED || kind == GT_VIEW_REVERSED_FILTERED) {
        *reversed = (kind == GT_VIEW_REVERSED_FILTERED);
        return filtered_view(g, view, *reversed, pull, push, active, nactive);
    }
    if (kind != GT_VIEW_DIRECTED && kind != GT_VIEW_REVERSED)
        return fail(GT_ERR_INVALID_ARG, "unknown view kind");
    *reversed = (kind == GT_VIEW_REVERSED);
    *pull = *reversed ? &g->out : &g->in;
    *push = *reversed ? &g->in : &g->out;
    return GT_OK;
}

gt_status check_vertex_out(gt_graph *g, const gt_vertex_out *r, const char *who) {
    if (!r || (!r->data && g->n)) return fail(GT_ERR_INVALID_ARG, "output map is null");
    if (r->graph_id && r->graph_id != g->id)
        return fail(GT_ERR_WRONG_GRAPH, std::string(who) + " map does not belong to this graph");
    if (r->len < g->n)
        return fail(GT_ERR_INVALID_ARG, "output map is shorter than the vertex bound");
    if (r->loc != GT_HOST && r->loc != GT_DEVICE) return fail(GT_ERR_INVALID_ARG, "bad location");
    return GT_OK;
}

gt_status check_params(gt_graph *g, const gt_power_iteration *p) {
    if (!p) return fail(GT_ERR_INVALID_ARG, "params is null");
    if (p->plan.len != g->n)
        return fail(GT_ERR_PLAN_COVER, "sweep: the plan must cover the output exactly");
    if (p->plan.chunks == 0) return fail(GT_ERR_INVALID_ARG, "plan has zero chunks");
    if (p->max_iter == 0 && !(p->epsilon > 0.0) && !std::isnan(p->epsilon))
        return fail(GT_ERR_INVALID_ARG,
                    "epsilon must be positive when max_iter is 0: the loop would never end");
    return GT_OK;
}

namespace {

// Everything a PageRank solve needs on the device, prepared once per call.
struct PrSolve {
    gt_graph *g = nullptr;
    Orient *pull = nullptr, *push = nullptr;
    PrPlan plan;
    PrArgs base{};
    double *rank64 = nullptr;  // the f64 map on the device (caller's or staged)
    const gt_prop *weight = nullptr;  // the weight map of a weighted solve
    const void *wsrc = nullptr;
    int wdtype = GT_F64;
    // Set when the sweeps run in the relabelled orientation: sweep vectors are
    // in new labels, rank64 stays in the caller's.
    const uint32_t *perm = nullptr;
    const uint32_t *iperm = nullptr;
    Relabeled *rel = nullptr;
    bool noalloc = false;  // every buffer is given: build nothing
};

unsigned vblocks(uint32_t n) {
    uint32_t b = (n + kBlock - 1) / kBlock;
    if (b > 2048) b = 2048;
    return b ? b : 1u;
}

// rank64 <- latest (sweep labels), widened; no-op when it is rank64 itself.
template <typename V>
gt_status pr_unpermute(PrSolve &S, const V *latest) {
    if ((const void *)latest == (const void *)S.rank64) return GT_OK;
    hipLaunchKernelGGL((unpermute<V>), dim3(vblocks(S.g->n)), dim3(kBlock), 0, S.g->stream, latest,
                       S.perm, S.g->n, S.rank64);
    GT_HIP(hipGetLastError());
    return GT_OK;
}

gt_status pr_prepare(gt_graph *g, const gt_view *view, const gt_prop *weight,
                     const gt_prop *pers, double damping, const gt_power_iteration *params,
                     gt_vertex_out *rank, PrSolve &S) {
    S.g = g;
    bool reversed = false;
    GT_TRY(check_vertex_out(g, rank, "pagerank: rank"));
    GT_TRY(check_params(g, params));
    const uint8_t *active = nullptr;
    uint32_t nactive = 0;
    GT_TRY(resolve_view(g, view, &S.pull, &S.push, &reversed, &active, &nactive));
    const uint32_t n = g->n;
    hipStream_t s = g->stream;
    // pagerank's uniform personalisation is 1 / num_vertices of the view
    // (centrality.rs:245-247), the filtered count on a filtered view.
    const double uniform = nactive == 0 ? 0.0 : 1.0 / (double)nactive;
    PrArgs &a = S.base;
    S.plan.masked = active != nullptr;
    a.active = active;
    a.out_off = S.push->off.as<uint32_t>();
    Orient *push_caller = S.push;  // caller's labels (degree staging)
    if (!active && n > 0) {
        Relabeled *R = nullptr;
        GT_TRY(ensure_relabeled(g, reversed ? 1 : 0, &R));
        S.pull = &R->pull;
        S.rel = R;
        S.perm = R->perm.as<uint32_t>();
        S.iperm = R->iperm.as<uint32_t>();
        a.out_off = R->push_off.as<uint32_t>();
    }
    a.iperm = S.iperm;
    GT_TRY(stage_vertex_prop(g, pers, g->ws.pers, &a.pers, &a.pval, &S.plan.dpers, uniform));
    a.pers_c = a.pers;
    if (S.plan.dpers && S.iperm) {
        GT_TRY(g->ws.pers2.reserve(sizeof(double) * n));
        hipLaunchKernelGGL(permute_f64, dim3(vblocks(n)), dim3(kBlock), 0, s, a.pers, S.iperm, n,
                           g->ws.pers2.as<double>());
        GT_HIP(hipGetLastError());
        a.pers = g->ws.pers2.as<double>();
    }
    if (weight && weight->kind != GT_MAP_UNITY) {
        S.plan.weighted = true;
        S.weight = weight;
        GT_TRY(stage_edge_source(g, weight, &S.wsrc, &S.wdtype));
        GT_TRY(stage_degree(g, weight, *push_caller, S.wsrc, S.wdtype, g->ws.deg));
        a.deg = g->ws.deg.as<double>();
        if (S.iperm) {
            GT_TRY(g->ws.deg2.reserve(sizeof(double) * n));
            hipLaunchKernelGGL(permute_f64, dim3(vblocks(n)), dim3(kBlock), 0, s, a.deg, S.iperm,
                               n, g->ws.deg2.as<double>());
            GT_HIP(hipGetLastError());
            a.deg = g->ws.deg2.as<double>();
        }
    }
    if (n == 0) return GT_OK;
    GT_TRY(ensure_schedule(g, *S.pull, n));
    note_layout(g, *S.pull);
    GT_TRY(prof_prepare(g));

    Workspace &ws = g->ws;
    GT_TRY(ws.partials.reserve(sizeof(double) * kPartials));
    GT_TRY(ws.long_part.reserve(sizeof(double) * std::max<uint32_t>(1, S.pull->nslots)));
    GT_TRY(ws.state.reserve(sizeof(SweepState)));
    GT_HIP(hipMemsetAsync(ws.state.ptr, 0, sizeof(SweepState), s));

    if (rank->loc == GT_DEVICE) {
        S.rank64 = rank->data;
    } else {
        GT_TRY(ws.y3.reserve(sizeof(double) * n));
        S.rank64 = ws.y3.as<double>();
        const uint64_t had = std::min<uint64_t>(rank->had, n);
        if (had)
            GT_HIP(hipMemcpyAsync(S.rank64, rank->data, sizeof(double) * had,
                                  hipMemcpyHostToDevice, s));
    }

    a.tiles = S.pull->tiles.as<Tile>();
    a.ntiles = S.pull->ntiles;
    a.nlong = S.pull->nlong;
    a.in_off = S.pull->off.as<uint32_t>();
    a.packed = S.pull->packed.as<uint32_t>();
    a.bits = S.pull->bits;
    a.tbase = S.pull->tbase.as<uint32_t>();
    a.short_row = (uint32_t)g->tune.short_row;
    a.damping = damping;
    a.eps_abs = params->epsilon;
    a.eps_rel = 0.0;
    a.max_iter = params->max_iter;
    a.partials = ws.partials.as<double>();
    a.long_part = ws.long_part.as<double>();
    a.long_rows = S.pull->long_rows.as<uint32_t>();
    a.long_slot = S.pull->long_slot.as<uint32_t>();
    a.st = ws.state.as<SweepState>();
    return GT_OK;
}

// The sliced split for the sweeps of value type V over `nhot`-label slices,
// or null: weighted, filtered, the tuning turns it off, the vector fits one
// slice, or no row is long enough.  Grows the partial slots to its pieces.
template <typename V>
gt_status sliced_for(PrSolve &S, uint32_t nhot, const Sliced **out) {
    *out = nullptr;
    gt_graph *g = S.g;
    if (!nhot || !S.rel || S.noalloc || S.plan.weighted || S.plan.masked ||
        g->tune.sliced_min_len < 0 ||
        g->n <= nhot)
        return GT_OK;
    const int grid = occupancy_cap((const void *)pr_sweep_sliced<V, false>, 1);
    Sliced *SL = nullptr;
    GT_TRY(ensure_sliced(g, *S.rel, nhot, sizeof(V), grid, &SL));
    if (!SL->active) return GT_OK;
    GT_TRY(g->ws.long_part.reserve(sizeof(double) * std::max<uint32_t>(1, SL->nout)));
    g->layout.slices = SL->nslices;
    g->layout.sliced_rows = SL->nlong;
    g->layout.sliced_entries = SL->heavy_entries;
    *out = SL;
    return GT_OK;
}

// Whether the sweeps of S can run on block tiles (blocks.hpp): relabelled
// and unfiltered, unless the tuning turns them off.  A weighted solve also
// needs every weight finite and non-negative (checked when they are staged).
bool blocks_apply(const PrSolve &S) {
    return S.rel && S.iperm && !S.plan.masked && !S.noalloc && S.g->tune.block_tile_nnz > 0;
}

// The block-sweep arguments of a solve over B.
BlkArgs blk_args(const PrSolve &S, const PrArgs &a, const Blocks &B) {
    BlkArgs b = blk_layout_args(S.g, B);
    b.st = a.st;
    b.pers = a.pers;
    b.pval = a.pval;
    b.damping = a.damping;
    b.eps_abs = a.eps_abs;
    b.eps_rel = a.eps_rel;
    b.max_iter = a.max_iter;
    return b;
}

// Init + sweeps on block tiles in value type V (FP32: the FP32 form's
// phase; FP64: the drop-in), on buffers x / c; returns the sweep count
// reached and leaves the latest iterate in x[count & 1], as pr_phase.  Sets
// *bailed when the phase must run on the wave engine instead: a sweep met an
// iterate whose mass its fixed-point sums cannot hold (not finite, or above
// kBlkMassMax), or (FP64) a row is longer than kBlkF64MaxRow.
template <typename V>
gt_status pr_phase_blocks(PrSolve &S, PrArgs a, V *x[2], V *c[2], uint64_t had,
                          double eps_pred, uint64_t cap, uint64_t *iter_out, bool *bailed) {
    constexpr bool F64 = std::is_same<V, double>::value;
    constexpr int mode = F64 ? kBlkF64 : kBlkF32;
    gt_graph *g = S.g;
    hipStream_t s = g->stream;
    *bailed = false;
    Blocks *B = nullptr;
    GT_TRY(ensure_blocks(g, *S.rel, (uint32_t)g->tune.block_tile_nnz,
                         (uint32_t)g->tune.block_long_min, &B));
    if (F64 && B->max_row > kBlkF64MaxRow) {
        *bailed = true;
        return GT_OK;
    }
    DevBuf &wbuf = F64 ? g->ws.blk_w64 : g->ws.blk_w;
    if (S.plan.weighted) {
        bool usable = false;
        GT_TRY(blk_stage_weights(g, *B, S.weight, S.wsrc, S.wdtype, F64, wbuf, &usable));
        if (!usable) {
            *bailed = true;
            return GT_OK;
        }
    }
    GT_TRY(blk_reserve(g, *B));
    g->prof_kind = F64 ? GT_SWEEP_PAGERANK_F64 : GT_SWEEP_PAGERANK_F32;
    a.x_next = x[0];
    a.c_next = c[0];
    dispatch_init<V>(S.plan, a, S.rank64, had, g->n, 0, s);
    BlkArgs b = blk_args(S, a, *B);
    if (S.plan.weighted) {
        b.wblk = wbuf.ptr;
        b.wdeg = a.deg;
    }
    blk_launch_start(b, s);
    GT_HIP(hipGetLastError());
    auto sweep = [&](uint64_t k) {
        const int par = (int)(k & 1);
        if constexpr (F64) {
            b.x_cur64 = x[par];
            b.x_next64 = x[par ^ 1];
            b.c_cur64 = c[par];
            b.c_next64 = c[par ^ 1];
        } else {
            b.x_cur = x[par];
            b.x_next = x[par ^ 1];
            b.c_cur = c[par];
            b.c_next = c[par ^ 1];
        }
        blk_launch_sweep(b, S.plan.dpers, g->tune.block_per_cu, mode, s);
    };
    auto fold = [&](uint64_t) { blk_launch_fold(b, S.plan.dpers, mode, s); };
    uint64_t launched = 0;
    GT_TRY(drive(g, a.st, eps_pred, cap, sweep, fold, &launched, g->prof_kind));
    *bailed = g->h_state->bail != 0;
    *iter_out = g->h_state->iter;
    return GT_OK;
}

// Init + sweeps in value type V on buffers x / c; returns the sweep count
// reached (st->iter) and leaves the latest iterate in x[count & 1].
template <typename V>
gt_status pr_phase(PrSolve &S, PrArgs a, V *x[2], V *c[2], uint64_t had, uint64_t iter0,
                   double eps_pred, uint64_t cap, uint64_t *iter_out) {
    gt_graph *g = S.g;
    hipStream_t s = g->stream;
    // The FP64 drop-in's sweeps run on block tiles unless tuned off.  Its
    // weighted and unweighted block sweeps share their shape and fixed-point
    // sums, so unit weights give the unweighted result bit for bit, as
    // centrality.rs's unity fast path does.
    const bool blk = std::is_same<V, float>::value || g->tune.block_f64 > 0;
    if (iter0 == 0 && cap != 1 && blk && blocks_apply(S)) {
        bool bailed = false;
        GT_TRY(pr_phase_blocks<V>(S, a, x, c, had, eps_pred, cap, iter_out, &bailed));
        if (!bailed) return GT_OK;
        g->layout.block_tiles = 0;
        g->layout.block_long_rows = 0;
        g->layout.block_groups = 0;
    }
    if (S.plan.weighted && !a.w) {
        // The wave engine's weights, in its packed order.
        DevBuf &wp = std::is_same<V, double>::value ? g->ws.wcsc64 : g->ws.wcsc;
        GT_TRY(stage_pull_weights<V>(g, S.weight, *S.pull, S.wsrc, S.wdtype, wp));
        a.w = wp.ptr;
    }
    g->prof_kind = std::is_same<V, double>::value ? GT_SWEEP_PAGERANK_F64 : GT_SWEEP_PAGERANK_F32;
    // The hot slice serves sweeps in relabelled order (the most gathered
    // vertices first); masked views keep the caller's labels.
    const uint32_t nhot =
        S.iperm && !S.plan.masked ? hot_len<V>(g->tune.sweep_hot_kb, g->n) : 0u;
    // With a vector wider than the slice, the sliced split serves the long
    // rows from LDS slices of that width (unweighted sweeps).
    const Sliced *SL = nullptr;
    GT_TRY(sliced_for<V>(S, nhot, &SL));
    if (SL) {
        a.nlong = SL->nlong;
        a.long_rows = SL->long_rows.as<uint32_t>();
        a.long_slot = SL->long_slot.as<uint32_t>();
        a.long_idx = SL->long_idx.as<uint32_t>();
        a.long_part = g->ws.long_part.as<double>();
    }
    const SlicedArgs sa = SL ? sliced_args(*SL) : SlicedArgs{};
    a.x_next = x[0];
    a.c_next = c[0];
    dispatch_init<V>(S.plan, a, S.rank64, had, g->n, iter0, s);
    GT_HIP(hipGetLastError());
    uint64_t launched = 0;
    PrArgs b = a;
    int sg = 0;
    auto sweep = [&](uint64_t k) {
        const int par = (int)(k & 1);
        b.x_cur = x[par];
        b.x_next = x[par ^ 1];
        b.c_cur = c[par];
        b.c_next = c[par ^ 1];
        sg = SL     ? dispatch_sweep_sliced<V>(S.plan, b, sa, nhot, (int)SL->grid, s)
             : nhot ? dispatch_sweep_hot<V>(S.plan, b, nhot, s)
                    : dispatch_sweep<V>(S.plan, b, g->tune.sweep_blocks_per_cu, s);
    };
    auto fold = [&](uint64_t) { dispatch_fold<V>(S.plan, b, sg, s); };
    if (cap == 1) {
        // Exactly one sweep (the FP64 refinement of the FP32 form): nothing
        // to decide, so no wait on the device here.
        GT_TRY(prof_begin(g, 0));
        sweep(0);
        GT_TRY(prof_end(g, 0));
        fold(0);
        GT_HIP(hipGetLastError());
        GT_TRY(prof_flush(g, 1));
        *iter_out = iter0 + 1;
        return GT_OK;
    }
    GT_TRY(drive(g, a.st, eps_pred, cap, sweep, fold, &launched,
                 iter0 > 0 ? GT_SWEEP_KINDS : g->prof_kind));
    *iter_out = g->h_state->iter;
    return GT_OK;
}

// Copy the f64 iterate `latest` into the caller's map and wait.
gt_status pr_deliver(gt_graph *g, const double *latest, gt_vertex_out *rank) {
    hipStream_t s = g->stream;
    const size_t b = sizeof(double) * g->n;
    if (rank->loc == GT_DEVICE) {
        if (latest != rank->data)
            GT_HIP(hipMemcpyAsync(rank->data, latest, b, hipMemcpyDeviceToDevice, s));
    } else {
        GT_HIP(hipMemcpyAsync(rank->data, latest, b, hipMemcpyDeviceToHost, s));
    }
    GT_HIP(hipStreamSynchronize(s));
    return GT_OK;
}

bool no_sweep_for(double eps) { return std::isnan(eps) || !(eps + 1.0 >= eps); }

}  // namespace

gt_status pagerank_f64_impl(gt_graph *g, const gt_view *view, const gt_prop *weight,
                            const gt_prop *pers, double damping,
                            const gt_power_iteration *params, gt_vertex_out *rank,
                            uint64_t *iters) {
    clear_error();
    if (!g) return fail(GT_ERR_INVALID_ARG, "null graph");
    if (!iters) return fail(GT_ERR_INVALID_ARG, "iters is null");
    PrSolve S;
    GT_TRY(pr_prepare(g, view, weight, pers, damping, params, rank, S));
    const uint32_t n = g->n;
    const bool skip = no_sweep_for(params->epsilon);
    if (n == 0) {
        *iters = skip ? 0 : 1;
        g->last_sweeps = *iters;
        return GT_OK;
    }
    Workspace &ws = g->ws;
    const size_t vb = sizeof(double) * (size_t)n;
    GT_TRY(ws.x[0].reserve(vb));
    GT_TRY(ws.x[1].reserve(vb));
    // Contributions carry four spare values: a slice fill of the sliced
    // sweep reads whole 16-byte groups.
    GT_TRY(ws.c[0].reserve(vb + 32));
    GT_TRY(ws.c[1].reserve(vb + 32));
    // Without relabelling the caller's map is the first iterate itself.
    double *x[2] = {S.iperm ? ws.x[0].as<double>() : S.rank64, ws.x[1].as<double>()};
    double *c[2] = {ws.c[0].as<double>(), ws.c[1].as<double>()};
    uint64_t sweeps = 0;
    if (skip) {
        // Seed only.
        PrArgs a = S.base;
        a.x_next = x[0];
        a.c_next = c[0];
        dispatch_init<double>(S.plan, a, S.rank64, rank->had, n, 0, g->stream);
        GT_HIP(hipGetLastError());
    } else {
        GT_TRY(pr_phase<double>(S, S.base, x, c, rank->had, 0, params->epsilon, params->max_iter,
                                &sweeps));
    }
    g->last_sweeps = sweeps;
    *iters = sweeps;
    if (S.iperm) {
        GT_TRY(pr_unpermute<double>(S, x[sweeps & 1]));
        return pr_deliver(g, S.rank64, rank);
    }
    return pr_deliver(g, x[sweeps & 1], rank);
}

// FP32 sweeps (values, weights, SpMV accumulation) with FP64 dangling mass,
// teleport update and L1 test, until delta < max(eps, 2^-22 * mass) -- below
// that the L1 change of an FP32 iterate is rounding noise -- then, when
// `refine` is non-zero, one FP64 sweep over the FP32 phase's last
// contributions (FP64 sums, teleport and output).
gt_status pagerank_mixed_impl(gt_graph *g, const gt_view *view, const gt_prop *weight,
                              const gt_prop *pers, double damping,
                              const gt_power_iteration *params, gt_vertex_out *rank,
                              uint64_t refine, uint64_t *iters) {
    clear_error();
    if (!g) return fail(GT_ERR_INVALID_ARG, "null graph");
    if (!iters) return fail(GT_ERR_INVALID_ARG, "iters is null");
    PrSolve S;
    GT_TRY(pr_prepare(g, view, weight, pers, damping, params, rank, S));
    const uint32_t n = g->n;
    const bool skip = no_sweep_for(params->epsilon);
    if (n == 0 || skip) {
        if (n) {
            // Seeding only: reuse the FP64 path.
            return pagerank_f64_impl(g, view, weight, pers, damping, params, rank, iters);
        }
        *iters = skip ? 0 : 1;
        g->last_sweeps = *iters;
        return GT_OK;
    }
    Workspace &ws = g->ws;
    hipStream_t s = g->stream;
    for (int k = 0; k < 2; ++k) {
        GT_TRY(ws.x32[k].reserve(sizeof(float) * n));
        GT_TRY(ws.c32[k].reserve(sizeof(float) * ((size_t)n + 4)));
    }
    float *x32[2] = {ws.x32[0].as<float>(), ws.x32[1].as<float>()};
    float *c32[2] = {ws.c32[0].as<float>(), ws.c32[1].as<float>()};
    PrArgs a32 = S.base;
    a32.eps_rel = std::ldexp(1.0, -22);
    uint64_t cap32 = params->max_iter;
    uint64_t it32 = 0;
    GT_TRY(pr_phase<float>(S, a32, x32, c32, rank->had, 0,
                           std::max(params->epsilon, std::ldexp(1.0, -22)), cap32, &it32));
    (void)s;
    uint64_t total = it32;
    const double *latest = S.rank64;
    // A NaN delta stops the loop as it stops the reference's (`NaN >= eps` is
    // false) without converging: like a capped run, no refinement sweep.
    const bool capped = (params->max_iter > 0 && it32 >= params->max_iter) ||
                        std::isnan(g->h_state->delta);
    if (refine == 0 || capped) {
        // Widen the FP32 iterate into the caller's f64 map (and labels).
        GT_TRY(pr_unpermute<float>(S, x32[it32 & 1]));
    } else {
        // One FP64 sweep from the FP32 phase's last contributions (already in
        // HBM): no init kernel, 4-byte gathers, FP64 sums and output.
        PrArgs ar = S.base;
        const bool on_blocks = g->layout.block_tiles != 0;
        if (S.plan.weighted && on_blocks) {
            // The map's weights in FP64 and block order, as the wave
            // engine's refinement uses them: an f32 map's, widened from the
            // FP32 sweeps' copy.
            if (S.wdtype == GT_F32 && weight->kind == GT_MAP_DENSE) {
                GT_TRY(blk_widen_weights(g, S.rel->blk, ws.blk_w, ws.blk_w64));
            } else {
                bool usable = false;
                GT_TRY(blk_stage_weights(g, S.rel->blk, weight, S.wsrc, S.wdtype, true,
                                         ws.blk_w64, &usable));
                if (!usable) return fail(GT_ERR_UNSUPPORTED, "block tiles: weights changed");
            }
        } else if (S.plan.weighted) {
            // An f32 map's weights are already staged as f32 in packed order
            // for the FP32 sweeps: widening them is exact.
            if (S.wdtype == GT_F32 && weight->kind == GT_MAP_DENSE)
                GT_TRY(widen_pull_weights(g, *S.pull, ws.wcsc, ws.wcsc64));
            else
                GT_TRY(stage_pull_weights<double>(g, weight, *S.pull, S.wsrc, S.wdtype, ws.wcsc64));
            ar.w = ws.wcsc64.ptr;
        }
        GT_TRY(ws.x[0].reserve(sizeof(double) * n));
        double *x64 = S.iperm ? ws.x[0].as<double>() : S.rank64;
        ar.x_cur = x32[it32 & 1];
        ar.c_cur = c32[it32 & 1];
        ar.x_next = x64;
        ar.max_iter = it32 + 1;
        g->prof_kind = GT_SWEEP_PAGERANK_F64;
        GT_TRY(prof_prepare(g));
        GT_TRY(prof_begin(g, 0));
        if (on_blocks) {
            // The FP32 sweeps ran on block tiles: so does the refinement
            // (fixed-point sums, FP64 update and output).
            BlkArgs b = blk_args(S, ar, S.rel->blk);
            b.x_cur = x32[it32 & 1];
            b.c_cur = c32[it32 & 1];
            b.x_next64 = x64;
            if (S.plan.weighted) {
                b.wblk = ws.blk_w64.ptr;
                b.wdeg = ar.deg;
            }
            blk_launch_sweep(b, S.plan.dpers, g->tune.block_per_cu, kBlkRefine, s);
            GT_TRY(prof_end(g, 0));
            blk_launch_fold(b, S.plan.dpers, kBlkRefine, s);
        } else {
            const uint32_t nhot = S.iperm && !S.plan.masked && !S.plan.weighted
                                      ? hot_len<float>(g->tune.sweep_hot_kb, n)
                                      : 0u;
            const Sliced *SL = nullptr;
            if (nhot) GT_TRY(sliced_for<float>(S, nhot, &SL));
            if (SL) ar.long_part = ws.long_part.as<double>();
            dispatch_refine(S.plan, ar, g->tune.sweep_blocks_per_cu, nhot, SL, s);
            GT_TRY(prof_end(g, 0));
        }
        GT_HIP(hipGetLastError());
        GT_TRY(prof_flush(g, 1));
        total = it32 + 1;
        latest = x64;
        if (S.iperm) {
            GT_TRY(pr_unpermute<double>(S, latest));
            latest = S.rank64;
        }
    }
    g->last_sweeps = total;
    *iters = total;
    return pr_deliver(g, latest, rank);
}

// ------------------------------------------------------- no allocation

// The scratch of gt_pagerank_f64_noalloc: its iterates, contributions,
// partials and loop state are carved from this array instead of allocated.
// Every device the library runs on holds its own copy until the process
// ends; calls on one device share it and keep the last call's contents, so
// calls that overlap on one device race.
constexpr size_t kNoAllocScratch = 64ull << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratch];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch[kMaxDevices] = {};  // device address of `scratch`, per device

// The current device's copy of `scratch`, looked up once per device.
gt_status scratch_base(void **out) {
    int dev = 0;
    GT_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        return fail(GT_ERR_UNSUPPORTED, "no-alloc pagerank: device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return hip_fail(e, "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return GT_OK;
}

__global__ __launch_bounds__(64) void zero_state(SweepState *st) {
    uint32_t *w = reinterpret_cast<uint32_t *>(st);
    if (threadIdx.x < sizeof(SweepState) / 4) w[threadIdx.x] = 0u;
}

}  // namespace

gt_status pagerank_f64_noalloc_impl(gt_graph *g, const gt_view *view, const gt_prop *pers,
                                    double damping, const gt_power_iteration *params,
                                    gt_vertex_out *rank, uint64_t *iters) {
    clear_error();
    if (!g) return fail(GT_ERR_INVALID_ARG, "null graph");
    if (!iters) return fail(GT_ERR_INVALID_ARG, "iters is null");
    GT_TRY(check_vertex_out(g, rank, "pagerank: rank"));
    GT_TRY(check_params(g, params));
    const int kind = view ? view->kind : GT_VIEW_DIRECTED;
    if (kind == GT_VIEW_DIRECTED_FILTERED || kind == GT_VIEW_REVERSED_FILTERED)
        return fail(GT_ERR_UNSUPPORTED,
                    "no-alloc pagerank: a filtered view builds its orientation per call; use "
                    "gt_pagerank_f64");
    PrSolve S;
    S.g = g;
    S.noalloc = true;
    bool reversed = false;
    const uint8_t *active = nullptr;
    uint32_t nactive = 0;
    GT_TRY(resolve_view(g, view, &S.pull, &S.push, &reversed, &active, &nactive));
    if (rank->loc != GT_DEVICE)
        return fail(GT_ERR_UNSUPPORTED, "no-alloc pagerank: the rank map must be in device memory");
    if (pers && pers->kind == GT_MAP_DENSE && !(pers->dtype == GT_F64 && pers->loc == GT_DEVICE))
        return fail(GT_ERR_UNSUPPORTED,
                    "no-alloc pagerank: a dense personalisation must be f64 in device memory");
    const uint32_t n = g->n;
    const bool skip = no_sweep_for(params->epsilon);
    if (n == 0) {
        *iters = skip ? 0 : 1;
        g->last_sweeps = *iters;
        return GT_OK;
    }
    Relabeled *R = &g->rel[reversed ? 1 : 0];
    if (!R->ready || !R->pull.scheduled || R->pull.sched_rows != (uint32_t)g->tune.tile_rows)
        return fail(GT_ERR_UNSUPPORTED,
                    "no-alloc pagerank: the graph's sweep schedule is not built; run "
                    "gt_pagerank_f64 once on this graph and view first");
    S.pull = &R->pull;
    S.rel = R;
    S.perm = R->perm.as<uint32_t>();
    S.iperm = R->iperm.as<uint32_t>();
    PrArgs &a = S.base;
    a.active = nullptr;
    a.out_off = R->push_off.as<uint32_t>();
    a.iperm = S.iperm;
    DevBuf unused;  // stage_vertex_prop stages nothing for the maps accepted above
    GT_TRY(stage_vertex_prop(g, pers, unused, &a.pers, &a.pval, &S.plan.dpers,
                             1.0 / (double)nactive));
    a.pers_c = a.pers;

    // The scratch, in 256-byte-aligned pieces: partials, loop state, split
    // row slots, two iterates and two contributions (each with four spare
    // values), the personalisation in sweep labels.
    void *base = nullptr;
    GT_TRY(scratch_base(&base));
    size_t used = 0;
    auto carve = [&](size_t bytes) {
        unsigned char *p = static_cast<unsigned char *>(base) + used;
        used += (bytes + 255) & ~size_t(255);
        return p;
    };
    double *partials = reinterpret_cast<double *>(carve(sizeof(double) * kPartials));
    SweepState *st = reinterpret_cast<SweepState *>(carve(sizeof(SweepState)));
    double *long_part =
        reinterpret_cast<double *>(carve(sizeof(double) * std::max<uint32_t>(1, S.pull->nslots)));
    double *x[2], *c[2];
    for (int k = 0; k < 2; ++k) x[k] = reinterpret_cast<double *>(carve(sizeof(double) * (n + 4ull)));
    for (int k = 0; k < 2; ++k) c[k] = reinterpret_cast<double *>(carve(sizeof(double) * (n + 4ull)));
    double *pers2 = S.plan.dpers ? reinterpret_cast<double *>(carve(sizeof(double) * n)) : nullptr;
    if (used > kNoAllocScratch) {
        char msg[200];
        std::snprintf(msg, sizeof msg,
                      "no-alloc pagerank: this graph needs %zu bytes of scratch; the maximum is "
                      "%zu (about 1.6 M vertices)",
                      used, kNoAllocScratch);
        return fail(GT_ERR_UNSUPPORTED, msg);
    }
    hipStream_t s = g->stream;
    if (S.plan.dpers) {
        hipLaunchKernelGGL(permute_f64, dim3(vblocks(n)), dim3(kBlock), 0, s, a.pers, S.iperm, n,
                           pers2);
        a.pers = pers2;
    }
    hipLaunchKernelGGL(zero_state, dim3(1), dim3(64), 0, s, st);
    GT_HIP(hipGetLastError());
    note_layout(g, *S.pull);
    S.rank64 = rank->data;
    a.tiles = S.pull->tiles.as<Tile>();
    a.ntiles = S.pull->ntiles;
    a.nlong = S.pull->nlong;
    a.in_off = S.pull->off.as<uint32_t>();
    a.packed = S.pull->packed.as<uint32_t>();
    a.bits = S.pull->bits;
    a.tbase = S.pull->tbase.as<uint32_t>();
    a.short_row = (uint32_t)g->tune.short_row;
    a.damping = damping;
    a.eps_abs = params->epsilon;
    a.eps_rel = 0.0;
    a.max_iter = params->max_iter;
    a.partials = partials;
    a.long_part = long_part;
    a.long_rows = S.pull->long_rows.as<uint32_t>();
    a.long_slot = S.pull->long_slot.as<uint32_t>();
    a.st = st;
    // Profiling would create events: it is off for the call.
    const bool profiling = g->profiling;
    g->profiling = false;
    uint64_t sweeps = 0;
    gt_status rc = GT_OK;
    if (skip) {
        PrArgs b = a;
        b.x_next = x[0];
        b.c_next = c[0];
        dispatch_init<double>(S.plan, b, S.rank64, rank->had, n, 0, s);
        if (hipGetLastError() != hipSuccess) rc = fail(GT_ERR_HIP, "no-alloc pagerank: init launch");
    } else {
        rc = pr_phase<double>(S, a, x, c, rank->had, 0, params->epsilon, params->max_iter,
                              &sweeps);
    }
    g->profiling = profiling;
    GT_TRY(rc);
    g->last_sweeps = sweeps;
    *iters = sweeps;
    GT_TRY(pr_unpermute<double>(S, x[sweeps & 1]));
    return pr_deliver(g, S.rank64, rank);
}

}  // namespace gt

using namespace gt;

extern "C" {

gt_status gt_pagerank_f64_noalloc(gt_graph *g, const gt_view *view, const gt_prop *weight,
                                  const gt_prop *pers, double damping,
                                  const gt_power_iteration *params, gt_vertex_out *rank,
                                  uint64_t *iters) {
    if (weight && weight->kind != GT_MAP_UNITY) {
        clear_error();
        return fail(GT_ERR_UNSUPPORTED,
                    "no-alloc pagerank: weights need staging buffers; pass Unity (or null)");
    }
    return pagerank_f64_noalloc_impl(g, view, pers, damping, params, rank, iters);
}

gt_status gt_pagerank_f64(gt_graph *g, const gt_view *view, const gt_prop *weight,
                          const gt_prop *pers, double damping, const gt_power_iteration *params,
                          gt_vertex_out *rank, uint64_t *iters) {
    return pagerank_f64_impl(g, view, weight, pers, damping, params, rank, iters);
}

gt_status gt_pagerank_mixed(gt_graph *g, const gt_view *view, const gt_prop *weight,
                            const gt_prop *pers, double damping, const gt_power_iteration *params,
                            gt_vertex_out *rank, uint64_t *iters) {
    return pagerank_mixed_impl(g, view, weight, pers, damping, params, rank, 1, iters);
}

gt_status gt_graph_set_profiling(gt_graph *g, int32_t enable) {
    clear_error();
    if (!g) return fail(GT_ERR_INVALID_ARG, "null graph");
    g->profiling = enable != 0;
    for (int k = 0; k < GT_SWEEP_KINDS; ++k) {
        g->prof_ms[k] = 0.0;
        g->prof_launches[k] = 0;
    }
    return GT_OK;
}

gt_status gt_sweep_stats(const gt_graph *g, int32_t kind, double *total_ms, uint64_t *launches,
                         uint64_t *last_sweeps) {
    clear_error();
    if (!g) return fail(GT_ERR_INVALID_ARG, "null graph");
    if (kind < 0 || kind >= GT_SWEEP_KINDS) return fail(GT_ERR_INVALID_ARG, "unknown sweep kind");
    if (total_ms) *total_ms = g->prof_ms[kind];
    if (launches) *launches = g->prof_launches[kind];
    if (last_sweeps) *last_sweeps = g->last_sweeps;
    return GT_OK;
}

}  // extern "C"
