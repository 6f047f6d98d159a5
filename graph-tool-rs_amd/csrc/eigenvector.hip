// Eigenvector centrality by power iteration on gfx950, FP64.
//
// Semantics follow crates/gt-algo/src/centrality.rs:405-497: per sweep t,
// s(i) = sum over in-entries (j, e) of w(e) * x_{t-1}(j) for active i (inactive
// slots carried), norm_t = sqrt(sum s^2), x_t = s / (norm_t == 0 ? 1 : norm_t),
// delta_t = sum over active i of |x_t(i) - x_{t-1}(i)|; stop when
// delta_t < epsilon or t == max_iter; return (norm_T, T).
//
// One launch per sweep, normalisation folded: launch L writes the unscaled
// y_L = A x_{L-1} (computed as (A y_{L-1}) / scale_{L-1}) and, in the same pass,
// delta_{L-1} = sum |y_{L-1}/scale_{L-1} - y_{L-2}/scale_{L-2}|.  The scales
// are known at launch time, so no second vertex pass is needed; the price is
// that convergence of sweep L-1 is learnt in launch L (one extra SpMV at the
// end of a solve).  Three y buffers rotate.  A small second kernel per sweep
// (ev_fold) finishes the split rows, folds the squared-norm and delta
// partials in index order and decides on the device.
#include <cmath>

#include "device_common.hpp"
#include "gt_internal.hpp"
#include "props.hpp"
#include "blocks.hpp"
#include "driver.hpp"
#include "ev_state.hpp"
#include "sweep_common.hpp"

namespace gt {

struct EvArgs {
    const Tile *tiles;
    uint32_t ntiles;
    uint32_t nlong;
    const uint32_t *in_off;
    const uint32_t *packed;   // packed entry stream of the wave tiles
    uint32_t bits;            // slot bits of `packed`
    const uint32_t *tbase;    // source base of each tile
    uint32_t short_row;        // rows with fewer entries are summed by one lane
    const double *w;          // weights in packed order
    const uint8_t *active;
    const double *y_cur;      // y_{L-1}
    const double *y_prev;     // y_{L-2}
    double *y_next;           // y_L
    double epsilon;
    uint64_t max_iter;
    double *partials;         // kPartials: sweep partials, then fold partials
    double *long_part;
    const uint32_t *long_rows;
    const uint32_t *long_slot;
    EvState *st;
};

template <bool MASK>
struct EvFin : TileAll {
    const EvArgs &a;
    double sc_cur, sc_prev;
    bool first;
    double sq = 0.0, dl = 0.0;
    __device__ EvFin(const EvArgs &args, double c, double p, bool f)
        : a(args), sc_cur(c), sc_prev(p), first(f) {}
    // What row i reads: its last two iterates (the first launch has one and
    // reads it twice).
    static constexpr bool kLoadsFirst = true;
    struct In {
        double yc, yp;
        bool act;
    };
    __device__ __forceinline__ In load(uint32_t i) const {
        In in;
        in.yc = a.y_cur[i];
        in.yp = (first ? a.y_cur : a.y_prev)[i];
        in.act = MASK ? a.active[i] != 0 : true;
        return in;
    }
    __device__ __forceinline__ void finish(uint32_t i, double acc, const In &in) {
        if (MASK && !in.act) {
            a.y_next[i] = in.yc;
            return;
        }
        const double s = acc / sc_cur;
        a.y_next[i] = s;
        sq += s * s;
        if (!first) dl += fabs(in.yc / sc_cur - in.yp / sc_prev);
    }
    __device__ __forceinline__ void operator()(uint32_t i, double acc) { finish(i, acc, load(i)); }
};

// The SpMV of one sweep (split rows aside), with per-workgroup squared-norm
// and delta partials; ev_fold completes the sweep.
template <bool WEIGHTED, bool MASK>
__global__ __launch_bounds__(kBlock) void ev_sweep(const EvArgs a) {
    if (a.st->done) return;
    __shared__ WaveLds<double> L;
    const bool first = a.st->launches == 0;
    EvFin<MASK> fin(a, a.st->sc_cur, a.st->sc_prev, first);
    spmv_waves<double, WEIGHTED>(a.tiles, a.ntiles, a.in_off, a.packed, a.bits, a.tbase,
                                 a.y_cur, a.w, a.long_part, a.short_row, L, fin);
    double sq = fin.sq, dl = fin.dl;
    block_sum2(sq, dl, L.red);
    if (threadIdx.x == 0) {
        a.partials[blockIdx.x] = sq;
        a.partials[gridDim.x + blockIdx.x] = dl;
    }
}

// The sweep over a relabelled graph with the hot slice of y in LDS
// (kHotWaves wavefronts, one workgroup per compute unit).
template <bool WEIGHTED>
__global__ __launch_bounds__(kHotBlock) void ev_sweep_hot(const EvArgs a, uint32_t nhot) {
    if (a.st->done) return;
    __shared__ WaveLds<double, kHotWaves, false> L;
    __shared__ double hot[kHotMax64];
    fill_hot(a.y_cur, nhot, hot);
    const bool first = a.st->launches == 0;
    EvFin<false> fin(a, a.st->sc_cur, a.st->sc_prev, first);
    spmv_waves<double, WEIGHTED, true>(a.tiles, a.ntiles, a.in_off, a.packed, a.bits, a.tbase,
                                       a.y_cur, a.w, a.long_part, a.short_row, L, fin,
                                       static_cast<const double *>(hot), nhot);
    double sq = fin.sq, dl = fin.dl;
    block_sum2_waves<kHotWaves>(sq, dl, L.red);
    if (threadIdx.x == 0) {
        a.partials[blockIdx.x] = sq;
        a.partials[gridDim.x + blockIdx.x] = dl;
    }
}

// Completes the preceding ev_sweep launch (`sgrid` workgroups): split rows,
// then in the last workgroup the index-order fold and the decision.
template <bool WEIGHTED, bool MASK>
__global__ __launch_bounds__(kBlock) void ev_fold(const EvArgs a, uint32_t sgrid) {
    if (a.st->done) return;
    __shared__ double red[8];
    __shared__ uint32_t flag;
    const int tid = threadIdx.x;
    const bool first = a.st->launches == 0;
    EvFin<MASK> lf(a, a.st->sc_cur, a.st->sc_prev, first);
    for (uint32_t k = blockIdx.x * kBlock + tid; k < a.nlong; k += gridDim.x * kBlock)
        lf(a.long_rows[k], long_row_sum(a.long_part, a.long_slot, k));
    double s2 = lf.sq, d = lf.dl;
    block_sum2(s2, d, red);
    double *fp = a.partials + 2 * kMaxSweepGrid;
    if (tid == 0) {
        fp[blockIdx.x] = s2;
        fp[gridDim.x + blockIdx.x] = d;
    }
    if (!last_to_arrive(&a.st->counter, &flag)) return;
    s2 = 0.0;
    d = 0.0;
    for (uint32_t k = tid; k < sgrid; k += kBlock) {
        s2 += a.partials[k];
        d += a.partials[sgrid + k];
    }
    for (uint32_t k = tid; k < gridDim.x; k += kBlock) {
        s2 += fp[k];
        d += fp[gridDim.x + k];
    }
    block_sum2(s2, d, red);
    if (tid == 0) ev_close(a.st, s2, d, a.epsilon, a.max_iter);
}

// y0[k] = the caller's slot iperm[k] (identity when null), with slots
// [had, n) seeded with the uniform start; resets the loop state.
__global__ __launch_bounds__(kBlock) void ev_init(const double *x, const uint32_t *iperm,
                                                  uint64_t had, uint32_t n, double uniform,
                                                  double *y0, EvState *st) {
    for (uint32_t k = blockIdx.x * kBlock + threadIdx.x; k < n; k += gridDim.x * kBlock) {
        const uint32_t o = iperm ? iperm[k] : k;
        y0[k] = o >= had ? uniform : x[o];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->sc_cur = 1.0;
        st->sc_prev = 1.0;
        st->norm_last = 0.0;
        st->delta = 0.0;
        st->eig = 0.0;
        st->iter = 0;
        st->done = 0;
        st->counter = 0;
        st->launches = 0;
        st->final_scale = 1.0;
        st->final_iter = 0;
        st->bail = 0;
    }
}

// x(o) = y_T(perm[o]) / scale for active o; inactive slots are carried
// unscaled.  perm null: identity.
template <bool MASK>
__global__ __launch_bounds__(kBlock) void ev_emit(const double *y, const uint8_t *active,
                                                  const uint32_t *perm, const EvState *st,
                                                  uint32_t n, double *x) {
    const double sc = st->final_scale;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const double v = y[perm ? perm[i] : i];
        x[i] = (MASK && !active[i]) ? v : v / sc;
    }
}

template <bool W, bool M>
static int ev_launch(const EvArgs &a, int per_cu, hipStream_t s) {
    const int sg = wave_grid(occupancy_cap((const void *)ev_sweep<W, M>, per_cu), a.ntiles);
    hipLaunchKernelGGL((ev_sweep<W, M>), dim3(sg), dim3(kBlock), 0, s, a);
    return sg;
}
template <bool W>
static int ev_launch_hot(const EvArgs &a, uint32_t nhot, hipStream_t s) {
    const int sg = fit_grid(occupancy_cap((const void *)ev_sweep_hot<W>, 1),
                            (a.ntiles + kHotWaves - 1) / kHotWaves);
    hipLaunchKernelGGL((ev_sweep_hot<W>), dim3(sg), dim3(kHotBlock), 0, s, a, nhot);
    return sg;
}
template <bool W, bool M>
static void ev_launch_fold(const EvArgs &a, int sg, hipStream_t s) {
    hipLaunchKernelGGL((ev_fold<W, M>), dim3(fold_grid(a.nlong)), dim3(kBlock), 0, s, a,
                       (uint32_t)sg);
}

static unsigned vgrid(uint32_t n) {
    uint32_t b = (n + kBlock - 1) / kBlock;
    if (b > 2048) b = 2048;
    return b ? b : 1;
}

gt_status eigenvector_impl(gt_graph *g, const gt_view *view, const gt_prop *weight,
                           const gt_power_iteration *params, gt_vertex_out *x, double *eig,
                           uint64_t *iters) {
    clear_error();
    if (!g) return fail(GT_ERR_INVALID_ARG, "null graph");
    if (!iters || !eig) return fail(GT_ERR_INVALID_ARG, "null output");
    Orient *pull = nullptr, *push = nullptr;
    bool reversed = false;
    GT_TRY(check_vertex_out(g, x, "eigenvector: centrality"));
    GT_TRY(check_params(g, params));
    const uint8_t *active = nullptr;
    uint32_t nactive = 0;
    GT_TRY(resolve_view(g, view, &pull, &push, &reversed, &active, &nactive));
    const bool masked = active != nullptr;
    const uint32_t n = g->n;
    hipStream_t s = g->stream;
    Workspace &ws = g->ws;
    // New slots get 1 / num_vertices of the view (centrality.rs:418-429).
    const double uniform = nactive == 0 ? 0.0 : 1.0 / (double)nactive;
    const double eps = params->epsilon;
    const bool skip = std::isnan(eps) || !(eps + 1.0 >= eps);
    const uint32_t *perm = nullptr, *iperm = nullptr;
    Relabeled *R = nullptr;
    if (!masked && n > 0) {
        GT_TRY(ensure_relabeled(g, reversed ? 1 : 0, &R));
        pull = &R->pull;
        perm = R->perm.as<uint32_t>();
        iperm = R->iperm.as<uint32_t>();
    }

    bool weighted = false;
    const double *wp = nullptr;
    const void *wsrc = nullptr;
    int wdt = GT_F64;
    if (weight && weight->kind != GT_MAP_UNITY) {
        weighted = true;
        GT_TRY(stage_edge_source(g, weight, &wsrc, &wdt));
    }
    *eig = 0.0;
    if (n == 0) {
        *iters = skip ? 0 : 1;
        g->last_sweeps = *iters;
        return GT_OK;
    }
    GT_TRY(ensure_schedule(g, *pull, n));
    note_layout(g, *pull);
    GT_TRY(prof_prepare(g));
    GT_TRY(ws.partials.reserve(sizeof(double) * kPartials));
    GT_TRY(ws.long_part.reserve(sizeof(double) * std::max<uint32_t>(1, pull->nslots)));
    GT_TRY(ws.state.reserve(sizeof(EvState)));
    const size_t vb = sizeof(double) * n;
    // xio: the caller's map on the device (their buffer, or a staged copy).
    double *xio = nullptr;
    if (x->loc == GT_DEVICE) {
        xio = x->data;
    } else {
        GT_TRY(ws.y3.reserve(vb));
        xio = ws.y3.as<double>();
        const uint64_t had = std::min<uint64_t>(x->had, n);
        if (had) GT_HIP(hipMemcpyAsync(xio, x->data, sizeof(double) * had, hipMemcpyHostToDevice, s));
    }
    double *y[3];
    GT_TRY(ws.y[0].reserve(vb));
    GT_TRY(ws.y[1].reserve(vb));
    GT_TRY(ws.y[2].reserve(vb));
    y[0] = ws.y[0].as<double>();
    y[1] = ws.y[1].as<double>();
    y[2] = ws.y[2].as<double>();
    EvState *st = ws.state.as<EvState>();
    hipLaunchKernelGGL(ev_init, dim3(vgrid(n)), dim3(kBlock), 0, s, xio, iperm, x->had, n, uniform,
                       y[0], st);
    GT_HIP(hipGetLastError());

    uint64_t T = 0;
    double *out = y[0];
    bool on_blocks = false;
    if (!skip && R && g->tune.block_tile_nnz > 0) {
        // Block tiles (blocks.hpp): the sweep of a relabelled graph, its row
        // sums in two-word fixed point, weighted by the map in block order.
        // The loop runs on the wave engine instead when a row is too long
        // for the low words, a weight is negative or not finite, or a launch
        // meets an iterate whose terms the fixed point cannot hold (its mass
        // not finite or above kBlkMassMax).
        Blocks *B = nullptr;
        GT_TRY(ensure_blocks(g, *R, (uint32_t)g->tune.block_tile_nnz,
                             (uint32_t)g->tune.block_long_min, &B));
        bool usable = B->max_row <= kBlkF64MaxRow;
        if (usable && weighted)
            GT_TRY(blk_stage_weights(g, *B, weight, wsrc, wdt, true, ws.blk_w64, &usable));
        if (usable) {
            GT_TRY(blk_reserve(g, *B));
            BlkArgs b = blk_layout_args(g, *B);
            if (weighted) b.wblk = ws.blk_w64.ptr;
            b.st = reinterpret_cast<SweepState *>(st);
            b.ev = st;
            b.eps_abs = eps;
            b.max_iter = params->max_iter;
            b.x_cur64 = y[0];
            blk_launch_start_ev(b, s);
            GT_HIP(hipGetLastError());
            auto sweep = [&](uint64_t k) {
                // Launch L = k + 1 reads y_{L-1}, y_{L-2}, writes y_L.
                const uint64_t L = k + 1;
                b.x_cur64 = y[(L - 1) % 3];
                b.c_cur64 = b.x_cur64;
                b.x_prev64 = k == 0 ? b.x_cur64 : y[(L + 1) % 3];
                b.x_next64 = y[L % 3];
                blk_launch_sweep(b, false, 1, kBlkEv, s);
            };
            auto fold = [&](uint64_t) { blk_launch_fold(b, false, kBlkEv, s); };
            uint64_t launched = 0;
            g->prof_kind = GT_SWEEP_EIGENVECTOR;
            GT_TRY(drive(g, reinterpret_cast<SweepState *>(st), eps, params->max_iter, sweep, fold,
                         &launched, GT_SWEEP_EIGENVECTOR));
            EvState h;
            GT_HIP(hipMemcpy(&h, st, sizeof(EvState), hipMemcpyDeviceToHost));
            if (!h.bail) {
                on_blocks = true;
                T = h.final_iter;
                *eig = h.eig;
                out = y[T % 3];
            } else {
                g->layout.block_tiles = 0;
                g->layout.block_long_rows = 0;
                g->layout.block_groups = 0;
                hipLaunchKernelGGL(ev_init, dim3(vgrid(n)), dim3(kBlock), 0, s, xio, iperm,
                                   x->had, n, uniform, y[0], st);
                GT_HIP(hipGetLastError());
            }
        }
    }
    if (!skip && !on_blocks && weighted) {
        // The wave engine's weights, in its packed order.
        GT_TRY(stage_pull_weights<double>(g, weight, *pull, wsrc, wdt, ws.wcsc64));
        wp = ws.wcsc64.as<double>();
    }
    if (!skip && !on_blocks) {
        EvArgs a{};
        a.tiles = pull->tiles.as<Tile>();
        a.ntiles = pull->ntiles;
        a.nlong = pull->nlong;
        a.in_off = pull->off.as<uint32_t>();
        a.packed = pull->packed.as<uint32_t>();
        a.bits = pull->bits;
        a.tbase = pull->tbase.as<uint32_t>();
        a.short_row = (uint32_t)g->tune.short_row;
        a.w = wp;
        a.active = active;
        a.epsilon = eps;
        a.max_iter = params->max_iter;
        a.partials = ws.partials.as<double>();
        a.long_part = ws.long_part.as<double>();
        a.long_rows = pull->long_rows.as<uint32_t>();
        a.long_slot = pull->long_slot.as<uint32_t>();
        a.st = st;
        EvArgs b = a;
        int sg = 0;
        const int pc = g->tune.sweep_blocks_per_cu;
        // Sweeps in relabelled order keep the most gathered entries of y in
        // LDS.
        const uint32_t nhot = iperm && !masked ? hot_len<double>(g->tune.sweep_hot_kb, n) : 0u;
        auto sweep = [&](uint64_t k) {
            // Launch L = k + 1 reads y_{L-1}, y_{L-2}, writes y_L.
            const uint64_t L = k + 1;
            b.y_cur = y[(L - 1) % 3];
            b.y_prev = y[(L + 1) % 3];
            b.y_next = y[L % 3];
            if (nhot)
                sg = weighted ? ev_launch_hot<true>(b, nhot, s) : ev_launch_hot<false>(b, nhot, s);
            else if (weighted && masked) sg = ev_launch<true, true>(b, pc, s);
            else if (weighted) sg = ev_launch<true, false>(b, pc, s);
            else if (masked) sg = ev_launch<false, true>(b, pc, s);
            else sg = ev_launch<false, false>(b, pc, s);
        };
        auto fold = [&](uint64_t) {
            if (weighted && masked) ev_launch_fold<true, true>(b, sg, s);
            else if (weighted) ev_launch_fold<true, false>(b, sg, s);
            else if (masked) ev_launch_fold<false, true>(b, sg, s);
            else ev_launch_fold<false, false>(b, sg, s);
        };
        // The driver reads done / delta / iter through the SweepState layout;
        // EvState keeps them at the same offsets.
        uint64_t launched = 0;
        g->prof_kind = GT_SWEEP_EIGENVECTOR;
        // One launch beyond the answer is needed to learn its delta.
        const uint64_t cap = params->max_iter;
        GT_TRY(drive(g, reinterpret_cast<SweepState *>(st), eps, cap, sweep, fold, &launched,
                     GT_SWEEP_EIGENVECTOR));
        EvState h;
        GT_HIP(hipMemcpy(&h, st, sizeof(EvState), hipMemcpyDeviceToHost));
        T = h.final_iter;
        *eig = h.eig;
        out = y[T % 3];
    }
    // xio never aliases a y buffer, so the permuting copy has no hazard.
    double *dst = xio;
    if (masked)
        hipLaunchKernelGGL((ev_emit<true>), dim3(vgrid(n)), dim3(kBlock), 0, s, out, active,
                           nullptr, st, n, dst);
    else
        hipLaunchKernelGGL((ev_emit<false>), dim3(vgrid(n)), dim3(kBlock), 0, s, out, nullptr,
                           perm, st, n, dst);
    GT_HIP(hipGetLastError());
    if (x->loc == GT_HOST)
        GT_HIP(hipMemcpyAsync(x->data, dst, vb, hipMemcpyDeviceToHost, s));
    GT_HIP(hipStreamSynchronize(s));
    g->last_sweeps = T;
    *iters = T;
    return GT_OK;
}

}  // namespace gt

extern "C" gt_status gt_eigenvector_f64(gt_graph *g, const gt_view *view, const gt_prop *weight,
                                        const gt_power_iteration *params, gt_vertex_out *x,
                                        double *eig, uint64_t *iters) {
    return gt::eigenvector_impl(g, view, weight, params, x, eig, iters);
}
