// Adjoint of the PageRank fixed point (extension: no counterpart in the
// reference).  For x = F(x) = (1-d) p + d (M x + (q.x) p), M[t,s] = w_e/deg_s
// over edges e: s -> t, q the sink indicator, and a loss with dL/dx = c:
//
//   lambda = c + d M^T lambda + d (p.lambda) q        (transposed system)
//   dL/dw_e = d * x_s / deg_s * (lambda_t - mu_s),     mu = M^T lambda
//   dL/dp   = ((1-d) + d (q.x)) lambda
//   dL/dd   = lambda . (x - p) / d
//
// The transposed system is solved by the same power iteration, pulling over
// the opposite orientation (relabelled like the forward sweep's, the most
// gathered targets first; lambda and mu are mapped back to the caller's
// labels for the gradient passes), until its L1 change drops below
// epsilon * ||dL/dx||_1 (the system is linear in dL/dx, so the tolerance
// scales with it; FP32 runs stop no later than 2^-22 * ||dL/dx||_1, below
// which an FP32 iterate's change is rounding noise): row s of the out-halves gathers lambda at its
// targets (times the out-ordered weight), so the wave engine and the fold
// kernel of the forward sweep are reused unchanged.  No iterate of the
// forward solve is stored: only x* is read.  mu = M^T lambda comes from one
// more pass of the same engine; the edge gradients are then one coalesced
// pass over edge ids, reading each edge's endpoints from the graph's edge
// list.
#include <cmath>
#include <type_traits>

#include "blocks.hpp"
#include "device_common.hpp"
#include "driver.hpp"
#include "gt_internal.hpp"
#include "props.hpp"
#include "sweep_common.hpp"

namespace gt {

struct AdjArgs {
    const Tile *tiles;
    uint32_t ntiles;
    uint32_t nlong;
    const uint32_t *row_off;   // push orientation (rows = sources)
    const uint32_t *packed;    // packed targets of the push orientation's wave tiles
    uint32_t bits;             // slot bits of `packed`
    const uint32_t *tbase;     // source base of each tile
    uint32_t short_row;        // rows with fewer entries are summed by one lane
    const void *w;             // weights in packed order (value type)
    const double *deg;         // weighted out-degree (weighted)
    const double *cvec;        // dL/dx, f64
    const double *pers;        // dense personalisation
    double pval;
    double damping;
    double eps;                // stop when delta < max(eps, eps_rel) * ||dL/dx||_1
    double eps_rel;
    uint64_t max_iter;
    const void *l_cur;
    void *l_next;
    double *partials;
    double *long_part;
    const uint32_t *long_rows;
    const uint32_t *long_slot;
    SweepState *st;            // p_sink holds p . lambda of the current iterate
};

template <typename V, bool WEIGHTED, bool DPERS>
struct AdjFin : TileAll {
    const AdjArgs &a;
    double pl;
    double dsum = 0.0;
    double psum = 0.0;
    __device__ AdjFin(const AdjArgs &args, double p) : a(args), pl(p) {}
    // What row s reads: its degree, dL/dx, lambda and personalisation.
    static constexpr bool kLoadsFirst = true;
    struct In {
        double dg, cv, pv;
        V lc;
    };
    __device__ __forceinline__ In load(uint32_t s) const {
        In in;
        in.dg = WEIGHTED ? a.deg[s] : static_cast<double>(a.row_off[s + 1] - a.row_off[s]);
        in.cv = a.cvec[s];
        in.pv = DPERS ? a.pers[s] : a.pval;
        in.lc = static_cast<const V *>(a.l_cur)[s];
        return in;
    }
    __device__ __forceinline__ void finish(uint32_t s, double acc, const In &in) {
        V *ln = static_cast<V *>(a.l_next);
        const double mu = in.dg != 0.0 ? acc / in.dg : 0.0;
        const double v = in.cv + a.damping * mu + (in.dg == 0.0 ? a.damping * pl : 0.0);
        const V vv = static_cast<V>(v);
        dsum += fabs(static_cast<double>(vv) - static_cast<double>(in.lc));
        ln[s] = vv;
        psum += in.pv * static_cast<double>(vv);
    }
    __device__ __forceinline__ void operator()(uint32_t s, double acc) { finish(s, acc, load(s)); }
};

// The SpMV of one transposed sweep (split rows aside), with per-workgroup
// delta and p.lambda partials; adj_fold completes the sweep.
template <typename V, bool WEIGHTED, bool DPERS>
__global__ __launch_bounds__(kBlock) void adj_sweep(const AdjArgs a) {
    if (a.st->done) return;
    __shared__ WaveLds<V> L;
    AdjFin<V, WEIGHTED, DPERS> fin(a, a.st->p_sink);
    spmv_waves<V, WEIGHTED>(a.tiles, a.ntiles, a.row_off, a.packed, a.bits, a.tbase,
                            static_cast<const V *>(a.l_cur), static_cast<const V *>(a.w),
                            a.long_part, a.short_row, L, fin);
    double d = fin.dsum, p = fin.psum;
    block_sum2(d, p, L.red);
    if (threadIdx.x == 0) {
        a.partials[blockIdx.x] = d;
        a.partials[gridDim.x + blockIdx.x] = p;
    }
}

// The FP32 transposed sweep with the hot slice of lambda in LDS (one
// workgroup of kHotWaves wavefronts per compute unit).
template <bool WEIGHTED, bool DPERS>
__global__ __launch_bounds__(kHotBlock) void adj_sweep_hot(const AdjArgs a, uint32_t nhot) {
    if (a.st->done) return;
    __shared__ WaveLds<float, kHotWaves, false> L;
    __shared__ float hot[kHotMax];
    const float *lc = static_cast<const float *>(a.l_cur);
    fill_hot(lc, nhot, hot);
    AdjFin<float, WEIGHTED, DPERS> fin(a, a.st->p_sink);
    spmv_waves<float, WEIGHTED, true>(a.tiles, a.ntiles, a.row_off, a.packed, a.bits, a.tbase, lc,
                                      static_cast<const float *>(a.w), a.long_part, a.short_row,
                                      L, fin, static_cast<const float *>(hot), nhot);
    double d = fin.dsum, p = fin.psum;
    block_sum2_waves<kHotWaves>(d, p, L.red);
    if (threadIdx.x == 0) {
        a.partials[blockIdx.x] = d;
        a.partials[gridDim.x + blockIdx.x] = p;
    }
}

// Completes the preceding adj_sweep launch (`sgrid` workgroups): split rows,
// then in the last workgroup the index-order fold and the decision.
template <typename V, bool WEIGHTED, bool DPERS>
__global__ __launch_bounds__(kBlock) void adj_fold(const AdjArgs a, uint32_t sgrid) {
    if (a.st->done) return;
    __shared__ double red[8];
    __shared__ uint32_t flag;
    const int tid = threadIdx.x;
    AdjFin<V, WEIGHTED, DPERS> lf(a, a.st->p_sink);
    for (uint32_t k = blockIdx.x * kBlock + tid; k < a.nlong; k += gridDim.x * kBlock)
        lf(a.long_rows[k], long_row_sum(a.long_part, a.long_slot, k));
    double d = lf.dsum, p = lf.psum;
    block_sum2(d, p, red);
    double *fp = a.partials + 2 * kMaxSweepGrid;
    if (tid == 0) {
        fp[blockIdx.x] = d;
        fp[gridDim.x + blockIdx.x] = p;
    }
    if (!last_to_arrive(&a.st->counter, &flag)) return;
    d = 0.0;
    p = 0.0;
    for (uint32_t k = tid; k < sgrid; k += kBlock) {
        d += a.partials[k];
        p += a.partials[sgrid + k];
    }
    for (uint32_t k = tid; k < gridDim.x; k += kBlock) {
        d += fp[k];
        p += fp[gridDim.x + k];
    }
    block_sum2(d, p, red);
    if (tid == 0) {
        SweepState *st = a.st;
        const uint64_t it = st->iter + 1;
        st->delta = d;
        st->p_sink = p;
        st->iter = it;
        st->launches += 1;
        const double thr = fmax(a.eps, a.eps_rel) * (st->aux > 0.0 ? st->aux : 1.0);
        st->done = (!(d >= thr) || (a.max_iter > 0 && it >= a.max_iter)) ? 1u : 0u;
        st->counter = 0;
    }
}

// lambda_0 = c; p . lambda_0 into the state.
template <typename V, bool DPERS>
__global__ __launch_bounds__(kBlock) void adj_init(const AdjArgs a, uint32_t n) {
    __shared__ double red[8];
    __shared__ uint32_t flag;
    V *l0 = static_cast<V *>(a.l_next);
    double p = 0.0, z = 0.0;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const V v = static_cast<V>(a.cvec[i]);
        l0[i] = v;
        p += (DPERS ? a.pers[i] : a.pval) * static_cast<double>(v);
        z += fabs(a.cvec[i]);
    }
    block_sum2(p, z, red);
    if (threadIdx.x == 0) {
        a.partials[blockIdx.x] = p;
        a.partials[gridDim.x + blockIdx.x] = z;
    }
    if (!last_to_arrive(&a.st->counter, &flag)) return;
    double t = 0.0, c1 = 0.0;
    for (uint32_t k = threadIdx.x; k < gridDim.x; k += kBlock) {
        t += a.partials[k];
        c1 += a.partials[gridDim.x + k];
    }
    block_sum2(t, c1, red);
    if (threadIdx.x == 0) {
        SweepState *st = a.st;
        *st = SweepState{};
        st->p_sink = t;
        st->aux = c1;  // ||dL/dx||_1, the scale of the stopping rule
    }
}

// mu_s = (M^T lambda)_s = (sum over s's out-entries of w_e lambda_t) / deg_s
// (0 when deg_s == 0): the same wave-engine pass as a transposed sweep,
// written to mu instead of iterating.
template <typename V, bool WEIGHTED>
struct MuFin : TileAll {
    const AdjArgs &a;
    double *mu;
    __device__ MuFin(const AdjArgs &args, double *m) : a(args), mu(m) {}
    __device__ __forceinline__ void operator()(uint32_t s, double acc) {
        const double dg =
            WEIGHTED ? a.deg[s] : static_cast<double>(a.row_off[s + 1] - a.row_off[s]);
        mu[s] = dg != 0.0 ? acc / dg : 0.0;
    }
};

template <typename V, bool WEIGHTED>
__global__ __launch_bounds__(kBlock) void adj_mu(const AdjArgs a, const V *lam, double *mu) {
    __shared__ WaveLds<V> L;
    MuFin<V, WEIGHTED> fin(a, mu);
    spmv_waves<V, WEIGHTED>(a.tiles, a.ntiles, a.row_off, a.packed, a.bits, a.tbase, lam,
                            static_cast<const V *>(a.w), a.long_part, a.short_row, L, fin);
}

// The split rows of adj_mu, one thread each.
template <typename V, bool WEIGHTED>
__global__ __launch_bounds__(kBlock) void adj_mu_long(const AdjArgs a, double *mu) {
    MuFin<V, WEIGHTED> fin(a, mu);
    for (uint32_t k = blockIdx.x * kBlock + threadIdx.x; k < a.nlong; k += gridDim.x * kBlock)
        fin(a.long_rows[k], long_row_sum(a.long_part, a.long_slot, k));
}

// dL/dw_e = d * x_s / deg_s * (lambda_t - mu_s) for every edge id e < bound,
// (s, t) = (row, entry) of e in the pull orientation's opposite (es / et:
// the edge list, swapped for a reversed view).  Ids with no edge, and edges
// whose deg_s is 0, get 0.  Reads and writes run in edge-id order.
//
// Per source the edge pass needs d x_s / deg_s and mu_s: grad_coef packs
// them into one 16-byte pair per vertex (in the caller's labels; mu arrives
// in the relabelled order), so an edge gathers that pair and lambda_t -- two
// gathers instead of five.  coef is (d x_s) / deg_s, the operations and
// order of the single formula; a source with deg_s = 0 has coef 0 (and mu
// 0), so its edges get 0 for every finite lambda.
template <bool WEIGHTED>
__global__ __launch_bounds__(kBlock) void grad_coef(const double *__restrict__ deg,
                                                    const uint32_t *__restrict__ row_off,
                                                    const double *__restrict__ mu_rel,
                                                    const uint32_t *__restrict__ perm,
                                                    const double *__restrict__ x, double damping,
                                                    uint32_t n, double2 *__restrict__ pair) {
    for (uint32_t v = blockIdx.x * kBlock + threadIdx.x; v < n; v += gridDim.x * kBlock) {
        const double dg = WEIGHTED ? deg[v] : static_cast<double>(row_off[v + 1] - row_off[v]);
        pair[v] = make_double2(dg != 0.0 ? damping * x[v] / dg : 0.0, mu_rel[perm[v]]);
    }
}

template <typename V, typename G>
__global__ __launch_bounds__(kBlock) void grad_edges(const uint32_t *__restrict__ es,
                                                     const uint32_t *__restrict__ et,
                                                     const double2 *__restrict__ pair,
                                                     const V *__restrict__ lam, uint64_t bound,
                                                     G *__restrict__ grad) {
    for (uint64_t e = (uint64_t)blockIdx.x * kBlock + threadIdx.x; e < bound;
         e += (uint64_t)gridDim.x * kBlock) {
        const uint32_t sv = es[e];
        G out = G(0);
        if (sv != kNoEdge) {
            const double2 p = pair[sv];
            out = static_cast<G>(p.x * (static_cast<double>(lam[et[e]]) - p.y));
        }
        grad[e] = out;
    }
}

// Sums for the vertex gradients: p_sink* = sum over sinks of x, and
// lambda . (x - p).  Folded by the last workgroup into st->aux / st->norm.
template <typename V, bool WEIGHTED, bool DPERS>
__global__ __launch_bounds__(kBlock) void grad_sums(const AdjArgs a, const V *lam,
                                                    const double *x, uint32_t n) {
    __shared__ double red[8];
    __shared__ uint32_t flag;
    double sk = 0.0, ld = 0.0;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const double dg =
            WEIGHTED ? a.deg[i] : static_cast<double>(a.row_off[i + 1] - a.row_off[i]);
        if (dg == 0.0) sk += x[i];
        ld += static_cast<double>(lam[i]) * (x[i] - (DPERS ? a.pers[i] : a.pval));
    }
    block_sum2(sk, ld, red);
    if (threadIdx.x == 0) {
        a.partials[blockIdx.x] = sk;
        a.partials[gridDim.x + blockIdx.x] = ld;
    }
    if (!last_to_arrive(&a.st->counter, &flag)) return;
    double s = 0.0, l = 0.0;
    for (uint32_t k = threadIdx.x; k < gridDim.x; k += kBlock) {
        s += a.partials[k];
        l += a.partials[gridDim.x + k];
    }
    block_sum2(s, l, red);
    if (threadIdx.x == 0) {
        a.st->aux = s;
        a.st->norm = l;
        a.st->counter = 0;
        __threadfence();
    }
}

template <typename V>
__global__ __launch_bounds__(kBlock) void grad_pers_k(const V *lam, const SweepState *st,
                                                      double damping, uint32_t n, double *gp) {
    const double f = (1.0 - damping) + damping * st->aux;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
        gp[i] = f * static_cast<double>(lam[i]);
}

// out[i] = in[idx[i]]: with idx = iperm a caller-label vector into relabelled
// order, with idx = perm back.
template <typename T>
__global__ __launch_bounds__(kBlock) void adj_gather(const T *__restrict__ in,
                                                     const uint32_t *__restrict__ idx, uint32_t n,
                                                     T *__restrict__ out) {
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
        out[i] = in[idx[i]];
}

__global__ __launch_bounds__(kBlock) void f32_to_f64_v(const float *a, uint32_t n, double *b) {
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
        b[i] = static_cast<double>(a[i]);
}

template <typename V, bool W, bool P>
static int adj_launch(const AdjArgs &a, int per_cu, hipStream_t s) {
    const int sg = wave_grid(occupancy_cap((const void *)adj_sweep<V, W, P>, per_cu), a.ntiles);
    hipLaunchKernelGGL((adj_sweep<V, W, P>), dim3(sg), dim3(kBlock), 0, s, a);
    return sg;
}
template <bool W, bool P>
static int adj_launch_hot(const AdjArgs &a, uint32_t nhot, hipStream_t s) {
    const int sg = fit_grid(occupancy_cap((const void *)adj_sweep_hot<W, P>, 1),
                            (a.ntiles + kHotWaves - 1) / kHotWaves);
    hipLaunchKernelGGL((adj_sweep_hot<W, P>), dim3(sg), dim3(kHotBlock), 0, s, a, nhot);
    return sg;
}
template <typename V, bool W, bool P>
static void adj_launch_fold(const AdjArgs &a, int sg, hipStream_t s) {
    hipLaunchKernelGGL((adj_fold<V, W, P>), dim3(fold_grid(a.nlong)), dim3(kBlock), 0, s, a,
                       (uint32_t)sg);
}

// Grid of the vertex passes that end in a ticket (adj_init, grad_sums).
static unsigned vgrid(uint64_t n) {
    uint64_t b = (n + kBlock - 1) / kBlock;
    if (b > kTicketGrid) b = kTicketGrid;
    return b ? (unsigned)b : 1u;
}

// A vertex map (f64/f32, host/device) as a device f64 array.
static gt_status stage_vec(gt_graph *g, const gt_prop *p, DevBuf &buf, const double **out,
                           const char *what) {
    const uint32_t n = g->n;
    if (!p || p->kind != GT_MAP_DENSE || (!p->data && n))
        return fail(GT_ERR_INVALID_ARG, std::string(what) + " must be a dense vertex map");
    if (p->len < n) return fail(GT_ERR_INVALID_ARG, std::string(what) + " is shorter than the vertex bound");
    if (p->dtype == GT_F64 && p->loc == GT_DEVICE) {
        *out = static_cast<const double *>(p->data);
        return GT_OK;
    }
    GT_TRY(buf.reserve(sizeof(double) * std::max<uint32_t>(1, n)));
    if (p->dtype == GT_F64) {
        GT_HIP(hipMemcpyAsync(buf.ptr, p->data, sizeof(double) * n, hipMemcpyHostToDevice, g->stream));
    } else if (p->dtype == GT_F32) {
        const float *f = static_cast<const float *>(p->data);
        if (p->loc == GT_HOST) {
            std::vector<double> tmp(f, f + n);
            GT_HIP(hipMemcpyAsync(buf.ptr, tmp.data(), sizeof(double) * n, hipMemcpyHostToDevice,
                                  g->stream));
            GT_HIP(hipStreamSynchronize(g->stream));
        } else {
            hipLaunchKernelGGL(f32_to_f64_v, dim3(vgrid(n)), dim3(kBlock), 0, g->stream, f, n,
                               buf.as<double>());
        }
    } else {
        return fail(GT_ERR_INVALID_ARG, std::string(what) + " dtype must be f64 or f32");
    }
    *out = buf.as<double>();
    return GT_OK;
}

template <typename V>
static gt_status backward_impl(gt_graph *g, const gt_view *view, const gt_prop *weight,
                               const gt_prop *pers, double damping,
                               const gt_power_iteration *params, const gt_prop *xm,
                               const gt_prop *cm, void *grad_w, int32_t gw_dtype,
                               int32_t grad_loc, double *grad_pers, double *grad_damping,
                               uint64_t *iters) {
    GT_TRY(check_params(g, params));
    Orient *pull = nullptr, *push = nullptr;
    bool reversed = false;
    const uint8_t *active = nullptr;
    uint32_t nactive = 0;
    GT_TRY(resolve_view(g, view, &pull, &push, &reversed, &active, &nactive));
    if (active) return fail(GT_ERR_UNSUPPORTED, "the adjoint on a filtered view");
    if (grad_w && gw_dtype != GT_F64 && gw_dtype != GT_F32)
        return fail(GT_ERR_INVALID_ARG, "grad_w dtype must be f64 or f32");
    if (grad_loc != GT_HOST && grad_loc != GT_DEVICE)
        return fail(GT_ERR_INVALID_ARG, "bad location");
    const uint32_t n = g->n;
    hipStream_t s = g->stream;
    Workspace &ws = g->ws;
    AdjArgs a{};
    const double uniform = n == 0 ? 0.0 : 1.0 / (double)n;
    bool dpers = false;
    GT_TRY(stage_vertex_prop(g, pers, ws.pers, &a.pers, &a.pval, &dpers, uniform));
    const double *x = nullptr, *c = nullptr;
    GT_TRY(stage_vec(g, xm, ws.ax, &x, "x"));
    GT_TRY(stage_vec(g, cm, ws.ac, &c, "dL/dx"));
    const bool weighted = weight && weight->kind != GT_MAP_UNITY;
    const void *wsrc = nullptr;
    int wdt = GT_F64;
    if (weighted) {
        GT_TRY(stage_edge_source(g, weight, &wsrc, &wdt));
        GT_TRY(stage_degree(g, weight, *push, wsrc, wdt, ws.deg));
        a.deg = ws.deg.as<double>();
    }
    if (n == 0) {
        if (grad_damping) *grad_damping = 0.0;
        if (iters) *iters = 0;
        return GT_OK;
    }
    // The transposed system and mu run over the relabelled push orientation:
    // targets renumbered by how often they are gathered (the relabelling of
    // the opposite view's PageRank), rows in the same order.  dL/dx, the
    // degrees and the personalisation are gathered into that order; lambda
    // and mu come back to the caller's labels for the gradient passes.
    Relabeled *R = nullptr;
    GT_TRY(ensure_relabeled(g, reversed ? 0 : 1, &R));
    Orient &rp = R->pull;
    const uint32_t *perm = R->perm.as<uint32_t>();
    const uint32_t *iperm = R->iperm.as<uint32_t>();
    GT_TRY(ensure_schedule(g, rp, n));
    note_layout(g, rp);
    GT_TRY(prof_prepare(g));
    GT_TRY(ws.partials.reserve(sizeof(double) * kPartials));
    GT_TRY(ws.long_part.reserve(sizeof(double) * std::max<uint32_t>(1, rp.nslots)));
    GT_TRY(ws.state.reserve(sizeof(SweepState) * 2));
    GT_TRY(ws.lam[0].reserve(sizeof(V) * n));
    GT_TRY(ws.lam[1].reserve(sizeof(V) * n));
    V *lam[2] = {ws.lam[0].as<V>(), ws.lam[1].as<V>()};
    // Caller-label arguments (the gradient passes).
    a.row_off = push->off.as<uint32_t>();
    a.short_row = (uint32_t)g->tune.short_row;
    a.cvec = c;
    a.damping = damping;
    a.eps = params->epsilon;
    a.eps_rel = std::is_same<V, float>::value ? std::ldexp(1.0, -22) : 0.0;
    a.max_iter = params->max_iter;
    a.partials = ws.partials.as<double>();
    a.long_part = ws.long_part.as<double>();
    a.st = ws.state.as<SweepState>();
    // Relabelled arguments (the sweeps and mu).
    AdjArgs r = a;
    r.tiles = rp.tiles.as<Tile>();
    r.ntiles = rp.ntiles;
    r.nlong = rp.nlong;
    r.row_off = rp.off.as<uint32_t>();
    r.packed = rp.packed.as<uint32_t>();
    r.bits = rp.bits;
    r.tbase = rp.tbase.as<uint32_t>();
    r.long_rows = rp.long_rows.as<uint32_t>();
    r.long_slot = rp.long_slot.as<uint32_t>();
    {
        const dim3 vg(vgrid(n)), bd(kBlock);
        GT_TRY(ws.arl[0].reserve(sizeof(double) * n));
        hipLaunchKernelGGL(adj_gather<double>, vg, bd, 0, s, c, iperm, n, ws.arl[0].as<double>());
        r.cvec = ws.arl[0].as<double>();
        if (weighted) {
            GT_TRY(ws.arl[1].reserve(sizeof(double) * n));
            hipLaunchKernelGGL(adj_gather<double>, vg, bd, 0, s, a.deg, iperm, n,
                               ws.arl[1].as<double>());
            r.deg = ws.arl[1].as<double>();
        }
        if (dpers) {
            GT_TRY(ws.arl[2].reserve(sizeof(double) * n));
            hipLaunchKernelGGL(adj_gather<double>, vg, bd, 0, s, a.pers, iperm, n,
                               ws.arl[2].as<double>());
            r.pers = ws.arl[2].as<double>();
        }
        GT_HIP(hipGetLastError());
    }
    auto init = [&]() -> gt_status {
        GT_HIP(hipMemsetAsync(a.st, 0, sizeof(SweepState), s));
        AdjArgs b = r;
        b.l_next = lam[0];
        if (dpers)
            hipLaunchKernelGGL((adj_init<V, true>), dim3(vgrid(n)), dim3(kBlock), 0, s, b, n);
        else
            hipLaunchKernelGGL((adj_init<V, false>), dim3(vgrid(n)), dim3(kBlock), 0, s, b, n);
        GT_HIP(hipGetLastError());
        return GT_OK;
    };
    GT_TRY(init());
    g->prof_kind = GT_SWEEP_ADJOINT;
    const int pc = g->tune.sweep_blocks_per_cu;
    constexpr bool kF64 = std::is_same<V, double>::value;
    constexpr int bmode = kF64 ? kBlkAdj64 : kBlkAdj32;
    // Block tiles (blocks.hpp) of the transposed orientation, unless the
    // tuning turns them off: the same sweep kernel with the adjoint's row
    // update.  The loop runs on the wave engine instead when a row is too
    // long for the two-word sums (FP64), a weight is negative or not finite,
    // or a launch meets an iterate whose terms the fixed point cannot hold.
    bool on_blocks = false;
    BlkArgs bk{};
    if (g->tune.block_tile_nnz > 0) {
        Blocks *B = nullptr;
        GT_TRY(ensure_blocks(g, *R, (uint32_t)g->tune.block_tile_nnz,
                             (uint32_t)g->tune.block_long_min, &B));
        bool usable = !kF64 || B->max_row <= kBlkF64MaxRow;
        DevBuf &wbuf = kF64 ? ws.blk_w64 : ws.blk_w;
        if (usable && weighted)
            GT_TRY(blk_stage_weights(g, *B, weight, wsrc, wdt, kF64, wbuf, &usable));
        if (usable) {
            GT_TRY(blk_reserve(g, *B));
            bk = blk_layout_args(g, *B);
            bk.st = a.st;
            bk.pers = r.pers;
            bk.pval = r.pval;
            bk.damping = damping;
            bk.eps_abs = a.eps;
            bk.eps_rel = a.eps_rel;
            bk.max_iter = a.max_iter;
            bk.cvec = r.cvec;
            bk.row_off = r.row_off;
            if (weighted) {
                bk.wblk = wbuf.ptr;
                bk.wdeg = r.deg;
            }
            bk.x_cur64 = r.cvec;  // the start mass: lambda_0 = dL/dx
            blk_launch_start_ev(bk, s);
            GT_HIP(hipGetLastError());
            auto bsweep = [&](uint64_t q) {
                const int par = (int)(q & 1);
                if constexpr (kF64) {
                    bk.x_cur64 = reinterpret_cast<const double *>(lam[par]);
                    bk.c_cur64 = bk.x_cur64;
                    bk.x_next64 = reinterpret_cast<double *>(lam[par ^ 1]);
                } else {
                    bk.x_cur = reinterpret_cast<const float *>(lam[par]);
                    bk.c_cur = bk.x_cur;
                    bk.x_next = reinterpret_cast<float *>(lam[par ^ 1]);
                }
                blk_launch_sweep(bk, dpers, g->tune.block_per_cu, bmode, s);
            };
            auto bfold = [&](uint64_t) { blk_launch_fold(bk, dpers, bmode, s); };
            uint64_t launched = 0;
            GT_TRY(drive(g, a.st, params->epsilon, params->max_iter, bsweep, bfold, &launched,
                         GT_SWEEP_ADJOINT));
            if (!g->h_state->bail) {
                on_blocks = true;
            } else {
                g->layout.block_tiles = 0;
                g->layout.block_long_rows = 0;
                g->layout.block_groups = 0;
                GT_TRY(init());
            }
        }
    }
    if (weighted && !on_blocks) {
        // The wave engine's weights, in its packed order.
        GT_TRY(stage_pull_weights<V>(g, weight, rp, wsrc, wdt, ws.wcsr));
        r.w = ws.wcsr.ptr;
    }
    AdjArgs b = r;
    int sg = 0;
    // FP32 sweeps keep the most gathered entries of lambda in LDS.
    const uint32_t nhot =
        std::is_same<V, float>::value ? hot_len<float>(g->tune.sweep_hot_kb, n) : 0u;
    auto sweep = [&](uint64_t q) {
        const int par = (int)(q & 1);
        b.l_cur = lam[par];
        b.l_next = lam[par ^ 1];
        if (nhot) {
            if (weighted && dpers) sg = adj_launch_hot<true, true>(b, nhot, s);
            else if (weighted) sg = adj_launch_hot<true, false>(b, nhot, s);
            else if (dpers) sg = adj_launch_hot<false, true>(b, nhot, s);
            else sg = adj_launch_hot<false, false>(b, nhot, s);
        } else {
            if (weighted && dpers) sg = adj_launch<V, true, true>(b, pc, s);
            else if (weighted) sg = adj_launch<V, true, false>(b, pc, s);
            else if (dpers) sg = adj_launch<V, false, true>(b, pc, s);
            else sg = adj_launch<V, false, false>(b, pc, s);
        }
    };
    auto fold = [&](uint64_t) {
        if (weighted && dpers) adj_launch_fold<V, true, true>(b, sg, s);
        else if (weighted) adj_launch_fold<V, true, false>(b, sg, s);
        else if (dpers) adj_launch_fold<V, false, true>(b, sg, s);
        else adj_launch_fold<V, false, false>(b, sg, s);
    };
    uint64_t launched = 0;
    if (!on_blocks)
        GT_TRY(drive(g, a.st, params->epsilon, params->max_iter, sweep, fold, &launched,
                     GT_SWEEP_ADJOINT));
    const uint64_t sweeps = g->h_state->iter;
    const V *lr = lam[sweeps & 1];  // relabelled order
    V *lf = lam[(sweeps & 1) ^ 1];  // the caller's labels
    hipLaunchKernelGGL(adj_gather<V>, dim3(vgrid(n)), dim3(kBlock), 0, s, lr, perm, n, lf);
    GT_HIP(hipGetLastError());

    // Edge gradients: mu = M^T lambda (one more engine pass), then one pass
    // over edge ids.
    if (grad_w) {
        void *gdst = grad_w;
        const size_t gb = (gw_dtype == GT_F64 ? 8 : 4) * (size_t)g->edge_bound;
        if (grad_loc == GT_HOST) {
            GT_TRY(ws.gw.reserve(std::max<size_t>(gb, 8)));
            gdst = ws.gw.ptr;
        }
        GT_TRY(ensure_edge_list(g));
        GT_TRY(ws.mu.reserve(sizeof(double2) * n));
        GT_TRY(ws.arl[3].reserve(sizeof(double) * n));
        double2 *pair = ws.mu.as<double2>();
        double *mr = ws.arl[3].as<double>();
        const dim3 mg(wave_grid(occupancy_cap((const void *)adj_mu<V, true>, pc), r.ntiles)),
            lg(fold_grid(r.nlong)), bd(kBlock);
        if (on_blocks) {
            // mu from one more pass over the block tiles, gathering the last
            // iterate (whose mass the last fold left for this pass).
            if constexpr (kF64) {
                bk.x_cur64 = reinterpret_cast<const double *>(lr);
                bk.c_cur64 = bk.x_cur64;
            } else {
                bk.x_cur = reinterpret_cast<const float *>(lr);
                bk.c_cur = bk.x_cur;
            }
            bk.mu_out = mr;
            blk_launch_sweep(bk, dpers, g->tune.block_per_cu, bmode, s);
            blk_launch_fold(bk, dpers, bmode, s);
        } else if (weighted) {
            hipLaunchKernelGGL((adj_mu<V, true>), mg, bd, 0, s, r, lr, mr);
            hipLaunchKernelGGL((adj_mu_long<V, true>), lg, bd, 0, s, r, mr);
        } else {
            hipLaunchKernelGGL((adj_mu<V, false>), mg, bd, 0, s, r, lr, mr);
            hipLaunchKernelGGL((adj_mu_long<V, false>), lg, bd, 0, s, r, mr);
        }
        const uint32_t *ro = push->off.as<uint32_t>();
        if (weighted)
            hipLaunchKernelGGL(grad_coef<true>, dim3(vgrid(n)), bd, 0, s, a.deg, ro, mr, perm, x,
                               damping, n, pair);
        else
            hipLaunchKernelGGL(grad_coef<false>, dim3(vgrid(n)), bd, 0, s, a.deg, ro, mr, perm, x,
                               damping, n, pair);
        // Rows of the push orientation are sources for a directed view and
        // targets for a reversed one.
        const uint32_t *es = (reversed ? g->etgt : g->esrc).as<uint32_t>();
        const uint32_t *et = (reversed ? g->esrc : g->etgt).as<uint32_t>();
        const uint64_t eb = g->edge_bound;
        const dim3 gd((unsigned)std::max<uint64_t>(1, std::min<uint64_t>(16384, (eb + kBlock - 1) / kBlock)));
        if (gw_dtype == GT_F64)
            hipLaunchKernelGGL((grad_edges<V, double>), gd, bd, 0, s, es, et, pair, lf, eb,
                               static_cast<double *>(gdst));
        else
            hipLaunchKernelGGL((grad_edges<V, float>), gd, bd, 0, s, es, et, pair, lf, eb,
                               static_cast<float *>(gdst));
        GT_HIP(hipGetLastError());
        if (grad_loc == GT_HOST && gb)
            GT_HIP(hipMemcpyAsync(grad_w, gdst, gb, hipMemcpyDeviceToHost, s));
    }
    // Vertex gradients.
    {
        const dim3 gd(vgrid(n)), bd(kBlock);
        if (weighted && dpers) hipLaunchKernelGGL((grad_sums<V, true, true>), gd, bd, 0, s, a, lf, x, n);
        else if (weighted) hipLaunchKernelGGL((grad_sums<V, true, false>), gd, bd, 0, s, a, lf, x, n);
        else if (dpers) hipLaunchKernelGGL((grad_sums<V, false, true>), gd, bd, 0, s, a, lf, x, n);
        else hipLaunchKernelGGL((grad_sums<V, false, false>), gd, bd, 0, s, a, lf, x, n);
        GT_HIP(hipGetLastError());
        if (grad_pers) {
            double *gp = grad_pers;
            if (grad_loc == GT_HOST) {
                GT_TRY(ws.gp.reserve(sizeof(double) * n));
                gp = ws.gp.as<double>();
            }
            hipLaunchKernelGGL((grad_pers_k<V>), gd, bd, 0, s, lf, a.st, damping, n, gp);
            GT_HIP(hipGetLastError());
            if (grad_loc == GT_HOST)
                GT_HIP(hipMemcpyAsync(grad_pers, gp, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        }
    }
    GT_HIP(hipMemcpyAsync(g->h_state, a.st, sizeof(SweepState), hipMemcpyDeviceToHost, s));
    GT_HIP(hipStreamSynchronize(s));
    if (grad_damping) *grad_damping = g->h_state->norm / damping;
    if (iters) *iters = sweeps;
    g->last_sweeps = sweeps;
    return GT_OK;
}

}  // namespace gt

using namespace gt;

extern "C" gt_status gt_pagerank_backward(gt_graph *g, const gt_view *view, const gt_prop *weight,
                                          const gt_prop *pers, double damping,
                                          const gt_power_iteration *params, const gt_prop *x,
                                          const gt_prop *dl_dx, int32_t precision, void *grad_w,
                                          int32_t grad_w_dtype, int32_t grad_loc,
                                          double *grad_pers, double *grad_damping,
                                          uint64_t *iters) {
    clear_error();
    if (!g) return fail(GT_ERR_INVALID_ARG, "null graph");
    if (precision == GT_F64)
        return backward_impl<double>(g, view, weight, pers, damping, params, x, dl_dx, grad_w,
                                     grad_w_dtype, grad_loc, grad_pers, grad_damping, iters);
    if (precision == GT_F32)
        return backward_impl<float>(g, view, weight, pers, damping, params, x, dl_dx, grad_w,
                                    grad_w_dtype, grad_loc, grad_pers, grad_damping, iters);
    return fail(GT_ERR_INVALID_ARG, "precision must be GT_F64 or GT_F32");
}
