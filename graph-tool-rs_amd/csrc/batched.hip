// Batched personalised PageRank: k one-hot personalisation vectors solved
// together as one SpMM per sweep.  Column c is pagerank_with
// (crates/gt-algo/src/centrality.rs:261-384) with pers = e_{seeds[c]},
// started from its personalisation (a fresh map, seed_new_slots :198-205),
// stopped on its own L1 change and frozen from then on, so its sweep count is
// the one its single-vector run would report.
//
// Layout: the state is [n][64] FP32, one 256-byte row per vertex; lane c of
// a wavefront owns column c, so every gathered source row is one coalesced
// 256-byte load.  Values, weights and the SpMM accumulation are FP32; the
// teleport update, the per-column dangling mass and the per-column L1 change
// are FP64.  Each wavefront walks its own tiles of at most kMmRows rows and
// the tuned number of in-entries (ppr_sweep_w); longer rows are split into
// chunks whose [64] partial sums a second kernel per sweep (ppr_fold)
// combines in chunk order before folding the per-workgroup, per-column
// partials in index order.  k > 64 runs as consecutive blocks of 64
// columns.
#include <cmath>

#include "device_common.hpp"
#include "driver.hpp"
#include "gt_internal.hpp"
#include "props.hpp"
#include "sweep_common.hpp"

namespace gt {

constexpr int kCols = 64;
constexpr uint32_t kNoSeed = 0xffffffffu;
// Rows of an SpMM wave tile, and gathers in flight per wavefront.
constexpr uint32_t kMmRows = 16;
constexpr int kMmGroup = 16;

struct BatchState {
    SweepState h;            // delta = max over active columns, iter = sweeps
    double p_sink[kCols];
    double delta[kCols];
    uint64_t iters[kCols];
    uint64_t active;         // bit c: column c still iterating
    uint64_t was_active;     // active before the last fold
};

struct BArgs {
    const Tile *tiles;
    uint32_t ntiles;
    uint32_t nlong;
    const uint32_t *in_off;
    const uint32_t *in_idx;
    const float *w;           // pull-order weights (weighted only)
    const uint32_t *out_off;  // unity out-degree
    const double *deg;        // weighted out-degree
    const uint32_t *seeds;    // [64], kNoSeed for an unused column
    double damping;
    double eps;
    uint64_t max_iter;
    const float *x_cur;
    float *x_next;
    const float *c_cur;
    float *c_next;
    double *partials;         // [kPartials][64]: sweep partials, then fold partials
    float *long_part;         // [nslots][64]
    const uint32_t *long_rows;
    const uint32_t *long_slot;
    BatchState *st;
};

// A converged column's values are copied into the other buffers on the sweep
// after it stopped (col_copy) and left alone afterwards: both buffers then
// hold its last iterate.
template <bool WEIGHTED>
__device__ __forceinline__ void finish_row(const BArgs &a, uint32_t r, float acc, int lane,
                                           bool col_on, bool col_copy, uint32_t seed,
                                           double psink, double &dsum, double &ssum) {
    const size_t o = (size_t)r * kCols + lane;
    if (!col_on) {
        if (col_copy) {
            a.x_next[o] = a.x_cur[o];
            a.c_next[o] = a.c_cur[o];
        }
        return;
    }
    const float xo = a.x_cur[o];
    const double pv = (r == seed) ? 1.0 : 0.0;
    const double nv = (1.0 - a.damping) * pv + a.damping * (psink * pv + (double)acc);
    const float nf = (float)nv;
    dsum += fabs((double)nf - (double)xo);
    a.x_next[o] = nf;
    const double dg = WEIGHTED ? a.deg[r] : (double)(a.out_off[r + 1] - a.out_off[r]);
    if (dg != 0.0) {
        a.c_next[o] = (float)((double)nf / dg);
    } else {
        a.c_next[o] = 0.f;
        ssum += (double)nf;
    }
}

__device__ __forceinline__ double readlane_f64(double x, uint32_t l) {
    const uint64_t b = __builtin_bit_cast(uint64_t, x);
    const uint32_t lo = __builtin_amdgcn_readlane((int)(uint32_t)b, l);
    const uint32_t hi = __builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), l);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// finish_row with the row's iterate (xo, this lane's column) and out-degree
// (dg) already loaded.
template <bool WEIGHTED>
__device__ __forceinline__ void finish_row_pre(const BArgs &a, uint32_t r, float acc, float xo,
                                               double dg, int lane, bool col_on, bool col_copy,
                                               uint32_t seed, double psink, double &dsum,
                                               double &ssum) {
    const size_t o = (size_t)r * kCols + lane;
    if (!col_on) {
        if (col_copy) {
            a.x_next[o] = xo;
            a.c_next[o] = a.c_cur[o];
        }
        return;
    }
    const double pv = (r == seed) ? 1.0 : 0.0;
    const double nv = (1.0 - a.damping) * pv + a.damping * (psink * pv + (double)acc);
    const float nf = (float)nv;
    dsum += fabs((double)nf - (double)xo);
    a.x_next[o] = nf;
    if (dg != 0.0) {
        a.c_next[o] = (float)((double)nf / dg);
    } else {
        a.c_next[o] = 0.f;
        ssum += (double)nf;
    }
}

// Adds w(j) * c[src(j)][lane] over the entries u0 + u, u < G, of a chunk
// whose sources (and weights) lane u holds in `my` (`myw`; lanes at or past
// `cnt` repeat the chunk's last entry): the sources are read out to scalars,
// so each gather is one 256-byte row at a scalar address, and all G are in
// flight before the first is used.  A converged column's lane loads nothing
// (its sums are never used; a masked lane costs the gather nothing, while a
// lane that loads, even a byte another lane reads, does).  Entries at or
// past `cnt` are not added.  `add(u, x)` takes the values in entry order.
template <bool WEIGHTED, int G, typename Add>
__device__ __forceinline__ void gather_group(const float *__restrict__ c, uint32_t my, float myw,
                                             uint32_t u0, uint32_t cnt, int lane, bool col_on,
                                             Add add) {
    cnt = __builtin_amdgcn_readfirstlane(cnt);
    u0 = __builtin_amdgcn_readfirstlane(u0);
    float v[G];
#pragma unroll
    for (int u = 0; u < G; ++u) v[u] = 0.f;
    if (col_on) {
#pragma unroll
        for (int u = 0; u < G; ++u) {
            const uint32_t s = __builtin_amdgcn_readlane(my, u0 + u);
            const float *row = c + (size_t)s * kCols;  // scalar: one base per gather
            v[u] = row[lane];
        }
    }
#pragma unroll
    for (int u = 0; u < G; ++u) {
        if (u0 + u >= cnt) break;
        const float w = WEIGHTED ? __builtin_bit_cast(float, __builtin_amdgcn_readlane(
                                                                 __builtin_bit_cast(int, myw),
                                                                 u0 + u))
                                 : 1.f;
        add(u0 + u, WEIGHTED ? w * v[u] : v[u]);
    }
}

// The SpMM sweep over wave tiles of at most kMmRows rows (or one chunk of a
// split row): each wavefront walks its own tiles.  A tile's sources are
// loaded 64 at a time (one coalesced load), its rows' iterates and
// out-degrees before the gathers; the gathers go out G at a time; each row's
// sum is kept in a register in entry order and parked in the wavefront's
// LDS rows when the row ends; the rows are finished after the entries.
template <bool WEIGHTED, int G>
__global__ __launch_bounds__(kBlock) void ppr_sweep_w(const BArgs a) {
    if (a.st->h.done) return;
    __shared__ float s_rows[kBlock / kWave][kMmRows][kCols];
    __shared__ double s_red[2][kBlock / kWave][kCols];
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wid = tid / kWave;
    float(*rows)[kCols] = s_rows[wid];
    const uint64_t active = a.st->active;
    const bool col_on = (active >> lane) & 1ull;
    const bool col_copy = !col_on && ((a.st->was_active >> lane) & 1ull);
    const uint32_t seed = a.seeds[lane];
    const double psink = a.st->p_sink[lane];
    double dsum = 0.0, ssum = 0.0;
    const uint32_t nwaves = gridDim.x * (kBlock / kWave);

    for (uint32_t ti = blockIdx.x * (kBlock / kWave) + wid; ti < a.ntiles; ti += nwaves) {
        const Tile t = wave_tile(a.tiles, ti);
        const uint32_t nb = t.nnz_begin, ne = t.nnz_end;
        if (t.row_end == kSplit) {
            // A chunk of a split row: its sum goes to the row's slot.
            float acc = 0.f;
            for (uint32_t j0 = nb; j0 < ne; j0 += kWave) {
                const uint32_t cnt = ne - j0 < (uint32_t)kWave ? ne - j0 : (uint32_t)kWave;
                const uint32_t j = j0 + ((uint32_t)lane < cnt ? (uint32_t)lane : cnt - 1);
                const uint32_t my = a.in_idx[j];
                const float myw = WEIGHTED ? a.w[j] : 1.f;
                for (uint32_t u0 = 0; u0 < cnt; u0 += G)
                    gather_group<WEIGHTED, G>(a.c_cur, my, myw, u0, cnt, lane, col_on,
                                              [&](uint32_t, float x) { acc += x; });
            }
            a.long_part[(size_t)t.row_begin * kCols + lane] = acc;
            continue;
        }
        const uint32_t rb = t.row_begin, nr = t.row_end - rb;  // 1 <= nr <= kMmRows
        // Lane q <= nr: the offset of row rb + q; lane q < nr: its out-degree.
        const uint32_t ql = (uint32_t)lane < nr ? (uint32_t)lane : nr;
        const uint32_t roff = a.in_off[rb + ql];
        const uint32_t qd = (uint32_t)lane < nr ? (uint32_t)lane : nr - 1;
        const double dgl = WEIGHTED ? a.deg[rb + qd]
                                    : (double)(a.out_off[rb + qd + 1] - a.out_off[rb + qd]);
        float xo[kMmRows];
#pragma unroll
        for (int q = 0; q < (int)kMmRows; ++q)
            xo[q] = a.x_cur[(size_t)(rb + ((uint32_t)q < nr ? (uint32_t)q : nr - 1)) * kCols + lane];
        uint32_t written = 0, cur = 0;
        float acc = 0.f;
        for (uint32_t j0 = nb; j0 < ne; j0 += kWave) {
            const uint32_t cnt = ne - j0 < (uint32_t)kWave ? ne - j0 : (uint32_t)kWave;
            const uint32_t j = j0 + ((uint32_t)lane < cnt ? (uint32_t)lane : cnt - 1);
            const uint32_t my = a.in_idx[j];
            const float myw = WEIGHTED ? a.w[j] : 1.f;
            // Row of entry j0 + lane: the rows q >= 1 starting at or before it.
            uint32_t rq = 0;
            for (uint32_t q = 1; q < nr; ++q)
                rq += (uint32_t)(__builtin_amdgcn_readlane(roff, q) <= j0 + (uint32_t)lane);
            for (uint32_t u0 = 0; u0 < cnt; u0 += G)
                gather_group<WEIGHTED, G>(a.c_cur, my, myw, u0, cnt, lane, col_on,
                                          [&](uint32_t u, float x) {
                                              const uint32_t q = __builtin_amdgcn_readlane(rq, u);
                                              if (q != cur) {
                                                  rows[cur][lane] = acc;
                                                  written |= 1u << cur;
                                                  acc = 0.f;
                                                  cur = q;
                                              }
                                              acc += x;
                                          });
        }
        if (ne > nb) {
            rows[cur][lane] = acc;
            written |= 1u << cur;
        }
        wave_sync();
#pragma unroll
        for (int q = 0; q < (int)kMmRows; ++q) {
            if ((uint32_t)q >= nr) break;
            const float sum = (written >> q) & 1u ? rows[q][lane] : 0.f;
            const double dg = readlane_f64(dgl, q);
            finish_row_pre<WEIGHTED>(a, rb + q, sum, xo[q], dg, lane, col_on, col_copy, seed,
                                     psink, dsum, ssum);
        }
        wave_sync();  // the LDS rows are reused by this wavefront's next tile
    }

    s_red[0][wid][lane] = dsum;
    s_red[1][wid][lane] = ssum;
    __syncthreads();
    if (wid == 0) {
        const double d = ((s_red[0][0][lane] + s_red[0][1][lane]) + s_red[0][2][lane]) +
                         s_red[0][3][lane];
        const double s = ((s_red[1][0][lane] + s_red[1][1][lane]) + s_red[1][2][lane]) +
                         s_red[1][3][lane];
        a.partials[((size_t)blockIdx.x) * kCols + lane] = d;
        a.partials[((size_t)gridDim.x + blockIdx.x) * kCols + lane] = s;
    }
}

// Completes the preceding ppr_sweep_w launch (`sgrid` workgroups): one
// wavefront per split row combines its chunk partials in chunk order and
// finishes the row, and workgroup b also sums its contiguous slice of the
// sweep's per-column partials; the last workgroup to arrive folds the
// workgroup sums in index order and updates the per-column state.  (Two
// levels, so no workgroup walks all of the sweep's partials alone.)
// Sums of rows b, b + stride, ... < e of two [rows][64] partial arrays (d
// from pd, s from ps), column `lane`, in row order, in rounds of 16 rows in
// flight (a round's loads past e read row b again and are not added).
__device__ __forceinline__ void sum_rows2(const double *pd, const double *ps, uint32_t b,
                                          uint32_t e, uint32_t stride, int lane, double &d,
                                          double &s) {
    for (; b < e; b += 16 * stride) {
        double vd[16], vs[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const uint32_t q = b + (uint32_t)i * stride;
            const size_t o = (size_t)(q < e ? q : b) * kCols + lane;
            vd[i] = pd[o];
            vs[i] = ps[o];
        }
#pragma unroll
        for (int i = 0; i < 16; ++i)
            if (b + (uint32_t)i * stride < e) {
                d += vd[i];
                s += vs[i];
            }
    }
}

template <bool WEIGHTED>
__global__ __launch_bounds__(kBlock) void ppr_fold(const BArgs a, uint32_t sgrid) {
    if (a.st->h.done) return;
    __shared__ double s_red[2][kBlock / kWave][kCols];
    __shared__ uint32_t flag;
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wid = tid / kWave;
    const uint64_t active = a.st->active;
    const bool col_on = (active >> lane) & 1ull;
    const bool col_copy = !col_on && ((a.st->was_active >> lane) & 1ull);
    const uint32_t seed = a.seeds[lane];
    const double psink = a.st->p_sink[lane];
    double d = 0.0, s = 0.0;
    const uint32_t b0 = (uint32_t)((uint64_t)sgrid * blockIdx.x / gridDim.x);
    const uint32_t b1 = (uint32_t)((uint64_t)sgrid * (blockIdx.x + 1) / gridDim.x);
    sum_rows2(a.partials, a.partials + (size_t)sgrid * kCols, b0 + wid, b1, kBlock / kWave, lane,
              d, s);
    const uint32_t waves = gridDim.x * (kBlock / kWave);
    for (uint32_t k = blockIdx.x * (kBlock / kWave) + wid; k < a.nlong; k += waves) {
        // Chunk order, in rounds of 16 chunk rows in flight (a round's loads
        // past the row read its last chunk again and are not added): a hub's
        // hundreds of chunks set the fold's critical path.
        float acc = 0.f;
        const uint32_t e = a.long_slot[k + 1];
        for (uint32_t q = a.long_slot[k]; q < e; q += 16) {
            float v[16];
#pragma unroll
            for (int i = 0; i < 16; ++i)
                v[i] = a.long_part[(size_t)(q + i < e ? q + i : e - 1) * kCols + lane];
#pragma unroll
            for (int i = 0; i < 16; ++i)
                if (q + i < e) acc += v[i];
        }
        finish_row<WEIGHTED>(a, a.long_rows[k], acc, lane, col_on, col_copy, seed, psink, d, s);
    }
    s_red[0][wid][lane] = d;
    s_red[1][wid][lane] = s;
    __syncthreads();
    double *fp = a.partials + 2 * (size_t)kMaxSweepGrid * kCols;
    if (wid == 0) {
        d = ((s_red[0][0][lane] + s_red[0][1][lane]) + s_red[0][2][lane]) + s_red[0][3][lane];
        s = ((s_red[1][0][lane] + s_red[1][1][lane]) + s_red[1][2][lane]) + s_red[1][3][lane];
        fp[((size_t)blockIdx.x) * kCols + lane] = d;
        fp[((size_t)gridDim.x + blockIdx.x) * kCols + lane] = s;
    }
    if (!last_to_arrive(&a.st->h.counter, &flag)) return;

    d = 0.0;
    s = 0.0;
    sum_rows2(fp, fp + (size_t)gridDim.x * kCols, wid, gridDim.x, kBlock / kWave, lane, d, s);
    s_red[0][wid][lane] = d;
    s_red[1][wid][lane] = s;
    __syncthreads();
    if (wid == 0) {
        const double dt = ((s_red[0][0][lane] + s_red[0][1][lane]) + s_red[0][2][lane]) +
                          s_red[0][3][lane];
        const double stt = ((s_red[1][0][lane] + s_red[1][1][lane]) + s_red[1][2][lane]) +
                           s_red[1][3][lane];
        BatchState *st = a.st;
        bool keep = false;
        if (col_on) {
            const uint64_t it = st->iters[lane] + 1;
            st->iters[lane] = it;
            st->delta[lane] = dt;
            st->p_sink[lane] = stt;
            keep = (dt >= a.eps) && !(a.max_iter > 0 && it >= a.max_iter);
        }
        const uint64_t mask = __ballot(keep);
        double dmax = col_on ? dt : 0.0;
#pragma unroll
        for (int o = kWave / 2; o > 0; o >>= 1) dmax = fmax(dmax, __shfl_xor(dmax, o, kWave));
        if (lane == 0) {
            st->was_active = active;
            st->active = mask;
            st->h.delta = dmax;
            st->h.iter += 1;
            st->h.launches += 1;
            st->h.done = mask == 0 ? 1u : 0u;
            st->h.counter = 0;
        }
    }
}

// x0 = e_seed per column, c0 = x0 / deg; p_sink for the first sweep is 1
// for a column whose seed is a sink.
template <bool WEIGHTED>
__global__ __launch_bounds__(kBlock) void ppr_init(const BArgs a, uint32_t n, uint64_t active) {
    const int lane = threadIdx.x & (kWave - 1);
    const uint32_t seed = a.seeds[lane];
    const uint32_t waves = gridDim.x * (kBlock / kWave);
    for (uint32_t r = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave; r < n; r += waves) {
        const size_t o = (size_t)r * kCols + lane;
        const float x = (r == seed) ? 1.f : 0.f;
        a.x_next[o] = x;
        const double dg = WEIGHTED ? a.deg[r] : (double)(a.out_off[r + 1] - a.out_off[r]);
        a.c_next[o] = dg != 0.0 ? (float)(1.0 / dg) * x : 0.f;
    }
    if (blockIdx.x == 0 && threadIdx.x < kWave) {
        BatchState *st = a.st;
        double ps = 0.0;
        if (seed != kNoSeed) {
            const double dg =
                WEIGHTED ? a.deg[seed] : (double)(a.out_off[seed + 1] - a.out_off[seed]);
            ps = dg == 0.0 ? 1.0 : 0.0;
        }
        st->p_sink[lane] = ps;
        st->delta[lane] = 0.0;
        st->iters[lane] = 0;
        if (lane == 0) {
            st->active = active;
            st->was_active = active;
            st->h = SweepState{};
        }
    }
}

// Copy column block [n][64] (sweep labels) into the caller's [n][ldx] layout
// at column c0: caller row o is sweep row perm[o] (perm null: identity).
__global__ __launch_bounds__(kBlock) void ppr_emit(const float *x, const uint32_t *perm,
                                                   uint32_t n, uint32_t kb, uint64_t ldx,
                                                   uint32_t c0, float *out) {
    const int lane = threadIdx.x & (kWave - 1);
    const uint32_t waves = gridDim.x * (kBlock / kWave);
    for (uint32_t r = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave; r < n; r += waves) {
        const size_t sr = perm ? perm[r] : r;
        if ((uint32_t)lane < kb) out[(size_t)r * ldx + c0 + lane] = x[sr * kCols + lane];
    }
}

// seeds[c] <- perm[seeds[c]] for the used columns: seeds into sweep labels.
__global__ void ppr_map_seeds(uint32_t *seeds, const uint32_t *perm) {
    const int c = threadIdx.x;
    if (c < kCols && seeds[c] != kNoSeed) seeds[c] = perm[seeds[c]];
}

// out[k] = in[iperm[k]]: a vertex map into sweep labels.
__global__ __launch_bounds__(kBlock) void ppr_permute_f64(const double *__restrict__ in,
                                                          const uint32_t *__restrict__ iperm,
                                                          uint32_t n, double *__restrict__ out) {
    for (uint32_t k = blockIdx.x * kBlock + threadIdx.x; k < n; k += gridDim.x * kBlock)
        out[k] = in[iperm[k]];
}

static unsigned row_waves_grid(uint32_t n) {
    uint64_t b = ((uint64_t)n + 3) / 4;
    if (b > 8192) b = 8192;
    return b ? (unsigned)b : 1u;
}

gt_status pagerank_batched_impl(gt_graph *g, const gt_view *view, const gt_prop *weight,
                                const uint32_t *seeds, uint32_t k, double damping,
                                const gt_power_iteration *params, float *x, uint64_t ldx,
                                int32_t x_loc, uint64_t *iters) {
    clear_error();
    if (!g) return fail(GT_ERR_INVALID_ARG, "null graph");
    if (!params) return fail(GT_ERR_INVALID_ARG, "params is null");
    if (params->plan.len != g->n)
        return fail(GT_ERR_PLAN_COVER, "sweep: the plan must cover the output exactly");
    if (k == 0) return GT_OK;
    if (!seeds || !x || !iters) return fail(GT_ERR_INVALID_ARG, "null argument");
    if (ldx < k) return fail(GT_ERR_INVALID_ARG, "ldx is smaller than k");
    if (x_loc != GT_HOST && x_loc != GT_DEVICE) return fail(GT_ERR_INVALID_ARG, "bad location");
    if (params->max_iter == 0 && !(params->epsilon > 0.0))
        return fail(GT_ERR_INVALID_ARG,
                    "epsilon must be positive when max_iter is 0: the loop would never end");
    const uint32_t n = g->n;
    for (uint32_t c = 0; c < k; ++c)
        if (seeds[c] >= n) return fail(GT_ERR_NO_SUCH_VERTEX, "seed vertex out of range");
    Orient *pull = nullptr, *push = nullptr;
    bool reversed = false;
    const uint8_t *active = nullptr;
    uint32_t nactive = 0;
    GT_TRY(resolve_view(g, view, &pull, &push, &reversed, &active, &nactive));
    if (active) return fail(GT_ERR_UNSUPPORTED, "batched personalisation on a filtered view");
    hipStream_t s = g->stream;
    Workspace &ws = g->ws;
    BArgs a{};
    // The sweeps run in the locality-friendly labelling: hub rows (256 bytes
    // each) sit together and stay cached across the entries that read them.
    Relabeled *R = nullptr;
    GT_TRY(ensure_relabeled(g, reversed ? 1 : 0, &R));
    Orient *push_caller = push;
    pull = &R->pull;
    const uint32_t *perm = R->perm.as<uint32_t>(), *iperm = R->iperm.as<uint32_t>();
    bool weighted = false;
    if (weight && weight->kind != GT_MAP_UNITY) {
        weighted = true;
        const void *src = nullptr;
        int dt = GT_F64;
        GT_TRY(stage_edge_source(g, weight, &src, &dt));
        GT_TRY(stage_degree(g, weight, *push_caller, src, dt, ws.deg));
        GT_TRY(stage_csc_weights<float>(g, weight, *pull, src, dt, ws.wcsc));
        GT_TRY(ws.deg2.reserve(sizeof(double) * std::max<uint32_t>(1, n)));
        hipLaunchKernelGGL(ppr_permute_f64, dim3(row_waves_grid(n)), dim3(kBlock), 0, s,
                           ws.deg.as<double>(), iperm, n, ws.deg2.as<double>());
        GT_HIP(hipGetLastError());
        a.deg = ws.deg2.as<double>();
        a.w = ws.wcsc.as<float>();
    }
    GT_TRY(ensure_schedule_mm(*pull, n, (uint32_t)g->tune.mm_tile_nnz, kMmRows));
    const Schedule &S = pull->mm;
    const int cap_w =
        occupancy_cap((const void *)ppr_sweep_w<true, kMmGroup>, g->tune.mm_blocks_per_cu);
    const int cap_u =
        occupancy_cap((const void *)ppr_sweep_w<false, kMmGroup>, g->tune.mm_blocks_per_cu);
    const int grid_w = wave_grid(cap_w, S.ntiles), grid_u = wave_grid(cap_u, S.ntiles);
    // At least 64 fold workgroups: each sums at most 32 of the sweep's rows of
    // partials before the last one folds the workgroup sums.
    const unsigned fgrid =
        (unsigned)fit_grid(kMaxSweepGrid, std::max<uint32_t>(64, (S.nlong + 3) / 4));
    GT_TRY(prof_prepare(g));
    const size_t vb = sizeof(float) * kCols * (size_t)n;
    GT_TRY(ws.bx[0].reserve(vb));
    GT_TRY(ws.bx[1].reserve(vb));
    GT_TRY(ws.bc[0].reserve(vb));
    GT_TRY(ws.bc[1].reserve(vb));
    GT_TRY(ws.bpart.reserve(sizeof(double) * kCols * kPartials));
    GT_TRY(ws.blong.reserve(sizeof(float) * kCols * std::max<uint32_t>(1, S.nslots)));
    GT_TRY(ws.bstate.reserve(sizeof(BatchState) + sizeof(uint32_t) * kCols));
    BatchState *st = ws.bstate.as<BatchState>();
    uint32_t *dseeds = reinterpret_cast<uint32_t *>(st + 1);
    float *xb[2] = {ws.bx[0].as<float>(), ws.bx[1].as<float>()};
    float *cb[2] = {ws.bc[0].as<float>(), ws.bc[1].as<float>()};
    DevBuf hostx;
    float *xout = x;
    if (x_loc == GT_HOST) {
        GT_TRY(hostx.reserve(sizeof(float) * ldx * (size_t)n));
        xout = hostx.as<float>();
    }
    a.tiles = S.tiles.as<Tile>();
    a.ntiles = S.ntiles;
    a.nlong = S.nlong;
    a.in_off = pull->off.as<uint32_t>();
    a.in_idx = pull->idx.as<uint32_t>();
    a.out_off = R->push_off.as<uint32_t>();
    a.seeds = dseeds;
    a.damping = damping;
    a.eps = params->epsilon;
    a.max_iter = params->max_iter;
    a.partials = ws.bpart.as<double>();
    a.long_part = ws.blong.as<float>();
    a.long_rows = S.long_rows.as<uint32_t>();
    a.long_slot = S.long_slot.as<uint32_t>();
    a.st = st;
    g->prof_kind = GT_SWEEP_BATCHED;
    uint64_t total = 0;
    for (uint32_t c0 = 0; c0 < k; c0 += kCols) {
        const uint32_t kb = std::min<uint32_t>(kCols, k - c0);
        uint32_t hs[kCols];
        for (uint32_t c = 0; c < (uint32_t)kCols; ++c) hs[c] = c < kb ? seeds[c0 + c] : kNoSeed;
        GT_HIP(hipMemcpyAsync(dseeds, hs, sizeof(hs), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(ppr_map_seeds, dim3(1), dim3(kCols), 0, s, dseeds, perm);
        const uint64_t mask = kb == 64 ? ~0ull : ((1ull << kb) - 1);
        BArgs ai = a;
        ai.x_next = xb[0];
        ai.c_next = cb[0];
        if (weighted)
            hipLaunchKernelGGL(ppr_init<true>, dim3(row_waves_grid(n)), dim3(kBlock), 0, s, ai, n,
                               mask);
        else
            hipLaunchKernelGGL(ppr_init<false>, dim3(row_waves_grid(n)), dim3(kBlock), 0, s, ai,
                               n, mask);
        GT_HIP(hipGetLastError());
        BArgs b = a;
        auto sweep = [&](uint64_t q) {
            const int par = (int)(q & 1);
            b.x_cur = xb[par];
            b.x_next = xb[par ^ 1];
            b.c_cur = cb[par];
            b.c_next = cb[par ^ 1];
            if (weighted)
                hipLaunchKernelGGL((ppr_sweep_w<true, kMmGroup>), dim3(grid_w), dim3(kBlock), 0,
                                   s, b);
            else
                hipLaunchKernelGGL((ppr_sweep_w<false, kMmGroup>), dim3(grid_u), dim3(kBlock), 0,
                                   s, b);
        };
        auto fold = [&](uint64_t) {
            if (weighted)
                hipLaunchKernelGGL(ppr_fold<true>, dim3(fgrid), dim3(kBlock), 0, s, b,
                                   (uint32_t)grid_w);
            else
                hipLaunchKernelGGL(ppr_fold<false>, dim3(fgrid), dim3(kBlock), 0, s, b,
                                   (uint32_t)grid_u);
        };
        uint64_t launched = 0;
        GT_TRY(drive(g, reinterpret_cast<SweepState *>(st), params->epsilon, params->max_iter,
                     sweep, fold, &launched, GT_SWEEP_BATCHED));
        const uint64_t sweeps = g->h_state->iter;
        total += sweeps;
        uint64_t hit[kCols];
        GT_HIP(hipMemcpyAsync(hit, st->iters, sizeof(hit), hipMemcpyDeviceToHost, s));
        hipLaunchKernelGGL(ppr_emit, dim3(row_waves_grid(n)), dim3(kBlock), 0, s, xb[sweeps & 1],
                           perm, n, kb, ldx, c0, xout);
        GT_HIP(hipGetLastError());
        GT_HIP(hipStreamSynchronize(s));
        for (uint32_t c = 0; c < kb; ++c) iters[c0 + c] = hit[c];
    }
    if (x_loc == GT_HOST)
        GT_HIP(hipMemcpyAsync(x, xout, sizeof(float) * ldx * (size_t)n, hipMemcpyDeviceToHost, s));
    GT_HIP(hipStreamSynchronize(s));
    g->last_sweeps = total;
    return GT_OK;
}

}  // namespace gt

extern "C" gt_status gt_pagerank_batched(gt_graph *g, const gt_view *view, const gt_prop *weight,
                                         const uint32_t *seeds, uint32_t k, double damping,
                                         const gt_power_iteration *params, float *x,
                                         uint64_t ldx, int32_t x_loc, uint64_t *iters) {
    return gt::pagerank_batched_impl(g, view, weight, seeds, k, damping, params, x, ldx, x_loc,
                                     iters);
}
