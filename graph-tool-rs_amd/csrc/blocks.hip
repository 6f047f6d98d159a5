// Block-tile sweeps of the FP32 PageRank phase: the layout builder, the sweep
// and fold kernels.  What a sweep computes is pagerank_with's update
// (crates/gt-algo/src/centrality.rs:345-363) with the FP32 form's types: the
// FP32 contributions c = x / deg of the last sweep are gathered, each row's
// sum is formed in 64-bit fixed point (blocks.hpp), and the teleport update,
// dangling mass and L1 change are FP64, as in the wave engine's FP32 sweep.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <mutex>
#include <thread>

#include "blocks.hpp"
#include "device_common.hpp"

namespace gt {

// Blocks of kBlkQuad groups per wavefront batch: 4 word loads of 16 bytes a
// lane, then 16 gathers in flight.
constexpr int kBlkBatch = 4;
// Wavefronts per sweep workgroup: 8 for the one-word sweeps (two or three
// workgroups share a CU), 16 for the two-word sweeps, whose 96 KB of
// accumulators leave room for one workgroup per CU.
constexpr int kBlkWaves = 8;
constexpr int kBlkThreads = kBlkWaves * kWave;
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
constexpr uint32_t kBlkLds = kBlkAcc + kBlkLongMax * kWave;
// The row accumulator of a long row in the tile's range: no sum reaches
// 2^62 in magnitude, so no sum equals it.
constexpr unsigned long long kBlkMarked = 1ull << 63;
static_assert(kBlkLongMax * 16 <= kBlkThreads, "16 threads sum each long row");
static_assert(2 * kBlkLds * 8 <= 160 * 1024, "two accumulator words fit one CU's LDS");
static_assert(kBlkNull >= kBlkRows && kBlkNull < kBlkAcc, "padding adds into a dummy");
static_assert(kBlkLds <= (1u << kBlkCodeBits), "accumulator indices fit their bits");
// Row codes of the sort key: a short row's r - r0, the j-th long row's
// kBlkAcc + j (both below 2^kBlkCodeBits).
constexpr uint32_t kBlkKeyLong0 = kBlkAcc;
constexpr uint32_t kBlkCodeMask = (1u << kBlkCodeBits) - 1u;
typedef const __attribute__((address_space(4))) uint32_t ConstU32;
typedef const __attribute__((address_space(4))) BlkTile ConstBlkTile;

// The fixed-point exponent for an iterate of L1 mass `mass` (<= kBlkMassMax):
// 2^K * mass < 2^62.  A source's terms sum to at most its |x| (each is
// x / deg, and it has deg of them), so every partial sum of a sweep's terms,
// scaled by 2^K, stays below 2^62.
__device__ __forceinline__ int blk_scale(double mass) {
    return mass > 0.0 ? 61 - ilogb(mass) : 0;
}

// An integer-valued double of magnitude below 2^62 as a two's complement
// 64-bit integer: its high and low 32-bit halves by floor and fma, each
// converted exactly.
__device__ __forceinline__ unsigned long long blk_i64(double h) {
    const double hh = floor(ldexp(h, -32));
    const double ll = fma(hh, -0x1p32, h);
    const uint32_t hi = static_cast<uint32_t>(static_cast<int32_t>(hh));
    const uint32_t lo = static_cast<uint32_t>(ll);
    return (static_cast<unsigned long long>(hi) << 32) | lo;
}

// round(t * 2^K) (ties to even) as a 64-bit integer, |t| * 2^K < 2^62
// (blk_scale).  An FP32 term -- or a product of two FP32 values, 48
// significant bits -- is exact in FP64, so only this rounding to the grid
// touches it.
__device__ __forceinline__ unsigned long long blk_fixed_d(double t, int K) {
    return blk_i64(rint(ldexp(t, K)));
}
__device__ __forceinline__ unsigned long long blk_fixed(float c, int K) {
    return blk_fixed_d(static_cast<double>(c), K);
}

// An FP64 term in two words: hi = round(c * 2^K) and lo = round(r *
// 2^kBlkLoBits), r = c * 2^K - hi (|r| <= 1/2, found exactly), so c = (hi +
// lo * 2^-kBlkLoBits) * 2^-K to 2^-(K + kBlkLoBits + 1).  The low words of
// a row (at most 2^23 terms) sum below 2^62.
constexpr int kBlkLoBits = 38;
__device__ __forceinline__ void blk_fixed2(double c, int K, unsigned long long &hi,
                                           unsigned long long &lo) {
    const double f = ldexp(c, K);
    const double h = rint(f);
    hi = blk_i64(h);
    lo = blk_i64(rint(ldexp(f - h, kBlkLoBits)));
}

// A row's value from its one or two summed words.
__device__ __forceinline__ double blk_value(unsigned long long hi, int K) {
    return ldexp(static_cast<double>(static_cast<long long>(hi)), -K);
}
__device__ __forceinline__ double blk_value2(unsigned long long hi, unsigned long long lo, int K) {
    return ldexp(static_cast<double>(static_cast<long long>(hi)) +
                     ldexp(static_cast<double>(static_cast<long long>(lo)), -kBlkLoBits),
                 -K);
}

template <int MODE> struct BlkTypes {
    typedef float X;   // the iterate a row update reads
    typedef float C;   // the gathered contribution
    typedef float Wt;  // an edge weight (weighted sweeps)
    static constexpr bool kTwo = false;
};
// The refinement gathers the FP32 contributions and weighs them with the
// map's FP64 weights, as the wave engine's refinement does.
template <> struct BlkTypes<kBlkRefine> : BlkTypes<kBlkF32> {
    typedef double Wt;
};
template <> struct BlkTypes<kBlkF64> {
    typedef double X;
    typedef double C;
    typedef double Wt;
    static constexpr bool kTwo = true;
};
template <> struct BlkTypes<kBlkEv> : BlkTypes<kBlkF64> {};
template <> struct BlkTypes<kBlkAdj32> : BlkTypes<kBlkF32> {};
template <> struct BlkTypes<kBlkAdj64> : BlkTypes<kBlkF64> {};

// The sweep is compiled for four wavefronts per SIMD (at most 128 VGPRs):
// two workgroups of 8 wavefronts per CU, or one of 16.
template <int MODE, bool WGT = false> struct BlkShape {
    static constexpr int kWaves = BlkTypes<MODE>::kTwo ? 16 : kBlkWaves;
    static constexpr int kThreads = kWaves * kWave;
    // Blocks per wavefront batch.  A weighted sweep holds each lane's
    // weights beside its terms: half the batch keeps it within 128 VGPRs
    // without spills (weighted FP32 sweep at R-MAT 22: 172.5 us, against
    // 178.9 with the full batch and 36 bytes of spills, 179.5 at three
    // wavefronts per SIMD).  Its thread count must be the unweighted
    // sweep's: the rows each thread finishes, and so the order of the
    // partial sums, follow it, and unit weights must give the unweighted
    // result bit for bit.
    static constexpr int kBatch = WGT ? 2 : kBlkBatch;
    static constexpr int kRows = (kBlkRows + kThreads - 1) / kThreads;  // finisher rows a thread
};


// Tile t's descriptor, as scalar loads (the table is read-only in a sweep).
__device__ __forceinline__ BlkTile blk_tile(const BlkTile *tiles, uint32_t t) {
    ConstBlkTile *c = (ConstBlkTile *)tiles + t;
    BlkTile r;
    r.r0 = c->r0;
    r.r1 = c->r1;
    r.g0 = c->g0;
    r.g1 = c->g1;
    r.lfirst = c->lfirst;
    r.nlong = c->nlong;
    return r;
}

// Row i's update from its sum `acc` and its inputs (iterate xc, out-degree
// dg, personalisation pv, damping d).  The FP32 sweep follows PrFin's FP32 finisher
// (the teleport update in FP64, the iterate and its contribution rounded to
// FP32), the FP64 sweep PrFin's FP64 one; both add the L1 mass of the new
// iterate.  The refinement follows RefineFin (the FP64 update is the
// output; delta against the FP32 iterate; no contribution, no dangling mass).
template <int MODE>
__device__ __forceinline__ void blk_finish(const BlkArgs &b, uint32_t i, double acc, double xc,
                                           double dg, double pv, double p_sink, double d,
                                           double &dsum, double &ssum, double &msum) {
    const double nv = (1.0 - d) * pv + d * (p_sink * pv + acc);
    if (MODE == kBlkRefine) {
        dsum += fabs(nv - xc);
        b.x_next64[i] = nv;
        return;
    }
    if (MODE == kBlkF64) {
        dsum += fabs(nv - xc);
        msum += fabs(nv);
        b.x_next64[i] = nv;
        if (dg != 0.0) {
            b.c_next64[i] = nv / dg;
        } else {
            b.c_next64[i] = 0.0;
            ssum += nv;
        }
        return;
    }
    const float nvv = static_cast<float>(nv);
    dsum += fabs(static_cast<double>(nvv) - xc);
    msum += fabs(static_cast<double>(nvv));
    b.x_next[i] = nvv;
    if (dg != 0.0) {
        b.c_next[i] = static_cast<float>(static_cast<double>(nvv) / dg);
    } else {
        b.c_next[i] = 0.0f;
        ssum += static_cast<double>(nvv);
    }
}

// The eigenvector's row update, as EvFin (eigenvector.hip): row i's new
// unscaled iterate sv = acc / sc_cur (acc sums y_{L-1}, times the weight
// when weighted, over the row), its square into the norm, the delta of
// sweep L-1 from the row's two previous iterates yc = y_{L-1}(i), yp =
// y_{L-2}(i), and |sv| * the times i is gathered (dg) into the mass the next
// launch's terms are bounded by (times the largest weight, in the fold).
__device__ __forceinline__ void blk_finish_ev(const BlkArgs &b, uint32_t i, double acc, double yc,
                                              double yp, uint32_t dg, double sc_cur,
                                              double sc_prev, double &dsum, double &ssum,
                                              double &msum) {
    const double sv = acc / sc_cur;
    b.x_next64[i] = sv;
    ssum += sv * sv;
    dsum += fabs(yc / sc_cur - yp / sc_prev);
    msum += fabs(sv) * static_cast<double>(dg);
}

// The adjoint's row update, as AdjFin (adjoint.hip): mu = acc / dg (0 for
// a sink) and lambda_next = dL/dx + d mu (+ d (p . lambda) for a sink),
// rounded to the sweep's type; the delta against lambda_cur = lc, p .
// lambda_next, and |lambda_next| times the times row i is gathered (the
// next launch's mass).  The pass that computes mu writes mu alone.
template <int MODE>
__device__ __forceinline__ void blk_finish_adj(const BlkArgs &b, uint32_t i, double acc, double lc,
                                               double dg, double pv, double pl, double d,
                                               double &dsum, double &psum, double &msum) {
    typedef typename BlkTypes<MODE>::X V;
    const double mu = dg != 0.0 ? acc / dg : 0.0;
    if (b.mu_out) {
        b.mu_out[i] = mu;
        return;
    }
    const double v = b.cvec[i] + d * mu + (dg == 0.0 ? d * pl : 0.0);
    const V vv = static_cast<V>(v);
    dsum += fabs(static_cast<double>(vv) - lc);
    if constexpr (BlkTypes<MODE>::kTwo)
        b.x_next64[i] = vv;
    else
        b.x_next[i] = vv;
    psum += pv * static_cast<double>(vv);
    msum += fabs(static_cast<double>(vv)) * static_cast<double>(b.odeg[i]);
}

// Sum of a 64-bit value over each aligned group of 16 lanes.
__device__ __forceinline__ unsigned long long sum16(unsigned long long v) {
#pragma unroll
    for (int o = 8; o > 0; o >>= 1) {
        const uint32_t lo = __shfl_xor(static_cast<uint32_t>(v), o, kWave);
        const uint32_t hi = __shfl_xor(static_cast<uint32_t>(v >> 32), o, kWave);
        v += (static_cast<unsigned long long>(hi) << 32) | lo;
    }
    return v;
}

// One sweep.  Workgroups take tiles from bs->next_tile until none is left,
// in row order (taking the tiles with the most work first measured slower:
// 168.6 against 151.9 us per sweep at R-MAT 22).
// Per tile: the wavefronts walk its groups interleaved (wavefront w takes
// groups w, w + 8, ...), so the workgroup's gathers move up the source
// labels together; each lane adds its fixed-point term into the row's LDS
// accumulator, or, for a long row, into its own lane's slot of that row
// (rotated by the row index, so neither the adds nor the later column sums
// meet on one bank).  Then the long rows' lane slots are summed into their
// int64 slots in global memory, and the tile's rows are finished, each
// accumulator read once and cleared for the next tile.  The tile's delta,
// dangling and mass partials go to its own entries of b.part, so the fold
// sums them in tile order whichever workgroup took the tile.
template <bool DPERS, int MODE, bool WGT>
__global__ __launch_bounds__((BlkShape<MODE, WGT>::kThreads)) __attribute__((amdgpu_waves_per_eu(4)))
void pr_sweep_blk(const BlkArgs b) {
    typedef BlkTypes<MODE> M;
    typedef typename M::Wt Wt;
    constexpr int NW = BlkShape<MODE, WGT>::kWaves;
    constexpr int NT = BlkShape<MODE, WGT>::kThreads;
    constexpr int NR = BlkShape<MODE, WGT>::kRows;
    constexpr int NB = BlkShape<MODE, WGT>::kBatch;
    constexpr bool EV = MODE == kBlkEv;
    constexpr bool UNEV = MODE == kBlkUFwd || MODE == kBlkUAdj;
    constexpr bool ADJ = MODE == kBlkAdj32 || MODE == kBlkAdj64 || MODE == kBlkUAdj;
    SweepState *st = b.st;
    // The refinement runs once, after the FP32 loop finished (its state says
    // done), and so does the adjoint's mu pass: they do not look at the flag.
    // A batch of graphs runs while one is iterating and none has bailed;
    // its mu pass (after the loop) runs over every tile.
    if constexpr (UNEV) {
        if (!b.mu_out && (b.u.active[0] == 0 || b.u.active[1] != 0)) return;
    } else if (MODE != kBlkRefine && !(ADJ && b.mu_out) && st->done) {
        return;
    }
    __shared__ unsigned long long acc[kBlkLds];
    // The FP64 sweep's low words (a dummy otherwise).
    __shared__ unsigned long long acc_lo[M::kTwo ? kBlkLds : 1];
    __shared__ double red[3 * NW];
    __shared__ uint32_t s_tile[2];
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & (kWave - 1);
    const uint32_t w = __builtin_amdgcn_readfirstlane(tid / kWave);
    // A batch of graphs takes its scale, dangling mass, damping and
    // personalisation per tile, from the tile's graph.
    const double mass = UNEV ? 0.0 : b.bs->mass;
    if (!UNEV && !(mass <= kBlkMassMax)) {
        if (blockIdx.x == 0 && tid == 0) {
            st->bail = 1u;
            st->done = 1u;
        }
        return;
    }
    int K = blk_scale(mass);
    double p_sink = (EV || UNEV) ? 0.0 : st->p_sink;
    double dmp = b.damping, pval = b.pval;
    const double sc_cur = EV ? b.ev->sc_cur : 1.0;
    const double sc_prev = EV ? b.ev->sc_prev : 1.0;
    uint32_t *ctr = &b.bs->next_tile;
    uint32_t nwork = b.ntiles;
    if constexpr (UNEV) {
        ctr = b.u.ctr;
        if (!b.mu_out) {
            nwork = __builtin_amdgcn_readfirstlane(*b.u.list_n);
            if (blockIdx.x == 0 && tid == 0) {
                *b.u.ctr_next = 0u;
                *b.u.list_n_next = 0u;
            }
        }
    }
    for (uint32_t i = tid; i < kBlkLds; i += NT) {
        acc[i] = 0ull;
        if (M::kTwo) acc_lo[i] = 0ull;
    }
    if (tid == 0) s_tile[0] = atomicAdd(ctr, 1u);
    __syncthreads();
    for (uint32_t it = 0;; ++it) {
        uint32_t t = __builtin_amdgcn_readfirstlane(s_tile[it & 1]);
        if (t >= nwork) break;
        if constexpr (UNEV) {
            ConstU32 *lst = (ConstU32 *)b.u.list;
            if (!b.mu_out) t = __builtin_amdgcn_readfirstlane(lst[t]);
            const uint32_t gi = __builtin_amdgcn_readfirstlane(b.u.tile_graph[t]);
            const double mg = b.u.mass[gi];
            // A graph whose mass the fixed point cannot hold sums garbage
            // here; its fold makes the batch bail.
            K = mg <= kBlkMassMax ? blk_scale(mg) : 0;
            p_sink = b.u.ps[gi];
            dmp = b.u.damp[gi];
            pval = 1.0 / static_cast<double>(b.u.voff[gi + 1] - b.u.voff[gi]);
        }
        const BlkTile T = blk_tile(b.tiles, t);
        // Groups come in blocks of kBlkQuad: lane l's 16-byte load of block q
        // holds its word of each of the block's groups.  Wavefront w takes
        // blocks w, w + NW, ..., NB blocks per batch; the next batch's
        // words and bases are loaded while this batch's gathers are in
        // flight.  A batch's blocks past the tile are neither loaded nor
        // gathered.  The words are read once per sweep: non-temporal loads
        // keep them from evicting the gathered contributions.
        const uint32_t q0 = T.g0 / kBlkQuad;
        const uint32_t nq = (T.g1 - T.g0) / kBlkQuad;
        ConstU32 *gbase = (ConstU32 *)b.gbase;
        const u32x4 *wq = reinterpret_cast<const u32x4 *>(b.words);
        u32x4 W[NB];
        uint32_t base[NB * kBlkQuad];
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            if (w + j * NW >= nq) continue;
            const uint32_t q = min(q0 + w + j * NW, b.qlast);
            W[j] = __builtin_nontemporal_load(wq + (size_t)q * kWave + lane);
#pragma unroll
            for (int k = 0; k < kBlkQuad; ++k) base[j * kBlkQuad + k] = gbase[q * kBlkQuad + k];
        }
        // This thread's rows of the tile (r0 + tid + 512 k) and what the
        // finisher reads of them, loaded now: they arrive under the gathers
        // and the finisher waits on no load.  Loads past the tile read a
        // clamped row and are not used.  pr is the personalisation, or for
        // the eigenvector y_{L-2}.
        typename M::X xr[NR];
        typename std::conditional<WGT || ADJ, double, uint32_t>::type dr[NR];
        double pr[NR];
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            // A wavefront whose 64 rows all lie past the tile loads nothing
            // (the tiles of the hub rows hold a few dozen rows).
            if (T.r0 + (tid & ~(kWave - 1u)) + k * NT >= T.r1) continue;
            const uint32_t i = min(T.r0 + tid + k * NT, b.n - 1);
            if constexpr (M::kTwo)
                xr[k] = b.x_cur64[i];
            else
                xr[k] = b.x_cur[i];
            if constexpr (WGT && !EV)
                dr[k] = b.wdeg[i];
            else if constexpr (ADJ)
                dr[k] = static_cast<double>(b.row_off[i + 1] - b.row_off[i]);
            else
                dr[k] = b.odeg[i];
            if constexpr (EV)
                pr[k] = b.x_prev64[i];
            else
                pr[k] = DPERS ? b.pers[i] : pval;
        }
        // The tile's long rows inside its row range are finished by the
        // fold: their row accumulators (no entry adds into them) are marked.
        if (tid < T.nlong) {
            const uint32_t r = b.long_rows[b.ltab[T.lfirst + tid]];
            if (r >= T.r0 && r < T.r1) acc[r - T.r0] = kBlkMarked;
        }
        for (uint32_t qi = w; qi < nq; qi += NW * NB) {
            typename M::C v[NB * kBlkQuad];
            uint32_t at[NB * kBlkQuad];
            // A weighted sweep's weights of this batch, stored lane-major as
            // the words are: lane l's kBlkQuad weights of block q side by
            // side.  Read once per sweep, like the words.
            Wt wt[WGT ? NB * kBlkQuad : 1];
            if constexpr (WGT) {
                const Wt *wb = static_cast<const Wt *>(b.wblk);
#pragma unroll
                for (int j = 0; j < NB; ++j) {
                    if (qi + j * NW >= nq) continue;
                    const uint32_t q = min(q0 + qi + j * NW, b.qlast);
                    const Wt *at4 = wb + ((size_t)q * kWave + lane) * kBlkQuad;
                    if constexpr (sizeof(Wt) == 4) {
                        const f32x4 w4 = __builtin_nontemporal_load(
                            reinterpret_cast<const f32x4 *>(at4));
#pragma unroll
                        for (int k = 0; k < kBlkQuad; ++k) wt[j * kBlkQuad + k] = w4[k];
                    } else {
                        const f64x2 wa = __builtin_nontemporal_load(
                            reinterpret_cast<const f64x2 *>(at4));
                        const f64x2 wc = __builtin_nontemporal_load(
                            reinterpret_cast<const f64x2 *>(at4) + 1);
                        wt[j * kBlkQuad + 0] = wa[0];
                        wt[j * kBlkQuad + 1] = wa[1];
                        wt[j * kBlkQuad + 2] = wc[0];
                        wt[j * kBlkQuad + 3] = wc[1];
                    }
                }
            }
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                // Blocks past the tile's last gather nothing (the last batch
                // of a wavefront is often partial).
                if (qi + j * NW >= nq) continue;
#pragma unroll
                for (int k = 0; k < kBlkQuad; ++k) {
                    const uint32_t word = W[j][k];
                    const uint32_t src = base[j * kBlkQuad + k] + (word >> kBlkCodeBits);
                    if constexpr (M::kTwo)
                        v[j * kBlkQuad + k] = b.c_cur64[src];
                    else
                        v[j * kBlkQuad + k] = b.c_cur[src];
                    at[j * kBlkQuad + k] = word & kBlkCodeMask;
                }
            }
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                // The next batch's block j, when the tile has it.
                if (qi + (NB + j) * NW >= nq) continue;
                const uint32_t q = min(q0 + qi + (NB + j) * NW, b.qlast);
                W[j] = __builtin_nontemporal_load(wq + (size_t)q * kWave + lane);
#pragma unroll
                for (int k = 0; k < kBlkQuad; ++k) base[j * kBlkQuad + k] = gbase[q * kBlkQuad + k];
            }
#pragma unroll
            for (int j = 0; j < NB; ++j)
                if (qi + j * NW < nq) {
#pragma unroll
                    for (int k = 0; k < kBlkQuad; ++k) {
                        const uint32_t a = at[j * kBlkQuad + k];
                        if constexpr (M::kTwo) {
                            unsigned long long hi, lo;
                            double term = v[j * kBlkQuad + k];
                            if constexpr (WGT) term *= wt[j * kBlkQuad + k];
                            blk_fixed2(term, K, hi, lo);
                            __hip_atomic_fetch_add(&acc[a], hi, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_WORKGROUP);
                            __hip_atomic_fetch_add(&acc_lo[a], lo, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_WORKGROUP);
                        } else {
                            double term = static_cast<double>(v[j * kBlkQuad + k]);
                            if constexpr (WGT) term *= static_cast<double>(wt[j * kBlkQuad + k]);
                            __hip_atomic_fetch_add(&acc[a], blk_fixed_d(term, K),
                                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                        }
                    }
                }
        }
        __syncthreads();
        if (tid == 0) s_tile[(it + 1) & 1] = atomicAdd(ctr, 1u);
        if (T.nlong) {
            // Long row j's 64 lane slots, 16 threads a row, 4 slots each.
            const uint32_t j = tid >> 4, part = tid & 15;
            unsigned long long sl = 0, sll = 0;
            if (j < T.nlong) {
#pragma unroll
                for (uint32_t l = 0; l < 4; ++l) {
                    const uint32_t q = kBlkAcc + j * kWave + part * 4 + l;
                    sl += acc[q];
                    acc[q] = 0ull;
                    if (M::kTwo) {
                        sll += acc_lo[q];
                        acc_lo[q] = 0ull;
                    }
                }
            }
            sl = sum16(sl);
            if (M::kTwo) sll = sum16(sll);
            if (j < T.nlong && part == 0) {
                const uint32_t slot = b.ltab[T.lfirst + j];
                atomicAdd(&b.slots[slot], sl);
                if (M::kTwo) atomicAdd(&b.slots_lo[slot], sll);
            }
        }
        double dsum = 0.0, ssum = 0.0, msum = 0.0;
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            const uint32_t i = T.r0 + tid + k * NT;
            if (i < T.r1) {
                const uint32_t q = i - T.r0;
                const unsigned long long a = acc[q];
                acc[q] = 0ull;
                unsigned long long al = 0;
                if (M::kTwo) {
                    al = acc_lo[q];
                    acc_lo[q] = 0ull;
                }
                if (a != kBlkMarked) {
                    const double v = M::kTwo ? blk_value2(a, al, K) : blk_value(a, K);
                    if constexpr (EV)
                        blk_finish_ev(b, i, v, xr[k], pr[k], dr[k], sc_cur, sc_prev, dsum, ssum,
                                      msum);
                    else if constexpr (ADJ)
                        blk_finish_adj<MODE>(b, i, v, static_cast<double>(xr[k]),
                                             static_cast<double>(dr[k]), pr[k], p_sink, dmp,
                                             dsum, ssum, msum);
                    else
                        blk_finish<MODE>(b, i, v, static_cast<double>(xr[k]),
                                         static_cast<double>(dr[k]), pr[k], p_sink, dmp, dsum,
                                         ssum, msum);
                }
            }
        }
        dsum = wave_sum(dsum);
        ssum = wave_sum(ssum);
        msum = wave_sum(msum);
        if (lane == 0) {
            red[w] = dsum;
            red[NW + w] = ssum;
            red[2 * NW + w] = msum;
        }
        __syncthreads();
        if (tid == 0) {
            double d = 0.0, s = 0.0, ms = 0.0;
            for (int k = 0; k < NW; ++k) {
                d += red[k];
                s += red[NW + k];
                ms += red[2 * NW + k];
            }
            b.part[t] = d;
            b.part[b.ntiles + t] = s;
            b.part[2 * b.ntiles + t] = ms;
        }
    }
}

__device__ __forceinline__ void block_sum3(double &a, double &b, double &c, double *red) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    a = wave_sum(a);
    b = wave_sum(b);
    c = wave_sum(c);
    if (lane == 0) {
        red[wid] = a;
        red[4 + wid] = b;
        red[8 + wid] = c;
    }
    __syncthreads();
    a = (red[0] + red[1]) + (red[2] + red[3]);
    b = (red[4] + red[5]) + (red[6] + red[7]);
    c = (red[8] + red[9]) + (red[10] + red[11]);
    __syncthreads();
}

// Completes the sweep: the long rows (one thread each; each slot is read and
// cleared for the next sweep), then, in the last workgroup to arrive, the
// tile partials in tile order and this kernel's in workgroup order, the
// convergence decision and the next sweep's mass.
template <bool DPERS, int MODE, bool WGT>
__global__ __launch_bounds__(kBlock) void pr_fold_blk(const BlkArgs b) {
    constexpr bool EV = MODE == kBlkEv;
    constexpr bool ADJ = MODE == kBlkAdj32 || MODE == kBlkAdj64;
    SweepState *st = b.st;
    if (MODE != kBlkRefine && !(ADJ && b.mu_out) && st->done) return;
    __shared__ double red[12];
    __shared__ uint32_t flag;
    const int K = blk_scale(b.bs->mass);
    const double p_sink = EV ? 0.0 : st->p_sink;
    const double sc_cur = EV ? b.ev->sc_cur : 1.0;
    const double sc_prev = EV ? b.ev->sc_prev : 1.0;
    double d = 0.0, s = 0.0, ms = 0.0;
    for (uint32_t k = blockIdx.x * kBlock + threadIdx.x; k < b.nlong; k += gridDim.x * kBlock) {
        const uint32_t r = b.long_rows[k];
        const unsigned long long hi = b.slots[k];
        b.slots[k] = 0ull;
        const double dg = WGT && !EV ? b.wdeg[r]
                          : ADJ       ? static_cast<double>(b.row_off[r + 1] - b.row_off[r])
                                      : static_cast<double>(b.odeg[r]);
        const double pv = DPERS ? b.pers[r] : b.pval;
        if constexpr (BlkTypes<MODE>::kTwo) {
            const unsigned long long lo = b.slots_lo[k];
            b.slots_lo[k] = 0ull;
            const double v = blk_value2(hi, lo, K);
            if constexpr (EV)
                blk_finish_ev(b, r, v, b.x_cur64[r], b.x_prev64[r], b.odeg[r], sc_cur, sc_prev, d,
                              s, ms);
            else if constexpr (ADJ)
                blk_finish_adj<MODE>(b, r, v, b.x_cur64[r], dg, pv, p_sink, b.damping, d, s, ms);
            else
                blk_finish<MODE>(b, r, v, b.x_cur64[r], dg, pv, p_sink, b.damping, d, s, ms);
        } else if constexpr (ADJ) {
            blk_finish_adj<MODE>(b, r, blk_value(hi, K), static_cast<double>(b.x_cur[r]), dg, pv,
                                 p_sink, b.damping, d, s, ms);
        } else {
            blk_finish<MODE>(b, r, blk_value(hi, K), static_cast<double>(b.x_cur[r]), dg, pv,
                             p_sink, b.damping, d, s, ms);
        }
    }
    // The mu pass has no partials and decides nothing.
    if (ADJ && b.mu_out) return;
    block_sum3(d, s, ms, red);
    double *fp = b.fold_part;
    if (threadIdx.x == 0) {
        fp[blockIdx.x] = d;
        fp[gridDim.x + blockIdx.x] = s;
        fp[2 * gridDim.x + blockIdx.x] = ms;
    }
    if (!last_to_arrive(&st->counter, &flag)) return;
    d = 0.0;
    s = 0.0;
    ms = 0.0;
    for (uint32_t k = threadIdx.x; k < b.ntiles; k += kBlock) {
        d += b.part[k];
        s += b.part[b.ntiles + k];
        ms += b.part[2 * b.ntiles + k];
    }
    for (uint32_t k = threadIdx.x; k < gridDim.x; k += kBlock) {
        d += fp[k];
        s += fp[gridDim.x + k];
        ms += fp[2 * gridDim.x + k];
    }
    block_sum3(d, s, ms, red);
    if (EV && threadIdx.x == 0) {
        // d: delta of the previous sweep; s: squared norm of this one.
        ev_close(b.ev, s, d, b.eps_abs, b.max_iter);
        b.bs->mass = b.wblk ? ms * __longlong_as_double(static_cast<long long>(b.bs->wmax)) : ms;
        b.bs->next_tile = 0;
    }
    if (ADJ && threadIdx.x == 0) {
        // As adj_fold: s is p . lambda of the new iterate; the stop rule is
        // relative to ||dL/dx||_1 (st->aux).
        const uint64_t it = st->iter + 1;
        st->delta = d;
        st->p_sink = s;
        st->iter = it;
        st->launches += 1;
        const double thr = fmax(b.eps_abs, b.eps_rel) * (st->aux > 0.0 ? st->aux : 1.0);
        st->done = (!(d >= thr) || (b.max_iter > 0 && it >= b.max_iter)) ? 1u : 0u;
        st->counter = 0;
        b.bs->mass = b.wblk ? ms * __longlong_as_double(static_cast<long long>(b.bs->wmax)) : ms;
        b.bs->next_tile = 0;
    }
    if (!EV && !ADJ && threadIdx.x == 0) {
        const uint64_t it = st->iter + 1;
        const double thr = fmax(b.eps_abs, b.eps_rel * st->aux);
        st->delta = d;
        st->p_sink = s;
        st->iter = it;
        st->launches += 1;
        st->done = (!(d >= thr) || (b.max_iter > 0 && it >= b.max_iter)) ? 1u : 0u;
        st->counter = 0;
        b.bs->mass = ms;
        b.bs->next_tile = 0;
    }
}

// The eigenvector's and the adjoint's start: zero the long-row sums, then
// the mass of the start vector v (sum of |v| * out-degree, times the largest
// weight when weighted), per workgroup and then in workgroup order by the
// last to arrive.
__global__ __launch_bounds__(kBlock) void ev_start_blk(const BlkArgs b) {
    __shared__ double red[4];
    __shared__ uint32_t flag;
    for (uint32_t k = blockIdx.x * kBlock + threadIdx.x; k < b.nlong; k += gridDim.x * kBlock) {
        b.slots[k] = 0ull;
        b.slots_lo[k] = 0ull;
    }
    double m = 0.0;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < b.n; i += gridDim.x * kBlock)
        m += fabs(b.x_cur64[i]) * static_cast<double>(b.odeg[i]);
    m = block_sum(m, red);
    if (threadIdx.x == 0) b.fold_part[blockIdx.x] = m;
    if (!last_to_arrive(&b.st->counter, &flag)) return;
    m = 0.0;
    for (uint32_t k = threadIdx.x; k < gridDim.x; k += kBlock) m += b.fold_part[k];
    m = block_sum(m, red);
    if (threadIdx.x == 0) {
        b.bs->mass = b.wblk ? m * __longlong_as_double(static_cast<long long>(b.bs->wmax)) : m;
        b.bs->next_tile = 0;
        b.st->bail = 0;
        b.st->counter = 0;
    }
}

__global__ __launch_bounds__(kBlock) void pr_start_blk(const BlkArgs b) {
    for (uint32_t k = blockIdx.x * kBlock + threadIdx.x; k < b.nlong; k += gridDim.x * kBlock) {
        b.slots[k] = 0ull;
        if (b.slots_lo) b.slots_lo[k] = 0ull;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        b.bs->mass = b.st->aux;
        b.bs->next_tile = 0;
        b.st->bail = 0;
    }
}

// The uneven batch's start: each graph's first mass, the sum over its
// vertices of |x0| (times the times the vertex is gathered when `gathered`,
// and then times the largest weight when weighted), one workgroup per graph
// in a fixed order; the long-row sums cleared.
__global__ __launch_bounds__(kBlock) void u_start_blk(const BlkArgs b, const float *x0,
                                                      int gathered) {
    __shared__ double red[4];
    const BlkUneven &u = b.u;
    for (uint32_t k = blockIdx.x * kBlock + threadIdx.x; k < b.nlong; k += gridDim.x * kBlock)
        b.slots[k] = 0ull;
    const double wm = gathered && b.wblk
                          ? __longlong_as_double(static_cast<long long>(b.bs->wmax))
                          : 1.0;
    for (uint32_t gi = blockIdx.x; gi < u.ngraphs; gi += gridDim.x) {
        double m = 0.0;
        for (uint32_t v = u.voff[gi] + threadIdx.x; v < u.voff[gi + 1]; v += kBlock)
            m += fabs(static_cast<double>(x0[v])) *
                 (gathered ? static_cast<double>(b.odeg[v]) : 1.0);
        m = block_sum(m, red);
        if (threadIdx.x == 0) u.mass[gi] = m * wm;
    }
}

// The uneven batch's fold: one wavefront per graph still iterating.  It sums
// the graph's tile partials in tile order (a strided share per lane, then the
// wavefront's tree), finishes the graph's long rows, and decides the graph's
// convergence as the wave engine's batch fold does: the forward stops when
// delta < eps, the adjoint when delta < max(eps, eps_rel) ||dL/dx_g||_1, both
// at max_iter sweeps when it is non-zero.  A graph that goes on lists its
// tiles for the next sweep, and its mass is the new iterate's.  A mass the
// fixed point cannot hold sets the bail word: every later launch returns,
// and the batch runs again on the wave engine.
template <int MODE, bool WGT>
__global__ __launch_bounds__(kBlock) void u_fold_blk(const BlkArgs b) {
    constexpr bool ADJ = MODE == kBlkUAdj;
    const BlkUneven &u = b.u;
    // The adjoint's mu pass finishes every graph's long rows and decides
    // nothing.
    if (!b.mu_out && (u.active[0] == 0 || u.active[1] != 0)) return;
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t gi =
        __builtin_amdgcn_readfirstlane(blockIdx.x * kWavesPerBlock + threadIdx.x / kWave);
    if (gi >= u.ngraphs || (!b.mu_out && u.done[gi])) return;
    const double mg = u.mass[gi];
    const int K = mg <= kBlkMassMax ? blk_scale(mg) : 0;
    const double dmp = u.damp[gi], ps = u.ps[gi];
    const double pv = 1.0 / static_cast<double>(u.voff[gi + 1] - u.voff[gi]);
    const uint32_t nt = b.ntiles;
    double d = 0.0, s = 0.0, ms = 0.0;
    for (uint32_t t = u.tb[gi] + lane; t < u.tb[gi + 1]; t += kWave) {
        d += b.part[t];
        s += b.part[nt + t];
        ms += b.part[2 * nt + t];
    }
    for (uint32_t k = u.lb[gi] + lane; k < u.lb[gi + 1]; k += kWave) {
        const uint32_t r = b.long_rows[k];
        const double v = blk_value(b.slots[k], K);
        b.slots[k] = 0ull;
        const double dg = WGT   ? b.wdeg[r]
                          : ADJ ? static_cast<double>(b.row_off[r + 1] - b.row_off[r])
                                : static_cast<double>(b.odeg[r]);
        if constexpr (ADJ)
            blk_finish_adj<MODE>(b, r, v, static_cast<double>(b.x_cur[r]), dg, pv, ps, dmp, d, s,
                                 ms);
        else
            blk_finish<MODE>(b, r, v, static_cast<double>(b.x_cur[r]), dg, pv, ps, dmp, d, s, ms);
    }
    if (ADJ && b.mu_out) return;
    d = wave_sum(d);
    s = wave_sum(s);
    ms = wave_sum(ms);
    bool more = false;
    if (lane == 0) {
        const uint64_t it = u.iter[gi] + 1;
        const double thr =
            ADJ ? fmax(b.eps_abs, b.eps_rel) * (u.scale[gi] > 0.0 ? u.scale[gi] : 1.0) : b.eps_abs;
        const double nm =
            ADJ && WGT ? ms * __longlong_as_double(static_cast<long long>(b.bs->wmax)) : ms;
        u.iter[gi] = it;
        u.delta[gi] = d;
        u.ps[gi] = s;
        u.mass[gi] = nm;
        if (!(nm <= kBlkMassMax)) atomicOr(u.active + 1, 1u);
        if (!(d >= thr) || (b.max_iter > 0 && it >= b.max_iter)) {
            u.done[gi] = 1u;
            atomicSub(u.active, 1u);
        } else {
            more = true;
        }
    }
    if (__ballot(more)) {
        const uint32_t t0 = u.tb[gi], cnt = u.tb[gi + 1] - t0;
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(u.list_n_next, cnt);
        base = __shfl(base, 0, kWave);
        for (uint32_t i = lane; i < cnt; i += kWave) u.list[base + i] = t0 + i;
    }
}

// odeg[i] = off[i + 1] - off[i]: the out-degree in sweep labels, read by the
// block sweeps' finishers in one load.
__global__ __launch_bounds__(kBlock) void blk_out_degree(const uint32_t *__restrict__ off,
                                                         uint32_t n, uint32_t *__restrict__ odeg) {
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
        odeg[i] = off[i + 1] - off[i];
}

// ------------------------------------------------------------------ host

namespace {

int device_cus() {
    static std::mutex mu;
    static int cus = -1;
    std::lock_guard<std::mutex> lk(mu);
    if (cus < 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            cus = 0;
    }
    return cus > 0 ? cus : 256;
}

// A tile as the host builds it: rows [r0, r1) with entries [e0, e1) of the
// orientation, or a chunk [e0, e1) of one long row (chunk set).
struct HostTile {
    uint32_t r0, r1;
    uint64_t e0, e1;
    uint32_t lfirst, nlong;
    bool chunk;
};

// Runs f(k) for k in [0, count) on the host's cores.
template <typename F>
void parallel_for(uint32_t count, F f) {
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (count < 4 || nt == 1) {
        for (uint32_t k = 0; k < count; ++k) f(k);
        return;
    }
    std::atomic<uint32_t> next{0};
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nt; ++t)
        pool.emplace_back([&]() {
            for (uint32_t k; (k = next.fetch_add(1)) < count;) f(k);
        });
    for (auto &th : pool) th.join();
}

// The block tiles of orientation o into B; `gathered_off` (device, n + 1
// offsets) gives the times each vertex is gathered as its row length there;
// with `bounds`, no tile spans one of those rows.  Waits for the stream.
gt_status build_blocks(gt_graph *g, const Orient &o, const uint32_t *gathered_off,
                       uint32_t tile_nnz, uint32_t long_min, const std::vector<uint32_t> *bounds,
                       Blocks &B) {
    B.ready = false;
    B.batch_bounds.clear();
    const uint32_t n = g->n;
    const std::vector<uint32_t> &off = o.h_off;
    const uint64_t m = off[n];
    std::vector<uint32_t> idx(m), eid(m);
    if (m) {
        GT_HIP(hipMemcpy(idx.data(), o.idx.ptr, sizeof(uint32_t) * m, hipMemcpyDeviceToHost));
        GT_HIP(hipMemcpy(eid.data(), o.eid.ptr, sizeof(uint32_t) * m, hipMemcpyDeviceToHost));
    }
    // Each row's entries by source (stable: parallel edges keep their
    // storage order), with their edge ids, where a row is not sorted yet (a
    // relabelled orientation's rows are).
    {
        constexpr uint32_t kRowsPerTask = 4096;
        parallel_for((n + kRowsPerTask - 1) / kRowsPerTask, [&](uint32_t k) {
            std::vector<std::pair<uint32_t, uint32_t>> tmp;
            const uint32_t r1 = std::min<uint32_t>(n, (k + 1) * kRowsPerTask);
            for (uint32_t r = k * kRowsPerTask; r < r1; ++r) {
                const uint64_t lo = off[r], hi = off[r + 1];
                if (std::is_sorted(idx.begin() + lo, idx.begin() + hi)) continue;
                tmp.resize(hi - lo);
                for (uint64_t e = lo; e < hi; ++e) tmp[e - lo] = {idx[e], eid[e]};
                std::stable_sort(tmp.begin(), tmp.end(),
                                 [](const std::pair<uint32_t, uint32_t> &a,
                                    const std::pair<uint32_t, uint32_t> &b) {
                                     return a.first < b.first;
                                 });
                for (uint64_t e = lo; e < hi; ++e) {
                    idx[e] = tmp[e - lo].first;
                    eid[e] = tmp[e - lo].second;
                }
            }
        });
    }

    // Tiles, in row order.
    std::vector<HostTile> ht;
    std::vector<uint32_t> ltab, long_rows, code(n);
    uint32_t r0 = 0, nl = 0, lfirst = 0;
    uint64_t ent = 0;
    auto close = [&](uint32_t r1) {
        if (r1 > r0) ht.push_back(HostTile{r0, r1, off[r0], off[r1], lfirst, nl, false});
        r0 = r1;
        ent = 0;
        nl = 0;
        lfirst = (uint32_t)ltab.size();
    };
    size_t bi = 0;
    for (uint32_t r = 0; r < n; ++r) {
        if (bounds) {
            bool cut = false;
            for (; bi < bounds->size() && (*bounds)[bi] <= r; ++bi) cut |= (*bounds)[bi] == r;
            if (cut) close(r);
        }
        const uint64_t L = (uint64_t)off[r + 1] - off[r];
        if (L > tile_nnz) {
            close(r);
            const uint32_t k = (uint32_t)long_rows.size();
            long_rows.push_back(r);
            for (uint64_t e = off[r]; e < off[r + 1]; e += tile_nnz) {
                ht.push_back(HostTile{r, r, e, std::min<uint64_t>(e + tile_nnz, off[r + 1]),
                                      (uint32_t)ltab.size(), 1, true});
                ltab.push_back(k);
            }
            r0 = r + 1;
            lfirst = (uint32_t)ltab.size();
            continue;
        }
        const bool lng = L > long_min;
        if (ent + L > tile_nnz || r + 1 - r0 > kBlkRows || (lng && nl == kBlkLongMax)) close(r);
        if (lng) {
            code[r] = kBlkKeyLong0 + nl++;
            ltab.push_back((uint32_t)long_rows.size());
            long_rows.push_back(r);
        } else {
            code[r] = r - r0;
        }
        ent += L;
    }
    close(n);
    const uint32_t ntiles = (uint32_t)ht.size();

    // Each tile's entries as (source << code bits | code), sorted; then its
    // groups of 64: a group ends after 64 entries or before a source that
    // lies kBlkSpan or more past the group's first.
    std::vector<uint64_t> key(m);
    std::vector<uint32_t> ngr(ntiles + 1, 0);
    parallel_for(ntiles, [&](uint32_t t) {
        const HostTile &T = ht[t];
        if (T.chunk) {
            for (uint64_t e = T.e0; e < T.e1; ++e)
                key[e] = ((uint64_t)idx[e] << kBlkCodeBits) | kBlkKeyLong0;
        } else {
            for (uint32_t r = T.r0; r < T.r1; ++r)
                for (uint64_t e = off[r]; e < off[r + 1]; ++e)
                    key[e] = ((uint64_t)idx[e] << kBlkCodeBits) | code[r];
        }
        std::sort(key.begin() + T.e0, key.begin() + T.e1);
        uint32_t groups = 0, cnt = 0;
        uint64_t base = 0;
        for (uint64_t e = T.e0; e < T.e1; ++e) {
            const uint64_t src = key[e] >> kBlkCodeBits;
            if (cnt == 0 || cnt == kWave || src - base >= kBlkSpan) {
                ++groups;
                cnt = 0;
                base = src;
            }
            ++cnt;
        }
        ngr[t] = (groups + kBlkQuad - 1) / kBlkQuad * kBlkQuad;
    });
    std::vector<uint64_t> g0(ntiles + 1, 0);
    for (uint32_t t = 0; t < ntiles; ++t) g0[t + 1] = g0[t] + ngr[t];
    const uint64_t ngroups = g0[ntiles];
    if (ngroups >= (1ull << 32) / kWave)
        return fail(GT_ERR_UNSUPPORTED, "block tiles: too many entry groups");
    B.h_tile_row.resize(ntiles);
    for (uint32_t t = 0; t < ntiles; ++t) B.h_tile_row[t] = ht[t].r0;
    B.h_long_rows = long_rows;
    B.bounds = bounds ? *bounds : std::vector<uint32_t>();
    // Padding groups (to whole blocks) hold null words over source 0.  A
    // graph without entries still gets one block: a sweep's first loads
    // read block 0 whether or not its tile has groups.
    const uint64_t nstore = std::max<uint64_t>(kBlkQuad, ngroups);
    std::vector<uint32_t> words(nstore * kWave, kBlkNull);
    std::vector<uint32_t> gbase(nstore, 0);
    // The edge id of each word, for weights staged in block order: an
    // entry's position in its row (rows are sorted by source) is the first
    // of its source there, plus the entries of the same source and row
    // before it in the tile's order.  `seen` checks that every entry of the
    // orientation is taken exactly once.
    std::vector<uint32_t> weid(nstore * kWave, 0);
    std::vector<uint8_t> seen(m, 0);
    std::atomic<bool> bad{false};
    parallel_for(ntiles, [&](uint32_t t) {
        const HostTile &T = ht[t];
        uint64_t gidx = g0[t];
        uint32_t cnt = 0;
        uint64_t base = 0;
        bool open = false;
        uint32_t dup = 0;
        for (uint64_t e = T.e0; e < T.e1; ++e) {
            const uint64_t src = key[e] >> kBlkCodeBits;
            const uint32_t c = (uint32_t)(key[e] & kBlkCodeMask);
            dup = (e > T.e0 && key[e] == key[e - 1]) ? dup + 1 : 0;
            uint64_t lo, hi;
            if (T.chunk) {
                lo = T.e0;
                hi = T.e1;
            } else {
                const uint32_t row = c < kBlkKeyLong0
                                         ? T.r0 + c
                                         : long_rows[ltab[T.lfirst + (c - kBlkKeyLong0)]];
                lo = off[row];
                hi = off[row + 1];
            }
            const uint64_t pos =
                (uint64_t)(std::lower_bound(idx.begin() + lo, idx.begin() + hi, (uint32_t)src) -
                           idx.begin()) + dup;
            if (pos >= hi || idx[pos] != src || seen[pos]) {
                bad = true;
                break;
            }
            seen[pos] = 1;
            if (!open || cnt == kWave || src - base >= kBlkSpan) {
                if (open) ++gidx;
                open = true;
                cnt = 0;
                base = src;
                gbase[gidx] = (uint32_t)src;
            }
            // A long row's entry adds into its lane's slot of the row.
            const uint32_t j = c - kBlkKeyLong0;
            const uint32_t at =
                j < kBlkLongMax ? kBlkAcc + j * kWave + ((cnt + j) & (kWave - 1)) : c;
            // Block gidx / 4 holds, for each lane, its word of the block's
            // four groups side by side.
            const uint64_t slot =
                (gidx / kBlkQuad) * (kWave * kBlkQuad) + cnt * kBlkQuad + gidx % kBlkQuad;
            words[slot] = ((uint32_t)(src - base) << kBlkCodeBits) | at;
            weid[slot] = eid[pos];
            ++cnt;
        }
    });
    if (bad) return fail(GT_ERR_UNSUPPORTED, "block tiles: an entry's edge id was not found");
    std::vector<BlkTile> tiles(std::max<uint32_t>(1, ntiles));
    for (uint32_t t = 0; t < ntiles; ++t) {
        const HostTile &T = ht[t];
        tiles[t] = BlkTile{T.r0, T.r1, (uint32_t)g0[t], (uint32_t)g0[t + 1], T.lfirst, T.nlong,
                           {0, 0}};
    }
    if (ltab.empty()) ltab.push_back(0);
    const uint32_t nlong = (uint32_t)long_rows.size();
    if (long_rows.empty()) long_rows.push_back(0);

    GT_TRY(B.tiles.reserve(sizeof(BlkTile) * tiles.size()));
    GT_TRY(B.words.reserve(sizeof(uint32_t) * words.size()));
    GT_TRY(B.eids.reserve(sizeof(uint32_t) * weid.size()));
    GT_TRY(B.gbase.reserve(sizeof(uint32_t) * gbase.size()));
    GT_TRY(B.ltab.reserve(sizeof(uint32_t) * ltab.size()));
    GT_TRY(B.long_rows.reserve(sizeof(uint32_t) * long_rows.size()));
    GT_TRY(B.odeg.reserve(sizeof(uint32_t) * std::max<uint32_t>(1, n)));
    if (n) {
        hipLaunchKernelGGL(blk_out_degree, dim3(std::min<uint32_t>(2048, (n + kBlock - 1) / kBlock)),
                           dim3(kBlock), 0, g->stream, gathered_off, n, B.odeg.as<uint32_t>());
        GT_HIP(hipGetLastError());
        GT_HIP(hipStreamSynchronize(g->stream));
    }
    GT_HIP(hipMemcpy(B.tiles.ptr, tiles.data(), sizeof(BlkTile) * tiles.size(),
                     hipMemcpyHostToDevice));
    GT_HIP(hipMemcpy(B.words.ptr, words.data(), sizeof(uint32_t) * words.size(),
                     hipMemcpyHostToDevice));
    GT_HIP(hipMemcpy(B.eids.ptr, weid.data(), sizeof(uint32_t) * weid.size(),
                     hipMemcpyHostToDevice));
    GT_HIP(hipMemcpy(B.gbase.ptr, gbase.data(), sizeof(uint32_t) * gbase.size(),
                     hipMemcpyHostToDevice));
    GT_HIP(hipMemcpy(B.ltab.ptr, ltab.data(), sizeof(uint32_t) * ltab.size(),
                     hipMemcpyHostToDevice));
    GT_HIP(hipMemcpy(B.long_rows.ptr, long_rows.data(), sizeof(uint32_t) * long_rows.size(),
                     hipMemcpyHostToDevice));
    B.ntiles = ntiles;
    B.nlong = nlong;
    B.max_row = 0;
    for (uint32_t r = 0; r < n; ++r) B.max_row = std::max(B.max_row, off[r + 1] - off[r]);
    B.ngroups = ngroups;
    B.tile_nnz = tile_nnz;
    B.long_min = long_min;
    B.ready = true;
    return GT_OK;
}

}  // namespace

gt_status ensure_blocks(gt_graph *g, Relabeled &R, uint32_t tile_nnz, uint32_t long_min,
                        Blocks **out) {
    Blocks &B = R.blk;
    *out = &B;
    if (B.ready && B.tile_nnz == tile_nnz && B.long_min == long_min) return GT_OK;
    return build_blocks(g, R.pull, R.push_off.as<uint32_t>(), tile_nnz, long_min, nullptr, B);
}

gt_status ensure_blocks_cut(gt_graph *g, Orient &o, const uint32_t *gathered_off,
                            uint32_t tile_nnz, uint32_t long_min,
                            const std::vector<uint32_t> &bounds, Blocks **out) {
    Blocks &B = o.blk;
    *out = &B;
    if (B.ready && B.tile_nnz == tile_nnz && B.long_min == long_min && B.bounds == bounds)
        return GT_OK;
    return build_blocks(g, o, gathered_off, tile_nnz, long_min, &bounds, B);
}

// A weighted sweep's weights in block order: out[k] = w(eids[k]) converted
// to V, for every word slot.  A real entry (not padding) whose weight is
// negative or not finite sets *bad: the block sweeps bound their sums by the
// iterate's mass, which holds for non-negative weights only.  Each XCD takes
// a contiguous eighth of the slots (workgroup b runs on XCD b % 8), so the
// weights its slots gather -- those of the rows of a few tiles -- stay in
// its L2; the slot and edge-id streams are read once and the output written
// once, non-temporally, so they do not evict them.  Each thread keeps eight
// slots a stride apart in flight.
template <typename S, typename V>
__global__ __launch_bounds__(kBlock) void blk_stage_w(const uint32_t *__restrict__ words,
                                                      const uint32_t *__restrict__ eids,
                                                      const void *w, int kind, double value,
                                                      uint64_t count, V *__restrict__ out,
                                                      uint32_t *bad, unsigned long long *wmax) {
    uint64_t begin = 0, end = count, first = (uint64_t)blockIdx.x * kBlock + threadIdx.x,
             stride = (uint64_t)gridDim.x * kBlock;
    if (gridDim.x % kXcds == 0) {
        const uint32_t x = blockIdx.x % kXcds;
        begin = count * x / kXcds;
        end = count * (x + 1) / kXcds;
        first = begin + (uint64_t)(blockIdx.x / kXcds) * kBlock + threadIdx.x;
        stride = (uint64_t)(gridDim.x / kXcds) * kBlock;
    }
    bool neg = false;
    double top = 0.0;
    for (uint64_t j0 = first; j0 < end; j0 += 8 * stride) {
        uint32_t e[8], wd[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const uint64_t j = j0 + u * stride;
            e[u] = j < end ? __builtin_nontemporal_load(eids + j) : 0u;
            wd[u] = j < end ? __builtin_nontemporal_load(words + j) : kBlkNull;
        }
        // Padding slots gather nothing (their weight is never used).
        double x[8];
#pragma unroll
        for (int u = 0; u < 8; ++u)
            x[u] = j0 + u * stride >= end || (wd[u] & kBlkCodeMask) == kBlkNull ? 0.0
                   : kind == GT_MAP_CONSTANT ? value
                                             : static_cast<double>(static_cast<const S *>(w)[e[u]]);
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const uint64_t j = j0 + u * stride;
            if (j >= end) continue;
            const V v = static_cast<V>(x[u]);
            __builtin_nontemporal_store(v, out + j);
            if ((wd[u] & kBlkCodeMask) != kBlkNull) {
                if (!(x[u] >= 0.0 && x[u] <= 1.7976931348623157e308)) neg = true;
                top = fmax(top, static_cast<double>(v));
            }
        }
    }
    if (__any(neg) && (threadIdx.x & (kWave - 1)) == 0) atomicOr(bad, 1u);
    // Non-negative doubles order as their bit patterns.
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) top = fmax(top, __shfl_xor(top, o, kWave));
    if ((threadIdx.x & (kWave - 1)) == 0 && !neg)
        atomicMax(wmax, static_cast<unsigned long long>(__double_as_longlong(top)));
}

gt_status blk_stage_weights(gt_graph *g, const Blocks &B, const gt_prop *w, const void *src,
                            int dtype, bool f64, DevBuf &out, bool *usable) {
    const uint64_t count = std::max<uint64_t>(kBlkQuad, B.ngroups) * kWave;
    GT_TRY(out.reserve((f64 ? sizeof(double) : sizeof(float)) * count));
    GT_TRY(g->ws.blk_state.reserve(sizeof(BlkState)));
    BlkState *bs = g->ws.blk_state.as<BlkState>();
    hipStream_t s = g->stream;
    static_assert(offsetof(BlkState, wmax) == offsetof(BlkState, pad) + 4, "pad, wmax adjacent");
    GT_HIP(hipMemsetAsync(&bs->pad, 0, sizeof(uint32_t) + sizeof(unsigned long long), s));
    const unsigned grid = (unsigned)std::min<uint64_t>(8192, (count + kBlock - 1) / kBlock);
    const uint32_t *words = B.words.as<uint32_t>(), *eids = B.eids.as<uint32_t>();
    const int kind = w->kind;
    const double value = w->value;
    if (dtype == GT_F32) {
        if (f64)
            hipLaunchKernelGGL((blk_stage_w<float, double>), dim3(grid), dim3(kBlock), 0, s, words,
                               eids, src, kind, value, count, out.as<double>(), &bs->pad, &bs->wmax);
        else
            hipLaunchKernelGGL((blk_stage_w<float, float>), dim3(grid), dim3(kBlock), 0, s, words,
                               eids, src, kind, value, count, out.as<float>(), &bs->pad, &bs->wmax);
    } else {
        if (f64)
            hipLaunchKernelGGL((blk_stage_w<double, double>), dim3(grid), dim3(kBlock), 0, s,
                               words, eids, src, kind, value, count, out.as<double>(), &bs->pad, &bs->wmax);
        else
            hipLaunchKernelGGL((blk_stage_w<double, float>), dim3(grid), dim3(kBlock), 0, s,
                               words, eids, src, kind, value, count, out.as<float>(), &bs->pad, &bs->wmax);
    }
    GT_HIP(hipGetLastError());
    uint32_t bad = 0;
    GT_HIP(hipMemcpyAsync(&bad, &bs->pad, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    GT_HIP(hipStreamSynchronize(s));
    *usable = bad == 0;
    return GT_OK;
}

__global__ __launch_bounds__(kBlock) void blk_widen_w(const float *__restrict__ in, uint64_t count,
                                                      double *__restrict__ out) {
    for (uint64_t k = (uint64_t)blockIdx.x * kBlock + threadIdx.x; k < count;
         k += (uint64_t)gridDim.x * kBlock)
        out[k] = static_cast<double>(in[k]);
}

gt_status blk_widen_weights(gt_graph *g, const Blocks &B, const DevBuf &f32, DevBuf &out) {
    const uint64_t count = std::max<uint64_t>(kBlkQuad, B.ngroups) * kWave;
    GT_TRY(out.reserve(sizeof(double) * count));
    const unsigned grid = (unsigned)std::min<uint64_t>(8192, (count + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(blk_widen_w, dim3(grid), dim3(kBlock), 0, g->stream, f32.as<float>(), count,
                       out.as<double>());
    GT_HIP(hipGetLastError());
    return GT_OK;
}

gt_status blk_reserve(gt_graph *g, const Blocks &B) {
    Workspace &ws = g->ws;
    GT_TRY(ws.blk_slots.reserve(sizeof(unsigned long long) * std::max<uint32_t>(1, B.nlong)));
    GT_TRY(ws.blk_slots_lo.reserve(sizeof(unsigned long long) * std::max<uint32_t>(1, B.nlong)));
    GT_TRY(ws.blk_part.reserve(sizeof(double) * 3 * std::max<uint32_t>(1, B.ntiles)));
    GT_TRY(ws.blk_state.reserve(sizeof(BlkState)));
    GT_TRY(ws.partials.reserve(sizeof(double) * kPartials));
    return GT_OK;
}

BlkArgs blk_layout_args(gt_graph *g, const Blocks &B) {
    BlkArgs b{};
    b.tiles = B.tiles.as<BlkTile>();
    b.ntiles = B.ntiles;
    b.nlong = B.nlong;
    b.qlast = B.ngroups ? (uint32_t)(B.ngroups / kBlkQuad - 1) : 0u;
    b.n = g->n;
    b.words = B.words.as<uint32_t>();
    b.gbase = B.gbase.as<uint32_t>();
    b.ltab = B.ltab.as<uint32_t>();
    b.long_rows = B.long_rows.as<uint32_t>();
    b.odeg = B.odeg.as<uint32_t>();
    b.slots = g->ws.blk_slots.as<unsigned long long>();
    b.slots_lo = g->ws.blk_slots_lo.as<unsigned long long>();
    b.part = g->ws.blk_part.as<double>();
    b.fold_part = g->ws.partials.as<double>();
    b.bs = g->ws.blk_state.as<BlkState>();
    g->layout.block_tiles = B.ntiles;
    g->layout.block_long_rows = B.nlong;
    g->layout.block_groups = B.ngroups;
    return b;
}

void blk_launch_start(const BlkArgs &b, hipStream_t s) {
    hipLaunchKernelGGL(pr_start_blk, dim3(fold_grid(b.nlong)), dim3(kBlock), 0, s, b);
}

void blk_launch_start_ev(const BlkArgs &b, hipStream_t s) {
    const int grid = fit_grid(1024, std::max((b.n + kBlock - 1) / kBlock, b.nlong / kBlock + 1));
    hipLaunchKernelGGL(ev_start_blk, dim3(grid), dim3(kBlock), 0, s, b);
}

template <int MODE, bool WGT>
void launch_mode_w(const BlkArgs &b, bool dpers, int grid, hipStream_t s) {
    constexpr int nt = BlkShape<MODE, WGT>::kThreads;
    if (dpers)
        hipLaunchKernelGGL((pr_sweep_blk<true, MODE, WGT>), dim3(grid), dim3(nt), 0, s, b);
    else
        hipLaunchKernelGGL((pr_sweep_blk<false, MODE, WGT>), dim3(grid), dim3(nt), 0, s, b);
}
template <int MODE>
void launch_mode(const BlkArgs &b, bool dpers, int grid, hipStream_t s) {
    if (b.wblk)
        launch_mode_w<MODE, true>(b, dpers, grid, s);
    else
        launch_mode_w<MODE, false>(b, dpers, grid, s);
}

void blk_launch_sweep(const BlkArgs &b, bool dpers, int per_cu, int mode, hipStream_t s) {
    // The two-word sweeps' accumulators take 96 KB: one workgroup per CU.
    const bool two = mode == kBlkF64 || mode == kBlkEv || mode == kBlkAdj64;
    const int per = two ? 1 : std::max(1, per_cu);
    const int grid = fit_grid(device_cus() * per, b.ntiles);
    switch (mode) {
        case kBlkF32: launch_mode<kBlkF32>(b, dpers, grid, s); break;
        case kBlkRefine: launch_mode<kBlkRefine>(b, dpers, grid, s); break;
        case kBlkF64: launch_mode<kBlkF64>(b, dpers, grid, s); break;
        case kBlkAdj32: launch_mode<kBlkAdj32>(b, dpers, grid, s); break;
        case kBlkAdj64: launch_mode<kBlkAdj64>(b, dpers, grid, s); break;
        default:
            launch_mode<kBlkEv>(b, false, grid, s);
            break;
    }
}

template <int MODE, bool WGT>
void fold_mode_w(const BlkArgs &b, bool dpers, int grid, hipStream_t s) {
    if (dpers)
        hipLaunchKernelGGL((pr_fold_blk<true, MODE, WGT>), dim3(grid), dim3(kBlock), 0, s, b);
    else
        hipLaunchKernelGGL((pr_fold_blk<false, MODE, WGT>), dim3(grid), dim3(kBlock), 0, s, b);
}
template <int MODE>
void fold_mode(const BlkArgs &b, bool dpers, int grid, hipStream_t s) {
    if (b.wblk)
        fold_mode_w<MODE, true>(b, dpers, grid, s);
    else
        fold_mode_w<MODE, false>(b, dpers, grid, s);
}

void blk_launch_fold(const BlkArgs &b, bool dpers, int mode, hipStream_t s) {
    const int grid = fold_grid(b.nlong);
    switch (mode) {
        case kBlkF32: fold_mode<kBlkF32>(b, dpers, grid, s); break;
        case kBlkRefine: fold_mode<kBlkRefine>(b, dpers, grid, s); break;
        case kBlkF64: fold_mode<kBlkF64>(b, dpers, grid, s); break;
        case kBlkAdj32: fold_mode<kBlkAdj32>(b, dpers, grid, s); break;
        case kBlkAdj64: fold_mode<kBlkAdj64>(b, dpers, grid, s); break;
        default:
            fold_mode<kBlkEv>(b, false, grid, s);
            break;
    }
}

void blk_launch_ustart(const BlkArgs &b, const float *x0, bool gathered, hipStream_t s) {
    const int grid = fit_grid(2048, std::max(b.u.ngraphs, b.nlong / kBlock + 1));
    hipLaunchKernelGGL(u_start_blk, dim3(grid), dim3(kBlock), 0, s, b, x0, gathered ? 1 : 0);
}

template <int MODE>
void usweep_mode(const BlkArgs &b, int grid, hipStream_t s) {
    if (b.wblk)
        hipLaunchKernelGGL((pr_sweep_blk<false, MODE, true>), dim3(grid),
                           dim3(BlkShape<MODE, true>::kThreads), 0, s, b);
    else
        hipLaunchKernelGGL((pr_sweep_blk<false, MODE, false>), dim3(grid),
                           dim3(BlkShape<MODE, false>::kThreads), 0, s, b);
}

void blk_launch_usweep(const BlkArgs &b, int per_cu, int mode, hipStream_t s) {
    const int grid = fit_grid(device_cus() * std::max(1, per_cu), b.ntiles);
    if (mode == kBlkUAdj)
        usweep_mode<kBlkUAdj>(b, grid, s);
    else
        usweep_mode<kBlkUFwd>(b, grid, s);
}

template <int MODE>
void ufold_mode(const BlkArgs &b, hipStream_t s) {
    const dim3 grid((b.u.ngraphs + kWavesPerBlock - 1) / kWavesPerBlock);
    if (b.wblk)
        hipLaunchKernelGGL((u_fold_blk<MODE, true>), grid, dim3(kBlock), 0, s, b);
    else
        hipLaunchKernelGGL((u_fold_blk<MODE, false>), grid, dim3(kBlock), 0, s, b);
}

void blk_launch_ufold(const BlkArgs &b, int mode, hipStream_t s) {
    if (mode == kBlkUAdj)
        ufold_mode<kBlkUAdj>(b, s);
    else
        ufold_mode<kBlkUFwd>(b, s);
}

}  // namespace gt
