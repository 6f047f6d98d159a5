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
// level-synchronous pulls over row tiles (at most 64 rows and `tile_nnz`
// entries; longer rows cut into pieces):
//   forward, level L -> L + 1: each entry (u -> w) of a row w with unreached
//     lanes U gathers u's level-L mask; an entry whose mask meets U adds u's
//     sigma row on those lanes.  Path counts are integers, so the sum is exact
//     in any order.
//   backward, level L: each entry (v -> u) of a row v with lanes at level L
//     gathers u's level-(L+1) mask; a DAG entry adds its factor to dep[v] on
//     those lanes, and the 64-lane sum of the factors to the entry's edge
//     accumulator.  Every entry and every vertex has one owner per launch, so
//     no atomics and a fixed summation order.  A finished vertex stores
//     g = (1 + dep) / sigma instead of dep, so a factor is one product,
//     sigma[v] * g[u], rather than the reference's (sigma[v] / sigma[u]) *
//     (1 + dep[u]) with its division: the same value up to two roundings.
// The wavefront gathers one mask per lane for 64 entries at a time, then walks
// the entries that hit with all 64 lanes, four at a time so that their row
// reads overlap.  Pieces of long rows leave partial rows that a fix-up kernel
// sums in piece order.  Edge sums are kept per entry position and scattered
// to edge ids at the end.
#include <algorithm>
#include <string>
#include <vector>

#include "device_common.hpp"
#include "gt_internal.hpp"
#include "props.hpp"

namespace gt {

constexpr uint32_t kBcUnreached = 0xffffffffu;
// Entries and rows per row tile.  Rows with more than kBcPieceNnz entries
// are cut into pieces of that many (the hits of a row are walked by one
// wavefront, so no wavefront walks more than a piece of one row).
constexpr uint32_t kBcTileNnz = 128;
constexpr uint32_t kBcTileRows = 64;
constexpr uint32_t kBcPieceNnz = 128;
// Forward levels launched between two reads of the level counter.
constexpr uint32_t kBcLevelsPerCheck = 4;
// Entries that hit, walked together (the backward's edge sums reduce four
// at once).
constexpr int kBcGroup = 4;

struct BcArgs {
    const uint32_t *off, *idx;  // the orientation walked (rows list the neighbours pulled)
    const uint4 *tiles;
    uint32_t ntiles, npieces;
    const uint32_t *long_rows, *long_first;
    uint32_t nlong;
    uint32_t n;
    uint64_t act;           // lanes holding a source in this batch
    uint64_t *seen;         // [n]
    uint64_t *mask[2];      // [n] each: level masks, level L in mask[L & 1]
    uint32_t *lo, *hi;      // [n]: first and last level of the vertex over the lanes
    uint32_t *dist;         // [n][64]
    double *sigma;          // [n][64]: shortest-path counts
    double *gq;             // [n][64]: (1 + dependency) / sigma, once the level is done
    double *part;           // [npieces][64]
    uint64_t *pmask;        // [npieces]
    double *vacc;           // [n]
    double *eacc;           // [entries]: per entry position
    uint32_t *level;        // deepest level reached so far (device counter)
    uint32_t *status;       // this batch's slot: deepest level + 1, or 0 if its
                            // forward needed more levels than were launched
    // Work counters, one slot per wavefront of a launch (no shared atomics):
    // [wave][3] = entries, lane hits, entry hits.
    unsigned long long *work;
};

// Most wavefronts of a level launch (the grid is capped at the resident
// workgroups, at most kMaxSweepGrid).
constexpr uint32_t kBcMaxWaves = kMaxSweepGrid * kWavesPerBlock;

// Streams a call runs batches on (1, 2 or 4), and the most device memory
// the other streams' batch state may take together (bytes).  At R-MAT 16,
// all pivots: 460.6 ms per call on one stream, 321.8 on two, 342.7 on four.
constexpr int kBcLanes = 2;
static_assert(kBcLanes == 1 || kBcLanes == 2 || kBcLanes == 4, "lanes");
constexpr double kBcLaneBudget = 16.0 * (1ull << 30);

__device__ __forceinline__ uint64_t readlane64(uint64_t x, int k) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, k);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), k);
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint32_t readlane32(uint32_t x, int k) {
    return (uint32_t)__builtin_amdgcn_readlane((int)x, k);
}

__device__ __forceinline__ uint64_t shfl64(uint64_t x, int src) {
    const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)x, src, kWave);
    const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(x >> 32), src, kWave);
    return ((uint64_t)hi << 32) | lo;
}

// x from the lane DPP control CTRL names, both halves of the double.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double x) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(x), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(x), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// Sum over a row of 16 lanes: pairs (quad_perm 1,0,3,2), quads (quad_perm
// 2,3,0,1), then rotations of the row by 4 and 8.  Every lane ends with the
// row total (in an order that depends on its position).
__device__ __forceinline__ double row16_sum(double x) {
    x += dpp_f64<0xB1>(x);
    x += dpp_f64<0x4E>(x);
    x += dpp_f64<0x124>(x);
    x += dpp_f64<0x128>(x);
    return x;
}

// The 64-lane totals of four values at once.  Exchanges across 32 and 16
// lanes halve the values a lane carries (it keeps value 2*(lane>=32) +
// (lane&16 != 0)), so one row sum finishes all four: the total of f[i] is
// taken from lane 16 i.  Lane kk[i] (if kk[i] >= 0) receives it in `out`.
__device__ __forceinline__ void wave_sum4(const double *f, const int *kk, double &out, int lane) {
    const bool h32 = lane & 32, h16 = lane & 16;
    double s0 = h32 ? f[2] : f[0], s1 = h32 ? f[3] : f[1];
    const double g0 = h32 ? f[0] : f[2], g1 = h32 ? f[1] : f[3];
    s0 += __shfl_xor(g0, 32, kWave);
    s1 += __shfl_xor(g1, 32, kWave);
    double k = h16 ? s1 : s0;
    const double g = h16 ? s0 : s1;
    k += __shfl_xor(g, 16, kWave);
    k = row16_sum(k);
    int src = lane;
    bool take = false;
#pragma unroll
    for (int i = 0; i < 4; ++i)
        if (kk[i] == lane) {
            src = 16 * i;
            take = true;
        }
    const double t = __shfl(k, src, kWave);
    if (take) out = t;
}

__device__ __forceinline__ int pop_hit(uint64_t &hits) {
    const int k = __builtin_ctzll(hits);
    hits &= hits - 1;
    return k;
}

// Row slot (within the tile) of this lane's entry j in the chunk starting at
// `base`: the last tile row whose first entry is <= j.  `rs` is the slot of
// the row holding `base`; row starts are o_r in lane r.
__device__ __forceinline__ uint32_t entry_slot(uint32_t rs, uint32_t nr, uint32_t o_r,
                                               uint32_t base, uint32_t j) {
    uint32_t slot = rs;
    for (uint32_t r = rs + 1; r < nr; ++r) {
        const uint32_t o = readlane32(o_r, r);
        if (o >= base + kWave) break;
        if (j >= o) slot = r;
    }
    return slot;
}

// Adds this wavefront's counts to its own slot (the launch's work pointer is
// per phase; the same wavefront index owns the same slot in every launch).
__device__ __forceinline__ void add_work(const BcArgs &a, uint32_t wave, uint32_t entries,
                                         uint32_t lane_hits, uint32_t entry_hits, int lane) {
    const uint32_t h = (uint32_t)wave_sum((double)lane_hits);
    if (lane == 0 && (entries | h | entry_hits)) {
        unsigned long long *w = a.work + 3 * (size_t)wave;
        w[0] += entries;
        w[1] += h;
        w[2] += entry_hits;
    }
}

// Raises the level counter to L + 1 once per workgroup that found a vertex
// of level L + 1 (a read first: most workgroups find it raised already).
__device__ __forceinline__ void mark_level(const BcArgs &a, bool found, uint32_t L,
                                           uint32_t *s_found) {
    if (found && (threadIdx.x & (kWave - 1)) == 0) *s_found = 1;
    __syncthreads();
    if (threadIdx.x == 0 && *s_found &&
        __hip_atomic_load(a.level, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < L + 1)
        __hip_atomic_store(a.level, L + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ------------------------------------------------------------------ forward

// Entries [jb, je) of a single row w with unreached lanes U (a piece of a
// long row): lanes add the sigma rows of predecessors at level L.
__device__ __forceinline__ void fwd_piece(const BcArgs &a, const uint64_t *cur, uint32_t jb,
                                          uint32_t je, uint32_t w, uint64_t U, double &acc,
                                          uint64_t &H, uint32_t &lh, uint32_t &eh, int lane) {
    for (uint32_t base = jb; base < je; base += kWave) {
        const uint32_t j = base + lane;
        uint32_t u = w;
        uint64_t F = 0;
        if (j < je) {
            u = a.idx[j];
            if (u != w) F = cur[u] & U;
        }
        uint64_t hits = __ballot(F != 0);
        eh += __popcll(hits);
        while (hits) {
            double v[kBcGroup];
#pragma unroll
            for (int i = 0; i < kBcGroup; ++i) {
                v[i] = 0.0;
                if (hits) {
                    const int k = pop_hit(hits);
                    const uint32_t uk = readlane32(u, k);
                    const uint64_t Fk = readlane64(F, k);
                    H |= Fk;
                    if ((Fk >> lane) & 1) {
                        v[i] = a.sigma[(size_t)uk * kWave + lane];
                        ++lh;
                    }
                }
            }
#pragma unroll
            for (int i = 0; i < kBcGroup; ++i) acc += v[i];
        }
    }
}

// A finished row w: its level-(L+1) lanes H get their path counts.
__device__ __forceinline__ void fwd_store(const BcArgs &a, uint32_t w, uint64_t H, double acc,
                                          uint32_t L, int lane) {
    if ((H >> lane) & 1) {
        a.dist[(size_t)w * kWave + lane] = L + 1;
        a.sigma[(size_t)w * kWave + lane] = acc;
    }
}

// Row bookkeeping after a level: mask out, seen, level range (by the lanes
// with `ok`, one row each).
__device__ __forceinline__ void fwd_rows(const BcArgs &a, uint64_t *nxt, uint32_t w, bool ok,
                                         uint64_t seen, uint64_t H, uint32_t L) {
    if (!ok) return;
    nxt[w] = H;
    if (H) {
        a.seen[w] = seen | H;
        if (a.lo[w] == kBcUnreached) a.lo[w] = L + 1;
        a.hi[w] = L + 1;
    }
}

// Level L -> L + 1.  Does nothing once a level came out empty.
__global__ __launch_bounds__(kBlock) void bc_forward(BcArgs a, uint32_t L) {
    // Uniform over the launch: only this launch raises the counter, to L + 1.
    if (__hip_atomic_load(a.level, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < L) return;
    __shared__ uint32_t s_found;
    if (threadIdx.x == 0) s_found = 0;
    __syncthreads();
    const int lane = threadIdx.x & (kWave - 1);
    const uint32_t wave = (blockIdx.x * kBlock + threadIdx.x) / kWave;
    const uint32_t nwaves = gridDim.x * (kBlock / kWave);
    const uint64_t *cur = a.mask[L & 1];
    uint64_t *nxt = a.mask[(L + 1) & 1];
    bool found = false;
    uint32_t entries = 0, lh = 0, eh = 0;
    for (uint32_t t = wave; t < a.ntiles; t += nwaves) {
        const uint4 tl = a.tiles[t];
        if (tl.y == kSplit) {  // a piece of a long row
            const uint32_t w = tl.x;
            const uint64_t U = ~a.seen[w] & a.act;
            double acc = 0.0;
            uint64_t H = 0;
            if (U) {
                entries += tl.w - tl.z;
                fwd_piece(a, cur, tl.z, tl.w, w, U, acc, H, lh, eh, lane);
            }
            a.part[(size_t)t * kWave + lane] = acc;
            if (lane == 0) a.pmask[t] = H;
            continue;
        }
        const uint32_t rb = tl.x, nr = tl.y - tl.x;
        const bool rl = (uint32_t)lane < nr;
        const uint32_t row = rb + (rl ? lane : 0);
        const uint64_t seen_r = rl ? a.seen[row] : ~0ull;
        const uint64_t U_r = ~seen_r & a.act;
        uint64_t H_r = 0;
        if (__ballot(U_r != 0)) {
            const uint32_t o_r = rl ? a.off[row] : tl.w;
            entries += tl.w - tl.z;
            int cur_row = -1;
            double acc = 0.0;
            uint64_t Hc = 0;
            uint32_t rs = 0;
            for (uint32_t base = tl.z; base < tl.w; base += kWave) {
                const uint32_t j = base + lane;
                const uint32_t slot = entry_slot(rs, nr, o_r, base, j);
                rs = readlane32(slot, kWave - 1);
                const uint64_t Uj = shfl64(U_r, (int)slot);
                uint32_t u = 0;
                uint64_t F = 0;
                if (j < tl.w && Uj) {
                    u = a.idx[j];
                    if (u != rb + slot) F = cur[u] & Uj;
                }
                uint64_t hits = __ballot(F != 0);
                eh += __popcll(hits);
                while (hits) {
                    double v[kBcGroup];
                    uint64_t Fk[kBcGroup];
                    int rk[kBcGroup];
                    int cnt = 0;
#pragma unroll
                    for (int i = 0; i < kBcGroup; ++i) {
                        v[i] = 0.0;
                        Fk[i] = 0;
                        rk[i] = -1;
                        if (hits) {
                            ++cnt;
                            const int k = pop_hit(hits);
                            const uint32_t uk = readlane32(u, k);
                            Fk[i] = readlane64(F, k);
                            rk[i] = (int)readlane32(slot, k);
                            if ((Fk[i] >> lane) & 1) {
                                v[i] = a.sigma[(size_t)uk * kWave + lane];
                                ++lh;
                            }
                        }
                    }
#pragma unroll
                    for (int i = 0; i < kBcGroup; ++i) {
                        if (i >= cnt) break;
                        if (rk[i] != cur_row) {
                            if (cur_row >= 0) {
                                fwd_store(a, rb + cur_row, Hc, acc, L, lane);
                                if (lane == cur_row) H_r = Hc;
                            }
                            cur_row = rk[i];
                            acc = 0.0;
                            Hc = 0;
                        }
                        acc += v[i];
                        Hc |= Fk[i];
                    }
                }
            }
            if (cur_row >= 0) {
                fwd_store(a, rb + cur_row, Hc, acc, L, lane);
                if (lane == cur_row) H_r = Hc;
            }
        }
        fwd_rows(a, nxt, row, rl, seen_r, H_r, L);
        found |= __ballot(H_r != 0) != 0;
    }
    add_work(a, wave, entries, lh, eh, lane);
    mark_level(a, found, L, &s_found);
}

// Sum of the partial rows [p0, p1) for this thread's lane, in a fixed order:
// wavefront q of the workgroup sums pieces p0 + q, p0 + q + 4, ..., then the
// four wavefront sums are added in wavefront order.  Block-uniform call.
// Sixteen pieces are loaded at a time: a hub of ~200 pieces sets the fix-up
// kernel's critical path.
__device__ __forceinline__ double sum_pieces(const double *part, uint32_t p0, uint32_t p1,
                                             double (*red)[kWave]) {
    constexpr int B = 16;
    const int lane = threadIdx.x & (kWave - 1);
    const int q = threadIdx.x / kWave;
    constexpr uint32_t S = kWavesPerBlock;
    double acc = 0.0;
    uint32_t p = p0 + q;
    for (; p + (B - 1) * S < p1; p += B * S) {
        double v[B];
#pragma unroll
        for (int i = 0; i < B; ++i) v[i] = part[(size_t)(p + i * S) * kWave + lane];
#pragma unroll
        for (int i = 0; i < B; ++i) acc += v[i];
    }
    for (; p < p1; p += S) acc += part[(size_t)p * kWave + lane];
    red[q][lane] = acc;
    __syncthreads();
    const double t = (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
    __syncthreads();
    return t;
}

// Finishes the long rows of level L -> L + 1 from their pieces (one
// workgroup per row).
__global__ __launch_bounds__(kBlock) void bc_forward_fix(BcArgs a, uint32_t L) {
    if (__hip_atomic_load(a.level, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < L) return;
    __shared__ uint32_t s_found;
    __shared__ double red[kWavesPerBlock][kWave];
    __shared__ unsigned long long s_mask;
    if (threadIdx.x == 0) s_found = 0;
    const int lane = threadIdx.x & (kWave - 1);
    uint64_t *nxt = a.mask[(L + 1) & 1];
    bool found = false;
    for (uint32_t k = blockIdx.x; k < a.nlong; k += gridDim.x) {
        const uint32_t w = a.long_rows[k];
        const uint32_t p0 = a.long_first[k], p1 = a.long_first[k + 1];
        if (threadIdx.x == 0) s_mask = 0;
        __syncthreads();
        uint64_t h = 0;
        for (uint32_t p = p0 + threadIdx.x; p < p1; p += kBlock) h |= a.pmask[p];
        for (int o = kWave / 2; o > 0; o >>= 1) h |= shfl64(h, lane ^ o);
        if (lane == 0 && h) atomicOr(&s_mask, (unsigned long long)h);
        const double acc = sum_pieces(a.part, p0, p1, red);  // contains the barriers
        const uint64_t H = s_mask;
        if (threadIdx.x < kWave) {
            fwd_store(a, w, H, acc, L, lane);
            fwd_rows(a, nxt, w, lane == 0, a.seen[w], H, L);
        }
        found |= H != 0;
        __syncthreads();
    }
    mark_level(a, found, L, &s_found);
}

// ----------------------------------------------------------------- backward

// Lanes of row v at level L (every lane calls it for the same v).
__device__ __forceinline__ uint64_t level_lanes(const BcArgs &a, uint32_t v, uint32_t L,
                                                int lane) {
    const uint32_t lo = a.lo[v], hi = a.hi[v];
    if (lo == kBcUnreached || L < lo || L > hi) return 0;
    return __ballot(a.dist[(size_t)v * kWave + lane] == L) & a.seen[v];
}

// Up to kBcGroup entries that hit: their factors (lanes are sources), each
// entry's 64-lane sum into the register of the lane that holds the entry, and
// their row slots.
struct BwdGroup {
    double f[kBcGroup];
    int rk[kBcGroup];
    int cnt;
};

__device__ __forceinline__ void bwd_group(const BcArgs &a, uint64_t &hits, uint32_t u,
                                          uint64_t S, uint32_t slot, uint32_t rb, double sv_row,
                                          bool one_row, double &emine, BwdGroup &G,
                                          uint32_t &lh, int lane) {
#pragma clang fp contract(off)
    int kk[kBcGroup];
    G.cnt = 0;
#pragma unroll
    for (int i = 0; i < kBcGroup; ++i) {
        kk[i] = -1;
        G.rk[i] = -1;
        G.f[i] = 0.0;
        if (hits) {
            const int k = pop_hit(hits);
            kk[i] = k;
            ++G.cnt;
            const uint32_t uk = readlane32(u, k);
            const uint64_t Sk = readlane64(S, k);
            G.rk[i] = (int)readlane32(slot, k);
            if ((Sk >> lane) & 1) {
                const double sv =
                    one_row ? sv_row : a.sigma[(size_t)(rb + G.rk[i]) * kWave + lane];
                G.f[i] = sv * a.gq[(size_t)uk * kWave + lane];
                ++lh;
            }
        }
    }
    wave_sum4(G.f, kk, emine, lane);
}

// A finished row v with lanes m at level L and dependencies dacc: the
// row's (1 + dep) / sigma for its predecessors, and its vertex sum into the
// register of lane `slot` (level 0 is the source itself and adds nothing).
__device__ __forceinline__ void bwd_store(const BcArgs &a, uint32_t v, uint64_t m, double dacc,
                                          uint32_t L, int lane, int slot, double &vrow) {
#pragma clang fp contract(off)
    const bool on = (m >> lane) & 1;
    if (on) {
        const size_t q = (size_t)v * kWave + lane;
        a.gq[q] = (1.0 + dacc) / a.sigma[q];
    }
    if (L > 0) {
        const double t = wave_sum(on ? dacc : 0.0);
        if (lane == slot) vrow = t;
    }
}

// Level L.  The deepest level, top, has no successors; it is read from the
// device counter.  `launched` is the number of forward levels launched: if
// the counter reached it, the forward may not be complete and the whole
// backward does nothing (the batch is run again).
__global__ __launch_bounds__(kBlock) void bc_backward(BcArgs a, uint32_t L, uint32_t launched) {
    const uint32_t top = __hip_atomic_load(a.level, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (L == 0 && blockIdx.x == 0 && threadIdx.x == 0) *a.status = top < launched ? top + 1 : 0;
    if (top >= launched || L > top) return;
    const int lane = threadIdx.x & (kWave - 1);
    const uint32_t wave = (blockIdx.x * kBlock + threadIdx.x) / kWave;
    const uint32_t nwaves = gridDim.x * (kBlock / kWave);
    const uint64_t *succ = a.mask[(L + 1) & 1];
    uint64_t *curm = a.mask[L & 1];
    uint32_t entries = 0, lh = 0, eh = 0;
    for (uint32_t t = wave; t < a.ntiles; t += nwaves) {
        const uint4 tl = a.tiles[t];
        if (tl.y == kSplit) {  // a piece of a long row
            const uint32_t v = tl.x;
            const uint64_t m = level_lanes(a, v, L, lane);
            double dacc = 0.0;
            if (m && L < top) {
                entries += tl.w - tl.z;
                const double sv = ((m >> lane) & 1) ? a.sigma[(size_t)v * kWave + lane] : 1.0;
                for (uint32_t base = tl.z; base < tl.w; base += kWave) {
                    const uint32_t j = base + lane;
                    uint32_t u = v;
                    uint64_t S = 0;
                    if (j < tl.w) {
                        u = a.idx[j];
                        if (u != v) S = succ[u] & m;
                    }
                    const uint64_t hit0 = __ballot(S != 0);
                    uint64_t hits = hit0;
                    eh += __popcll(hits);
                    double emine = 0.0;
                    while (hits) {
                        BwdGroup G;
                        bwd_group(a, hits, u, S, 0u, v, sv, true, emine, G, lh, lane);
#pragma unroll
                        for (int i = 0; i < kBcGroup; ++i) dacc += G.f[i];
                    }
                    if ((hit0 >> lane) & 1) a.eacc[j] += emine;
                }
            }
            a.part[(size_t)t * kWave + lane] = dacc;
            continue;
        }
        const uint32_t rb = tl.x, nr = tl.y - tl.x;
        const bool rl = (uint32_t)lane < nr;
        const uint32_t row = rb + (rl ? lane : 0);
        // Lanes of each row at level L (rows in their level range only).
        const uint32_t lo_r = rl ? a.lo[row] : kBcUnreached;
        const uint32_t hi_r = rl ? a.hi[row] : 0;
        const bool inr = lo_r != kBcUnreached && L >= lo_r && L <= hi_r;
        uint64_t m_r = 0;
        {
            const uint64_t seen_r = inr ? a.seen[row] : 0;
            uint64_t rows = __ballot(inr);
            while (rows) {
                uint32_t d[kBcGroup];
                int rr[kBcGroup];
#pragma unroll
                for (int i = 0; i < kBcGroup; ++i) {
                    rr[i] = -1;
                    d[i] = kBcUnreached;
                    if (rows) {
                        rr[i] = pop_hit(rows);
                        d[i] = a.dist[(size_t)(rb + rr[i]) * kWave + lane];
                    }
                }
#pragma unroll
                for (int i = 0; i < kBcGroup; ++i) {
                    if (rr[i] >= 0) {
                        const uint64_t m = __ballot(d[i] == L) & readlane64(seen_r, rr[i]);
                        if (lane == rr[i]) m_r = m;
                    }
                }
            }
        }
        if (rl) curm[row] = m_r;
        const uint64_t live = __ballot(m_r != 0);
        if (!live) continue;
        uint64_t done = 0;  // rows whose dependencies were stored
        double vrow = 0.0;
        if (L < top) {
            const uint32_t o_r = rl ? a.off[row] : tl.w;
            entries += tl.w - tl.z;
            int cur_row = -1;
            double dacc = 0.0;
            uint32_t rs = 0;
            for (uint32_t base = tl.z; base < tl.w; base += kWave) {
                const uint32_t j = base + lane;
                const uint32_t slot = entry_slot(rs, nr, o_r, base, j);
                rs = readlane32(slot, kWave - 1);
                const uint64_t mj = shfl64(m_r, (int)slot);
                uint32_t u = 0;
                uint64_t S = 0;
                if (j < tl.w && mj) {
                    u = a.idx[j];
                    if (u != rb + slot) S = succ[u] & mj;
                }
                const uint64_t hit0 = __ballot(S != 0);
                uint64_t hits = hit0;
                eh += __popcll(hits);
                double emine = 0.0;
                while (hits) {
                    BwdGroup G;
                    bwd_group(a, hits, u, S, slot, rb, 0.0, false, emine, G, lh, lane);
#pragma unroll
                    for (int i = 0; i < kBcGroup; ++i) {
                        if (i >= G.cnt) break;
                        if (G.rk[i] != cur_row) {
                            if (cur_row >= 0) {
                                bwd_store(a, rb + cur_row, readlane64(m_r, cur_row), dacc, L,
                                          lane, cur_row, vrow);
                                done |= 1ull << cur_row;
                            }
                            cur_row = G.rk[i];
                            dacc = 0.0;
                        }
                        dacc += G.f[i];
                    }
                }
                if ((hit0 >> lane) & 1) a.eacc[j] += emine;
            }
            if (cur_row >= 0) {
                bwd_store(a, rb + cur_row, readlane64(m_r, cur_row), dacc, L, lane, cur_row,
                          vrow);
                done |= 1ull << cur_row;
            }
        }
        // Rows at level L without a successor on the DAG: no dependency.
        uint64_t leaves = live & ~done;
        while (leaves) {
            const int r = pop_hit(leaves);
            if ((readlane64(m_r, r) >> lane) & 1) {
                const size_t q = (size_t)(rb + r) * kWave + lane;
                a.gq[q] = 1.0 / a.sigma[q];
            }
        }
        if (rl && ((done >> lane) & 1) && L > 0) a.vacc[row] += vrow;
    }
    add_work(a, wave, entries, lh, eh, lane);
}

// Finishes the long rows of level L from their pieces (one workgroup per row).
__global__ __launch_bounds__(kBlock) void bc_backward_fix(BcArgs a, uint32_t L, uint32_t launched) {
    const uint32_t top = __hip_atomic_load(a.level, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (top >= launched || L > top) return;
    __shared__ double red[kWavesPerBlock][kWave];
    const int lane = threadIdx.x & (kWave - 1);
    uint64_t *curm = a.mask[L & 1];
    for (uint32_t k = blockIdx.x; k < a.nlong; k += gridDim.x) {
        const uint32_t v = a.long_rows[k];
        const uint64_t m = level_lanes(a, v, L, lane);  // the same in every wavefront
        if (threadIdx.x == 0) curm[v] = m;
        if (!m) continue;
        const double dacc = sum_pieces(a.part, a.long_first[k], a.long_first[k + 1], red);
        if (threadIdx.x < kWave) {
            double vrow = 0.0;
            bwd_store(a, v, m, dacc, L, lane, 0, vrow);
            if (lane == 0 && L > 0) a.vacc[v] += vrow;
        }
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

// ---------------------------------------------------- merged orientation

// Whole blocks: row r = out-half of r, then in-half of r (undirected.rs:110-112).
__global__ __launch_bounds__(kBlock) void bc_merge_off(const uint32_t *oa, const uint32_t *ob,
                                                       uint32_t n, uint32_t *off,
                                                       uint32_t *first_len) {
    for (uint32_t r = blockIdx.x * kBlock + threadIdx.x; r <= n; r += gridDim.x * kBlock) {
        off[r] = oa[r] + ob[r];
        if (r < n) first_len[r] = oa[r + 1] - oa[r];
    }
}

// One wavefront per row, lanes over its entries.
__global__ __launch_bounds__(kBlock) void bc_merge_rows(const uint32_t *oa, const uint32_t *ia,
                                                        const uint32_t *ea, const uint32_t *ob,
                                                        const uint32_t *ib, const uint32_t *eb,
                                                        uint32_t n, const uint32_t *off,
                                                        uint32_t *idx, uint32_t *eid) {
    const int lane = threadIdx.x & (kWave - 1);
    const uint32_t wave = (blockIdx.x * kBlock + threadIdx.x) / kWave;
    const uint32_t nwaves = gridDim.x * (kBlock / kWave);
    for (uint32_t r = wave; r < n; r += nwaves) {
        uint32_t w = off[r];
        for (uint32_t j = oa[r] + lane; j < oa[r + 1]; j += kWave) {
            idx[w + (j - oa[r])] = ia[j];
            eid[w + (j - oa[r])] = ea[j];
        }
        w += oa[r + 1] - oa[r];
        for (uint32_t j = ob[r] + lane; j < ob[r + 1]; j += kWave) {
            idx[w + (j - ob[r])] = ib[j];
            eid[w + (j - ob[r])] = eb[j];
        }
    }
}

// ------------------------------------------------------------------ output

__global__ __launch_bounds__(kBlock) void bc_emit_vertices(const double *vacc, uint32_t n,
                                                           bool halve, double *out) {
    for (uint32_t v = blockIdx.x * kBlock + threadIdx.x; v < n; v += gridDim.x * kBlock)
        out[v] = halve ? vacc[v] / 2.0 : vacc[v];
}

// acc += other, entry by entry (the second stream's sums into the first's).
__global__ __launch_bounds__(kBlock) void bc_add(double *acc, const double *other, uint64_t len) {
    for (uint64_t e = blockIdx.x * (uint64_t)kBlock + threadIdx.x; e < len;
         e += (uint64_t)gridDim.x * kBlock)
        acc[e] += other[e];
}

__global__ __launch_bounds__(kBlock) void bc_fill(double *out, uint64_t len) {
    for (uint64_t e = blockIdx.x * (uint64_t)kBlock + threadIdx.x; e < len;
         e += (uint64_t)gridDim.x * kBlock)
        out[e] = 0.0;
}

// Edge sums from entry positions, one wavefront per row.  Without first_len
// (a directed view) each edge id appears once.  Over whole blocks, pass 0
// writes the out-half entries and pass 1 adds the in-half entries and halves
// (each edge has one entry in each half).
__global__ __launch_bounds__(kBlock) void bc_emit_edges(const uint32_t *off, const uint32_t *eid,
                                                        const uint32_t *first_len,
                                                        const double *acc, uint32_t n, int pass,
                                                        double *out) {
    const int lane = threadIdx.x & (kWave - 1);
    const uint32_t wave = (blockIdx.x * kBlock + threadIdx.x) / kWave;
    const uint32_t nwaves = gridDim.x * (kBlock / kWave);
    for (uint32_t r = wave; r < n; r += nwaves) {
        const uint32_t b = off[r], e = off[r + 1];
        if (!first_len) {
            for (uint32_t j = b + lane; j < e; j += kWave) out[eid[j]] = acc[j];
        } else if (pass == 0) {
            for (uint32_t j = b + lane; j < b + first_len[r]; j += kWave) out[eid[j]] = acc[j];
        } else {
            for (uint32_t j = b + first_len[r] + lane; j < e; j += kWave)
                out[eid[j]] = (out[eid[j]] + acc[j]) / 2.0;
        }
    }
}

namespace {

unsigned lgrid(uint64_t n) {
    uint64_t b = (n + kBlock - 1) / kBlock;
    if (b > 4096) b = 4096;
    return b ? (unsigned)b : 1u;
}

unsigned wgrid(uint64_t rows) {
    uint64_t b = (rows + kWavesPerBlock - 1) / kWavesPerBlock;
    if (b > 4096) b = 4096;
    return b ? (unsigned)b : 1u;
}

// Merge the out-half (a) and in-half (b) orientations into whole blocks.
gt_status merge_blocks(gt_graph *g, const Orient &a, const Orient &b, Orient &o) {
    const uint32_t n = g->n;
    hipStream_t s = g->stream;
    const uint64_t m = (uint64_t)a.h_off[n] + b.h_off[n];
    GT_TRY(o.off.reserve(sizeof(uint32_t) * ((size_t)n + 1)));
    GT_TRY(o.first_len.reserve(sizeof(uint32_t) * std::max<size_t>(n, 1)));
    GT_TRY(o.idx.reserve(sizeof(uint32_t) * std::max<uint64_t>(m, 1)));
    GT_TRY(o.eid.reserve(sizeof(uint32_t) * std::max<uint64_t>(m, 1)));
    hipLaunchKernelGGL(bc_merge_off, dim3(lgrid((uint64_t)n + 1)), dim3(kBlock), 0, s,
                       a.off.as<uint32_t>(), b.off.as<uint32_t>(), n, o.off.as<uint32_t>(),
                       o.first_len.as<uint32_t>());
    if (n)
        hipLaunchKernelGGL(bc_merge_rows, dim3(wgrid(n)), dim3(kBlock), 0, s,
                           a.off.as<uint32_t>(), a.idx.as<uint32_t>(), a.eid.as<uint32_t>(),
                           b.off.as<uint32_t>(), b.idx.as<uint32_t>(), b.eid.as<uint32_t>(), n,
                           o.off.as<uint32_t>(), o.idx.as<uint32_t>(), o.eid.as<uint32_t>());
    GT_HIP(hipGetLastError());
    o.h_off.resize((size_t)n + 1);
    for (size_t r = 0; r <= n; ++r) o.h_off[r] = a.h_off[r] + b.h_off[r];
    o.scheduled = false;
    o.mm.ready = false;
    o.heavy_ready = false;
    o.bcs.ready = false;
    return GT_OK;
}

// Row tiles of an orientation (from its host offsets), built once per tile
// size.
gt_status ensure_bc_schedule(gt_graph *g, Orient &o, uint32_t tile_nnz) {
    BcSchedule &S = o.bcs;
    if (S.ready && S.tile_nnz == tile_nnz) return GT_OK;
    const uint32_t n = g->n;
    const std::vector<uint32_t> &off = o.h_off;
    std::vector<uint32_t> tiles, lrows, lfirst;
    lfirst.push_back(0);
    auto is_long = [&](uint32_t r) { return off[r + 1] - off[r] > kBcPieceNnz; };
    for (uint32_t r = 0; r < n; ++r) {
        if (!is_long(r)) continue;
        lrows.push_back(r);
        for (uint32_t b = off[r]; b < off[r + 1]; b += kBcPieceNnz)
            tiles.insert(tiles.end(), {r, kSplit, b, std::min(off[r + 1], b + kBcPieceNnz)});
        lfirst.push_back((uint32_t)(tiles.size() / 4));
    }
    const uint32_t npieces = (uint32_t)(tiles.size() / 4);
    for (uint32_t r = 0; r < n;) {
        if (is_long(r)) {
            ++r;
            continue;
        }
        const uint32_t rb = r, eb = off[r];
        while (r < n && !is_long(r) && r - rb < kBcTileRows && off[r + 1] - eb <= tile_nnz) ++r;
        tiles.insert(tiles.end(), {rb, r, eb, off[r]});
    }
    hipStream_t s = g->stream;
    GT_TRY(S.tiles.reserve(std::max<size_t>(tiles.size(), 4) * sizeof(uint32_t)));
    GT_TRY(S.long_rows.reserve(std::max<size_t>(lrows.size(), 1) * sizeof(uint32_t)));
    GT_TRY(S.long_first.reserve(lfirst.size() * sizeof(uint32_t)));
    if (!tiles.empty())
        GT_HIP(hipMemcpyAsync(S.tiles.ptr, tiles.data(), tiles.size() * sizeof(uint32_t),
                              hipMemcpyHostToDevice, s));
    if (!lrows.empty())
        GT_HIP(hipMemcpyAsync(S.long_rows.ptr, lrows.data(), lrows.size() * sizeof(uint32_t),
                              hipMemcpyHostToDevice, s));
    GT_HIP(hipMemcpyAsync(S.long_first.ptr, lfirst.data(), lfirst.size() * sizeof(uint32_t),
                          hipMemcpyHostToDevice, s));
    GT_HIP(hipStreamSynchronize(s));
    S.ntiles = (uint32_t)(tiles.size() / 4);
    S.npieces = npieces;
    S.nlong = (uint32_t)lrows.size();
    S.tile_nnz = tile_nnz;
    S.ready = true;
    return GT_OK;
}

struct BcView {
    Orient *fwd = nullptr;  // rows list the predecessors (in_edges of the view)
    Orient *bwd = nullptr;  // rows list the successors (out_edges of the view)
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
        bv->fwd = &g->in;
        bv->bwd = &g->out;
        return GT_OK;
    case GT_VIEW_REVERSED:
        bv->fwd = &g->out;
        bv->bwd = &g->in;
        return GT_OK;
    case GT_VIEW_DIRECTED_FILTERED:
    case GT_VIEW_REVERSED_FILTERED:
        GT_TRY(filtered_view(g, view, kind == GT_VIEW_REVERSED_FILTERED, &pull, &push, &act,
                             &nact));
        bv->fwd = pull;
        bv->bwd = push;
        bv->host_vmask = view->vmask;
        return GT_OK;
    case GT_VIEW_UNDIRECTED:
        if (!g->und_ready) {
            GT_TRY(merge_blocks(g, g->out, g->in, g->und));
            g->und_ready = true;
        }
        bv->fwd = bv->bwd = &g->und;
        bv->undirected = true;
        return GT_OK;
    case GT_VIEW_UNDIRECTED_FILTERED:
        GT_TRY(filtered_view(g, view, false, &pull, &push, &act, &nact));
        GT_TRY(merge_blocks(g, *push, *pull, g->ws.fund));
        bv->fwd = bv->bwd = &g->ws.fund;
        bv->undirected = true;
        bv->host_vmask = view->vmask;
        return GT_OK;
    default:
        return fail(GT_ERR_INVALID_ARG, "unknown view kind");
    }
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

void fill_sched(BcArgs &a, const Orient &o) {
    a.off = o.off.as<uint32_t>();
    a.idx = o.idx.as<uint32_t>();
    a.tiles = o.bcs.tiles.as<uint4>();
    a.ntiles = o.bcs.ntiles;
    a.npieces = o.bcs.npieces;
    a.long_rows = o.bcs.long_rows.as<uint32_t>();
    a.long_first = o.bcs.long_first.as<uint32_t>();
    a.nlong = o.bcs.nlong;
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
    const std::vector<uint32_t> &bo = bv.bwd->h_off;
    std::vector<uint32_t> src;
    if (pivots) {
        for (uint64_t i = 0; i < npivots; ++i) {
            const uint32_t v = pivots[i];
            if (v >= n) return fail(GT_ERR_NO_SUCH_VERTEX, "betweenness: pivot out of range");
            if (bv.host_vmask && !bv.host_vmask[v])
                return fail(GT_ERR_INVALID_ARG, "betweenness: pivot is not a vertex of the view");
            if (bo[v + 1] > bo[v]) src.push_back(v);
        }
    } else {
        for (uint32_t v = 0; v < n; ++v)
            if ((!bv.host_vmask || bv.host_vmask[v]) && bo[v + 1] > bo[v]) src.push_back(v);
    }

    // Batch sources that are close in the graph: their BFS levels overlap, so
    // one DAG entry serves more lanes at once.  Order: breadth-first over the
    // view's out_edges from the vertex with the most, restarting at the next
    // unvisited source in index order (sums are independent of the order up
    // to rounding).
    if (src.size() > kWave) {
        std::vector<uint32_t> hidx(bo[n]);
        if (bo[n])
            GT_HIP(hipMemcpyAsync(hidx.data(), bv.bwd->idx.ptr, bo[n] * sizeof(uint32_t),
                                  hipMemcpyDeviceToHost, s));
        GT_HIP(hipStreamSynchronize(s));
        std::vector<uint8_t> want(n, 0), seen(n, 0);
        for (uint32_t v : src) want[v] += want[v] < 255;
        uint32_t start = src[0];
        for (uint32_t v : src)
            if (bo[v + 1] - bo[v] > bo[start + 1] - bo[start]) start = v;
        std::vector<uint32_t> order, queue;
        order.reserve(src.size());
        queue.reserve(n);
        auto bfs = [&](uint32_t r) {
            size_t head = queue.size();
            queue.push_back(r);
            seen[r] = 1;
            while (head < queue.size()) {
                const uint32_t u = queue[head++];
                for (uint8_t c = want[u]; c > 0; --c) order.push_back(u);
                for (uint32_t j = bo[u]; j < bo[u + 1]; ++j)
                    if (!seen[hidx[j]]) {
                        seen[hidx[j]] = 1;
                        queue.push_back(hidx[j]);
                    }
            }
        };
        bfs(start);
        for (uint32_t v : src)
            if (!seen[v]) bfs(v);
        // Pivots listed more than 255 times keep their extra copies at the end.
        if (order.size() < src.size()) {
            std::vector<uint32_t> cnt(n, 0);
            for (uint32_t v : src) ++cnt[v];
            for (uint32_t v = 0; v < n; ++v)
                for (uint32_t c = 255; c < cnt[v]; ++c) order.push_back(v);
        }
        src.swap(order);
    }

    const uint32_t tile_nnz = kBcTileNnz;
    GT_TRY(ensure_bc_schedule(g, *bv.fwd, tile_nnz));
    if (bv.bwd != bv.fwd) GT_TRY(ensure_bc_schedule(g, *bv.bwd, tile_nnz));

    Workspace &ws = g->ws;
    const size_t rows = (size_t)std::max<uint32_t>(n, 1);
    const uint64_t bentries = bo[n];
    const size_t np = std::max<size_t>(std::max(bv.fwd->bcs.npieces, bv.bwd->bcs.npieces), 1);
    // A batch's state (rows, masks, pieces, work slots) and the accumulators
    // in B[0..8]; the sources and per-batch status in ws.btw[9].
    auto state = [&](DevBuf *B, BcArgs &st, unsigned long long **slots) -> gt_status {
        GT_TRY(B[0].reserve(rows * kWave * sizeof(double)));       // sigma
        GT_TRY(B[1].reserve(rows * kWave * sizeof(double)));       // gq
        GT_TRY(B[2].reserve(rows * kWave * sizeof(uint32_t)));     // dist
        GT_TRY(B[3].reserve(rows * 3 * sizeof(uint64_t) + 64));    // seen, mask[2], level
        GT_TRY(B[8].reserve(2 * 3 * kBcMaxWaves * sizeof(unsigned long long)));  // work slots
        GT_TRY(B[4].reserve(rows * 2 * sizeof(uint32_t)));         // lo, hi
        GT_TRY(B[5].reserve(rows * sizeof(double)));               // vacc
        GT_TRY(B[6].reserve(std::max<uint64_t>(bentries, 1) * sizeof(double)));  // eacc
        GT_TRY(B[7].reserve(np * kWave * sizeof(double) + np * sizeof(uint64_t)));
        st = BcArgs{};
        st.n = n;
        st.sigma = B[0].as<double>();
        st.gq = B[1].as<double>();
        st.dist = B[2].as<uint32_t>();
        st.seen = B[3].as<uint64_t>();
        st.mask[0] = st.seen + rows;
        st.mask[1] = st.mask[0] + rows;
        st.level = reinterpret_cast<uint32_t *>(st.mask[1] + rows);
        *slots = B[8].as<unsigned long long>();
        st.lo = B[4].as<uint32_t>();
        st.hi = st.lo + rows;
        st.vacc = B[5].as<double>();
        st.eacc = B[6].as<double>();
        st.part = B[7].as<double>();
        st.pmask = reinterpret_cast<uint64_t *>(st.part + np * kWave);
        return GT_OK;
    };
    // Lane 0 runs on the call's stream; lanes 1 .. kBcLanes - 1 on streams of
    // their own, each with its own state and accumulators.  Batch b runs on
    // lane b mod kBcLanes, so several batches' level launches overlap; the
    // other lanes' sums are added to lane 0's at the end, in lane order.
    // Calls that time their level groups with events, or have too few
    // batches, use lane 0 only.
    struct Lane {
        hipStream_t st;
        BcArgs fa, ba;
        unsigned long long *work;
    };
    const size_t nbatch = (src.size() + kWave - 1) / kWave;
    const bool prof = g->profiling;
    const double lane_bytes = (double)rows * (kWave * 20.0 + 48.0) + 8.0 * (double)bentries +
                              (double)np * (kWave * 8.0 + 8.0);
    int nlanes = 1;
    if (!prof)
        while (nlanes < kBcLanes && nbatch >= 2 * (size_t)nlanes + 1 &&
               lane_bytes * nlanes <= kBcLaneBudget)
            nlanes *= 2;
    Lane lanes[4];
    BcArgs base{};
    unsigned long long *work_slots = nullptr;
    GT_TRY(state(ws.btw, base, &work_slots));
    GT_TRY(ws.btw[9].reserve((src.size() + src.size() / kWave + 2) * sizeof(uint32_t)));
    uint32_t *dsrc = ws.btw[9].as<uint32_t>();
    if (!src.empty())
        GT_HIP(hipMemcpyAsync(dsrc, src.data(), src.size() * sizeof(uint32_t),
                              hipMemcpyHostToDevice, s));
    for (int k = 0; k < nlanes; ++k) {
        Lane &ln = lanes[k];
        BcArgs st = base;
        ln.work = work_slots;
        ln.st = s;
        if (k > 0) {
            hipStream_t &side = g->side_stream[k - 1];
            if (!side) GT_HIP(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
            ln.st = side;
            GT_TRY(state(ws.btwx[k - 1], st, &ln.work));
        }
        ln.fa = st;
        ln.ba = st;
        fill_sched(ln.fa, *bv.fwd);
        fill_sched(ln.ba, *bv.bwd);
        ln.fa.work = ln.work;
        ln.ba.work = ln.work + 3 * (size_t)kBcMaxWaves;
        GT_HIP(hipMemsetAsync(st.vacc, 0, rows * sizeof(double), s));
        GT_HIP(hipMemsetAsync(st.eacc, 0, std::max<uint64_t>(bentries, 1) * sizeof(double), s));
        GT_HIP(hipMemsetAsync(ln.work, 0, 2 * 3 * kBcMaxWaves * sizeof(unsigned long long), s));
    }

    const int fcap = occupancy_cap((const void *)bc_forward, 8);
    const int bcap = occupancy_cap((const void *)bc_backward, 8);
    const BcArgs &fa0 = lanes[0].fa, &ba0 = lanes[0].ba;
    const unsigned fgrid = (unsigned)std::max<uint64_t>(
        1, std::min<uint64_t>(fcap, ((uint64_t)fa0.ntiles + kWavesPerBlock - 1) / kWavesPerBlock));
    const unsigned bgrid = (unsigned)std::max<uint64_t>(
        1, std::min<uint64_t>(bcap, ((uint64_t)ba0.ntiles + kWavesPerBlock - 1) / kWavesPerBlock));
    const unsigned fixf = std::max<uint32_t>(1, std::min<uint32_t>(fa0.nlong, 4096));
    const unsigned fixb = std::max<uint32_t>(1, std::min<uint32_t>(ba0.nlong, 4096));

    hipEvent_t ev[4] = {};
    if (prof)
        for (auto &e : ev) GT_HIP(hipEventCreate(&e));
    uint64_t launches_f = 0, launches_b = 0;
    double fms = 0.0, bms = 0.0;
    uint32_t *dstatus = dsrc + src.size();
    if (nbatch) GT_HIP(hipMemsetAsync(dstatus, 0, nbatch * sizeof(uint32_t), s));
    std::vector<uint32_t> status(nbatch, 0);
    // The other lanes start after the set-up above; the call's stream waits
    // for them before the outputs.
    // On every way out of the call the other lanes' streams are drained (an
    // error part-way must not leave them writing the workspace the next call
    // reuses) and the events released.
    struct LaneScope {
        hipStream_t side[3] = {};
        hipEvent_t fork = nullptr, join[3] = {};
        ~LaneScope() {
            for (hipStream_t st : side)
                if (st) (void)hipStreamSynchronize(st);
            if (fork) (void)hipEventDestroy(fork);
            for (hipEvent_t e : join)
                if (e) (void)hipEventDestroy(e);
        }
    } scope;
    hipEvent_t &fork = scope.fork;
    hipEvent_t *join = scope.join;
    if (nlanes > 1) {
        GT_HIP(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
        GT_HIP(hipEventRecord(fork, s));
        for (int k = 1; k < nlanes; ++k) {
            GT_HIP(hipEventCreateWithFlags(&join[k - 1], hipEventDisableTiming));
            GT_HIP(hipStreamWaitEvent(lanes[k].st, fork, 0));
            scope.side[k - 1] = lanes[k].st;
        }
    }

    auto setup = [&](Lane &ln, size_t b) -> gt_status {
        const size_t first = b * kWave;
        const uint32_t cnt = (uint32_t)std::min<size_t>(kWave, src.size() - first);
        ln.fa.act = ln.ba.act = cnt == kWave ? ~0ull : ((1ull << cnt) - 1);
        ln.ba.status = dstatus + b;
        hipLaunchKernelGGL(bc_reset, dim3(lgrid(n)), dim3(kBlock), 0, ln.st, ln.fa);
        hipLaunchKernelGGL(bc_seed, dim3(1), dim3(kWave), 0, ln.st, ln.fa, dsrc + first, cnt);
        GT_HIP(hipGetLastError());
        return GT_OK;
    };
    auto forward = [&](Lane &ln, uint32_t L0, uint32_t count) -> gt_status {
        if (prof) GT_HIP(hipEventRecord(ev[0], s));
        for (uint32_t L = L0; L < L0 + count; ++L) {
            hipLaunchKernelGGL(bc_forward, dim3(fgrid), dim3(kBlock), 0, ln.st, ln.fa, L);
            if (ln.fa.nlong)
                hipLaunchKernelGGL(bc_forward_fix, dim3(fixf), dim3(kBlock), 0, ln.st, ln.fa, L);
            ++launches_f;
        }
        GT_HIP(hipGetLastError());
        if (prof) {
            GT_HIP(hipEventRecord(ev[1], s));
            GT_HIP(hipEventSynchronize(ev[1]));
            float ms = 0.f;
            GT_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
            fms += ms;
        }
        return GT_OK;
    };
    auto backward = [&](Lane &ln, uint32_t from, uint32_t launched) -> gt_status {
        if (prof) GT_HIP(hipEventRecord(ev[2], s));
        for (uint32_t l = from + 1; l-- > 0;) {
            hipLaunchKernelGGL(bc_backward, dim3(bgrid), dim3(kBlock), 0, ln.st, ln.ba, l,
                               launched);
            if (ln.ba.nlong)
                hipLaunchKernelGGL(bc_backward_fix, dim3(fixb), dim3(kBlock), 0, ln.st, ln.ba, l,
                                   launched);
            ++launches_b;
        }
        GT_HIP(hipGetLastError());
        if (prof) {
            GT_HIP(hipEventRecord(ev[3], s));
            GT_HIP(hipEventSynchronize(ev[3]));
            float ms = 0.f;
            GT_HIP(hipEventElapsedTime(&ms, ev[2], ev[3]));
            bms += ms;
        }
        return GT_OK;
    };
    // A batch with the host in the loop: forward levels in groups until one
    // comes out empty, then exactly the backward levels needed.
    auto run_sync = [&](size_t b, uint32_t *probe) -> gt_status {
        Lane &ln = lanes[0];
        GT_TRY(setup(ln, b));
        uint32_t L = 0, top = 0;
        for (;;) {
            GT_TRY(forward(ln, L, *probe));
            L += *probe;
            GT_HIP(hipMemcpyAsync(&g->h_state->counter, ln.fa.level, sizeof(uint32_t),
                                  hipMemcpyDeviceToHost, s));
            GT_HIP(hipStreamSynchronize(s));
            top = g->h_state->counter;
            if (top < L) break;  // a level came out empty
            *probe *= 2;
        }
        GT_TRY(backward(ln, top, 0xffffffffu));
        status[b] = top + 1;
        *probe = std::max<uint32_t>(kBcLevelsPerCheck, top + 2);
        return GT_OK;
    };
    // The other batches run without the host: `depth` forward levels (the
    // launches past the last level return at once), then depth + 1 backward
    // levels that read the depth from the device.  A batch whose BFS went
    // deeper leaves status 0 and is run again with the host in the loop.
    uint32_t probe = kBcLevelsPerCheck;
    uint64_t rerun = 0;
    for (size_t b = 0; b < nbatch; ++b) {
        if (b == 0) {
            GT_TRY(run_sync(b, &probe));
            continue;
        }
        Lane &ln = lanes[b % (size_t)nlanes];
        GT_TRY(setup(ln, b));
        GT_TRY(forward(ln, 0, probe));
        GT_TRY(backward(ln, probe, probe));
    }
    for (int k = 1; k < nlanes; ++k) {
        GT_HIP(hipEventRecord(join[k - 1], lanes[k].st));
        GT_HIP(hipStreamWaitEvent(s, join[k - 1], 0));
    }
    if (nbatch > 1) {
        std::vector<uint32_t> dev_status(nbatch);
        GT_HIP(hipMemcpyAsync(dev_status.data(), dstatus, nbatch * sizeof(uint32_t),
                              hipMemcpyDeviceToHost, s));
        GT_HIP(hipStreamSynchronize(s));
        for (size_t b = 1; b < nbatch; ++b) {
            status[b] = dev_status[b];
            if (status[b] == 0) {
                ++rerun;
                GT_TRY(run_sync(b, &probe));
            }
        }
    }
    uint64_t levels_total = 0;
    for (uint32_t st : status) levels_total += st;
    if (prof) {
        for (auto &e : ev) (void)hipEventDestroy(e);
        g->prof_ms[GT_SWEEP_BC_FORWARD] += fms;
        g->prof_launches[GT_SWEEP_BC_FORWARD] += launches_f;
        g->prof_ms[GT_SWEEP_BC_BACKWARD] += bms;
        g->prof_launches[GT_SWEEP_BC_BACKWARD] += launches_b;
    }

    for (int k = 1; k < nlanes; ++k) {
        const BcArgs &o = lanes[k].fa;
        hipLaunchKernelGGL(bc_add, dim3(lgrid(n)), dim3(kBlock), 0, s, base.vacc, o.vacc,
                           (uint64_t)n);
        hipLaunchKernelGGL(bc_add, dim3(lgrid(bentries)), dim3(kBlock), 0, s, base.eacc, o.eacc,
                           bentries);
        GT_HIP(hipGetLastError());
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
        const Orient &o = *bv.bwd;
        const uint32_t *fl = bv.undirected ? o.first_len.as<uint32_t>() : nullptr;
        for (int pass = 0; pass < (bv.undirected ? 2 : 1); ++pass)
            if (n)
                hipLaunchKernelGGL(bc_emit_edges, dim3(wgrid(n)), dim3(kBlock), 0, s,
                                   o.off.as<uint32_t>(), o.eid.as<uint32_t>(), fl, base.eacc, n,
                                   pass, edst);
        GT_HIP(hipGetLastError());
        if (out->loc == GT_HOST)
            GT_HIP(hipMemcpyAsync(out->edge, edst, eb * sizeof(double), hipMemcpyDeviceToHost, s));
    }
    GT_HIP(hipGetLastError());
    if (out->loc == GT_HOST && n)
        GT_HIP(hipMemcpyAsync(out->vertex, vdst, n * sizeof(double), hipMemcpyDeviceToHost, s));
    std::vector<unsigned long long> slots(2 * 3 * (size_t)kBcMaxWaves * nlanes);
    for (int k = 0; k < nlanes; ++k)
        GT_HIP(hipMemcpyAsync(slots.data() + (size_t)k * 2 * 3 * kBcMaxWaves, lanes[k].work,
                              2 * 3 * kBcMaxWaves * sizeof(unsigned long long),
                              hipMemcpyDeviceToHost, s));
    GT_HIP(hipStreamSynchronize(s));
    unsigned long long work[6] = {0, 0, 0, 0, 0, 0};
    for (int k = 0; k < nlanes; ++k)
        for (size_t w = 0; w < kBcMaxWaves; ++w)
            for (int p = 0; p < 2; ++p)
                for (int c = 0; c < 3; ++c)
                    work[3 * p + c] +=
                        slots[(size_t)k * 2 * 3 * kBcMaxWaves + (p * kBcMaxWaves + w) * 3 + c];
    if (stats) {
        stats->sources = src.size();
        stats->batches = nbatch;
        stats->batches_rerun = rerun;
        stats->levels = levels_total;
        stats->forward_entries = work[0];
        stats->forward_hits = work[1];
        stats->forward_hit_entries = work[2];
        stats->backward_entries = work[3];
        stats->backward_hits = work[4];
        stats->backward_hit_entries = work[5];
        stats->forward_ms = fms;
        stats->backward_ms = bms;
        stats->streams = (uint64_t)nlanes;
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
