// The pull-SpMV engine shared by the PageRank sweeps (and the FP32 form's FP64
// refinement), the eigenvector sweep, the adjoint's transposed sweep and mu
// pass, and the uneven batch's sweeps.
//
// Work is cut into wave tiles of <= kWaveTileNnz in-entries and <= kWaveTileRows
// rows; a row longer than kWaveTileNnz is cut into kWaveTileNnz-entry chunks
// whose partial sums go to slots of long_part.  Inside a tile the entries are
// stored sorted by source, each packed as ((source - base) << bits) | slot,
// where base is the tile's smallest source (tbase[tile]) and slot is the
// entry's position in row order: the 64 lanes of every gather instruction
// read ascending addresses, so lanes that touch the same cache line share one
// request, and the gathered value is dropped into LDS at `slot`, which
// restores row order for the row sums.  The packing costs no bytes over a
// plain 4-byte index stream.  It needs every tile's sources to span fewer
// than 2^(32 - bits) labels -- always true below 2^22 vertices, and true for
// a batch of graphs of that size whatever their total.  Otherwise (bits == 0,
// every base 0) the entries are the plain sources in row order and an
// entry's slot is its position.
//
// Every wavefront walks its own tiles (wave w of the grid takes tiles w,
// w + W, ...) with no workgroup barrier:
//   phase 1  kWaveTileNnz / 64 packed words per lane in flight (coalesced),
//            as many gathers, each value times the packed-order weight when
//            weighted, scattered into the wavefront's LDS tile; the next
//            tile's packed words and row offsets are loaded while the
//            gathers are in flight (see spmv_waves);
//   phase 2  one lane per row when the row has fewer than `short_row` entries
//            (a tuning parameter), the whole wavefront (shuffle tree) for
//            longer rows;
//   phase 3  `fin(row, sum)`, lane q for rows q, q + 64, ...: the rows'
//            global reads and writes go out coalesced.
// A split chunk is summed straight from registers into its slot.
//
// The finisher `fin` is called as fin(row, sum) for every whole row, and
// brackets every tile: fin.tile(t) (wave-uniform) before it -- returning false
// skips the tile, loads included -- and fin.tile_end(t) after it.  TileAll is
// the pair for finishers that need neither.  A finisher that skips tiles may
// also be given a list of tiles to walk (`tlist`, ntiles entries) in place of
// tiles 0 .. ntiles - 1.
//
// With HOT, the first `nhot` entries of the source vector are also in LDS
// (`hot`, filled by the caller before the walk): a gather below that label
// reads LDS and sends no request to L2.  After relabelling by gather
// frequency those are the most gathered vertices.  The values are the same,
// so the sums are too.
#pragma once

#include <algorithm>
#include <type_traits>

#include "device_common.hpp"
#include "gt_internal.hpp"

namespace gt {

static_assert(kWavesPerBlock * kWave == kBlock, "workgroups are whole wavefronts");

// Per-wavefront tiles of a workgroup of NW wavefronts, each with its row sums
// (ROWS; without, the row sums stay in registers and the LDS goes to a hot
// slice), and the workgroup's reduction slots (two per wavefront).  V is the
// tile's element type: the sums' type, or, for an unweighted sweep with
// register row sums, the narrower gathered type.
template <typename V, int NW = kWavesPerBlock, bool ROWS = true>
struct WaveLds {
    V vals[NW][kWaveTileNnz];
    V rowsum[NW][ROWS ? kWaveTileRows : 1];
    double red[2 * NW];
};

// Two workgroup sums at once over NW wavefronts, in wavefront order.
template <int NW>
__device__ __forceinline__ void block_sum2_waves(double &a, double &b, double *red) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    a = wave_sum(a);
    b = wave_sum(b);
    if (lane == 0) {
        red[wid] = a;
        red[NW + wid] = b;
    }
    __syncthreads();
    a = 0.0;
    b = 0.0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        a += red[w];
        b += red[NW + w];
    }
    __syncthreads();
}

// Default tile hooks: every tile is processed, nothing is done at its end.
// A finisher deriving from TileAll keeps tile() as it is (the engine then
// loads a wavefront's next tile before it is done with the current one).
struct TileAll {
    __device__ __forceinline__ bool tile(uint32_t) const { return true; }
    __device__ __forceinline__ void tile_end(uint32_t) const {}
};

// A finisher may split its per-row work in two: `In load(row)` issues the
// row's loads (any valid label) and `finish(row, sum, in)` completes it.  The
// hot-slice walk (register row sums) then issues a tile's row loads with
// its gathers, so their latency overlaps the gathers' instead of following
// the row sums, round after round: before the gathers, only the rounds the
// tile has (kLoadsFirst), or after them, all four without a branch.
template <typename F, typename = void>
struct has_row_loads : std::false_type {};
template <typename F>
struct has_row_loads<F, std::void_t<typename F::In>> : std::true_type {};
template <typename F, bool>
struct RowIn {
    struct type {};
    static constexpr bool first = false;
};
template <typename F>
struct RowIn<F, true> {
    using type = typename F::In;
    static constexpr bool first = F::kLoadsFirst;
};

// Orders this wavefront's LDS accesses: writes before the call are visible to
// its lanes' reads after it.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// A raw buffer resource over `bytes` bytes at `base`: a load at a byte offset
// at or past `bytes` returns 0 (the hardware's range check), so a masked lane
// needs no branch.  Loads without branches let the compiler count exactly how
// many are in flight, and wait only for the ones a value comes from.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_rsrc(const void *base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes,
                                             0x00020000);
}
constexpr int kNonTemporal = 2;  // cache policy of a buffer load: streamed once

__device__ __forceinline__ float buffer_load(__amdgpu_buffer_rsrc_t r, uint32_t off, float *) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0,
                                                                          kNonTemporal));
}
__device__ __forceinline__ double buffer_load(__amdgpu_buffer_rsrc_t r, uint32_t off, double *) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0,
                                                                           kNonTemporal));
}

// What a wavefront loads for a tile before its gathers: the descriptor, the
// packed words and, for a tile of whole rows, each row's offsets (lane q of
// step i holds row 64 i + q).  Every load is unconditional (lanes past the
// tile read 0 through the buffer range check).
struct WavePre {
    static constexpr int K = kWaveTileNnz / kWave;
    static constexpr int R = kWaveTileRows / kWave;
    Tile t;
    uint32_t base;  // source label of packed value 0 (wave-uniform)
    uint32_t p[K];
    uint32_t lo[R], hi[R];
};

// The tile table is read-only for a sweep: read through the constant address
// space, a wave-uniform descriptor is a scalar load into SGPRs, and the buffer
// resources built from it stay scalar too.
typedef const __attribute__((address_space(4))) Tile ConstTile;

__device__ __forceinline__ Tile wave_tile(const Tile *__restrict__ tiles, uint32_t ti) {
    ConstTile *c = (ConstTile *)tiles + ti;
    Tile t;
    t.row_begin = c->row_begin;
    t.row_end = c->row_end;
    t.nnz_begin = c->nnz_begin;
    t.nnz_end = c->nnz_end;
    return t;
}

// Entry i of a device list, as a scalar load.
__device__ __forceinline__ uint32_t list_entry(const uint32_t *list, uint32_t i) {
    return ((const __attribute__((address_space(4))) uint32_t *)list)[i];
}

// Loads of tile `t` (its descriptor and source base already read).
__device__ __forceinline__ void wave_fetch(const Tile &t, uint32_t base,
                                           const uint32_t *__restrict__ off,
                                           const uint32_t *__restrict__ packed, int lane,
                                           WavePre &f) {
    f.t = t;
    f.base = base;
    const uint32_t nb = f.t.nnz_begin, cnt = f.t.nnz_end - f.t.nnz_begin;
    const __amdgpu_buffer_rsrc_t rp = buffer_rsrc(packed + nb, 4 * cnt);
#pragma unroll
    for (int k = 0; k < WavePre::K; ++k)
        f.p[k] = __builtin_amdgcn_raw_buffer_load_b32(rp, 4 * (lane + k * kWave), 0, kNonTemporal);
    // A split chunk's row_begin is a partial slot, not a row: no offsets.
    const bool split = f.t.row_end == kSplit;
    const uint32_t rb = split ? 0u : f.t.row_begin;
    const uint32_t rows = split ? 0u : f.t.row_end - f.t.row_begin;
    const __amdgpu_buffer_rsrc_t ro = buffer_rsrc(off + rb, split ? 0u : 4 * (rows + 1));
#pragma unroll
    for (int i = 0; i < WavePre::R; ++i) {
        const uint32_t q = lane + i * kWave;
        f.lo[i] = __builtin_amdgcn_raw_buffer_load_b32(ro, 4 * q, 0, 0);
        f.hi[i] = __builtin_amdgcn_raw_buffer_load_b32(ro, 4 * (q + 1), 0, 0);
    }
}

// V is the value type of the sums; the gathered source vector may be narrower
// (S), each value widened as it is summed, and the LDS tile holds TT (V, or S
// for an unweighted sweep with register row sums).  With a
// finisher that takes every tile (one deriving from TileAll, which must not
// override tile()), a wavefront loads its next tile's packed words and row
// offsets right after the current tile's gathers, so their latency overlaps
// the current tile's sums.
//
// LDSONLY: every source is below `nhot` (a slice of the sliced sweep, `src`
// and `hot` both starting at the slice), so no global gather is issued.
// A wavefront walks tiles first, first + stride, ...; by default first is its
// index in the grid and stride the grid's wavefront count.
template <typename V, bool WEIGHTED, bool HOT = false, bool LDSONLY = false, typename TT, int NW,
          bool ROWS, typename S, typename Fin>
__device__ __forceinline__ void spmv_waves(const Tile *__restrict__ tiles, uint32_t ntiles,
                                           const uint32_t *__restrict__ off,
                                           const uint32_t *__restrict__ packed, uint32_t bits,
                                           const uint32_t *__restrict__ tbase,
                                           const S *__restrict__ src, const V *__restrict__ w,
                                           double *__restrict__ long_part, uint32_t short_row,
                                           WaveLds<TT, NW, ROWS> &L, Fin &fin,
                                           const S *hot = nullptr, uint32_t nhot = 0,
                                           const uint32_t *tlist = nullptr,
                                           uint32_t first = 0xffffffffu, uint32_t stride = 0) {
    // The LDS tile holds TT: V, or -- unweighted, row sums in registers -- the
    // narrower gathered type S, widened exactly when a row is summed.
    static_assert(std::is_same<TT, V>::value || (std::is_same<TT, S>::value && !WEIGHTED && !ROWS),
                  "a narrower tile holds unweighted gathered values only");
    constexpr int K = WavePre::K;
    constexpr bool PIPE = std::is_base_of<TileAll, Fin>::value;
    constexpr bool PRE = !ROWS && has_row_loads<Fin>::value;
    using In = typename RowIn<Fin, PRE>::type;
    constexpr bool FIRST = RowIn<Fin, PRE>::first;
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const uint32_t slot_mask = (1u << bits) - 1u;
    TT *vals = L.vals[wid];
    TT *rowsum = L.rowsum[wid];  // ROWS only (TT == V)
    static_assert(!LDSONLY || HOT, "an LDS-only walk gathers from the hot slice");
    const uint32_t nwaves = stride ? stride : gridDim.x * NW;
    WavePre cur, nxt;
    uint32_t ti = first != 0xffffffffu ? __builtin_amdgcn_readfirstlane(first)
                                       : blockIdx.x * NW + wid;
    if (PIPE && ti < ntiles)
        wave_fetch(wave_tile(tiles, ti), list_entry(tbase, ti), off, packed, lane, cur);
    for (; ti < ntiles; ti += nwaves) {
        // The tile: the list's ti-th entry when walking a list (finishers
        // that skip tiles only), else ti.
        const uint32_t tid = !PIPE && tlist ? list_entry(tlist, ti) : ti;
        if (!PIPE) {
            if (!fin.tile(tid)) continue;
            wave_fetch(wave_tile(tiles, tid), list_entry(tbase, tid), off, packed, lane, cur);
        }
        // The next tile's descriptor (the last tile reads its own again).
        const uint32_t tin = ti + nwaves < ntiles ? ti + nwaves : ntiles - 1;
        const Tile tnext = wave_tile(tiles, tin);
        const uint32_t bnext = list_entry(tbase, tin);
        const Tile t = cur.t;
        const uint32_t nb = t.nnz_begin, cnt = t.nnz_end - t.nnz_begin;
        // 0. The rows' own loads, when the finisher takes them first: only
        //    the rounds of 64 rows the tile has (none for a split chunk).
        //    Issued before the gathers, they are done when the gathers are,
        //    whichever rounds were taken.
        In in[WavePre::R];
        if constexpr (PRE && FIRST) {
            const uint32_t nr = t.row_end == kSplit ? 0u : t.row_end - t.row_begin;
#pragma unroll
            for (int i = 0; i < WavePre::R; ++i) {
                if ((uint32_t)i * kWave >= nr) break;
                const uint32_t q = lane + i * kWave;
                in[i] = fin.load(t.row_begin + (q < nr ? q : nr - 1u));
            }
        }
        // 1. The gathers (global ones only above the hot slice).
        S g[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const uint32_t j = lane + k * kWave;
            const uint32_t sx = (cur.p[k] >> bits) + cur.base;
            g[k] = S(0);
            if (!LDSONLY && j < cnt && (!HOT || sx >= nhot)) g[k] = src[sx];
        }
        // 2. The weights, then the next tile's words and offsets (the last
        //    tile loads itself again): both without branches, so waiting
        //    for the gathers leaves them in flight.
        V wt[K];
        if (WEIGHTED) {
            const __amdgpu_buffer_rsrc_t rw = buffer_rsrc(w + nb, sizeof(V) * cnt);
#pragma unroll
            for (int k = 0; k < K; ++k)
                wt[k] = buffer_load(rw, sizeof(V) * (lane + k * kWave), (V *)nullptr);
        }
        // Otherwise after the gathers: every round, without branches (a
        //    split chunk loads row 0 and uses nothing), so waiting for the
        //    gathers leaves them in flight.
        if constexpr (PRE && !FIRST) {
            const bool split = t.row_end == kSplit;
            const uint32_t nr = split ? 1u : t.row_end - t.row_begin;
            const uint32_t r0 = split ? 0u : t.row_begin;
#pragma unroll
            for (int i = 0; i < WavePre::R; ++i) {
                const uint32_t q = lane + i * kWave;
                in[i] = fin.load(r0 + (q < nr ? q : nr - 1u));
            }
        }
        if (PIPE) wave_fetch(tnext, bnext, off, packed, lane, nxt);
        // Nothing that reads a gathered value moves above the loads issued so far.
        __builtin_amdgcn_sched_barrier(0);
        // 3. The hot slice (every lane reads LDS; a lane above it reads slot 0).
        V v[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const uint32_t j = lane + k * kWave;
            const uint32_t sx = (cur.p[k] >> bits) + cur.base;
            S x = g[k];
            if (LDSONLY) {
                x = hot[sx < nhot ? sx : 0u];
            } else if (HOT) {
                const S h = hot[sx < nhot ? sx : 0u];
                x = sx < nhot ? h : x;
            }
            v[k] = j < cnt ? static_cast<V>(x) : V(0);
            if (WEIGHTED) v[k] *= wt[k];
        }
        if (t.row_end == kSplit) {
            V acc = V(0);
#pragma unroll
            for (int k = 0; k < K; ++k) acc += v[k];
            const double tot = wave_sum(static_cast<double>(acc));
            if (lane == 0) long_part[t.row_begin] = tot;
            fin.tile_end(tid);
        } else {
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const uint32_t j = lane + k * kWave;
                if (j < cnt) vals[bits ? (cur.p[k] & slot_mask) : j] = static_cast<TT>(v[k]);
            }
            wave_sync();
            const uint32_t rb = t.row_begin, rows = t.row_end - t.row_begin;
            // Row q's sum in rowsum[q] (ROWS), else row 64 i + lane's in rs[i].
            V rs[WavePre::R];
#pragma unroll
            for (int i = 0; i < WavePre::R; ++i) {
                rs[i] = V(0);
                const uint32_t q0 = i * kWave;
                if (q0 >= rows) continue;
                const uint32_t q = q0 + lane;
                uint32_t lo = 0, hi = 0;
                bool longrow = false;
                if (q < rows) {
                    lo = cur.lo[i] - nb;
                    hi = cur.hi[i] - nb;
                    if (hi - lo < short_row) {
                        V acc = V(0);
                        for (uint32_t k = lo; k < hi; ++k) acc += static_cast<V>(vals[k]);
                        if (ROWS)
                            rowsum[q] = acc;
                        else
                            rs[i] = acc;
                    } else {
                        longrow = true;
                    }
                }
                uint64_t lmask = __ballot(longrow);
                while (lmask) {
                    const int r = __builtin_ctzll(lmask);
                    lmask &= lmask - 1;
                    const uint32_t rlo = __shfl(lo, r, kWave), rhi = __shfl(hi, r, kWave);
                    V acc = V(0);
                    for (uint32_t k = rlo + lane; k < rhi; k += kWave)
                        acc += static_cast<V>(vals[k]);
                    acc = wave_sum(acc);  // every lane holds the same total
                    if (ROWS) {
                        if (lane == 0) rowsum[q0 + r] = acc;
                    } else if (lane == r) {
                        rs[i] = acc;
                    }
                }
            }
            if (ROWS) {
                wave_sync();
                for (uint32_t q = lane; q < rows; q += kWave)
                    fin(rb + q, static_cast<double>(rowsum[q]));
            } else if constexpr (PRE) {
#pragma unroll
                for (int i = 0; i < WavePre::R; ++i) {
                    const uint32_t q = lane + i * kWave;
                    if (q < rows) fin.finish(rb + q, static_cast<double>(rs[i]), in[i]);
                }
            } else {
                // One copy of the finisher: the sums rotate through rs[0].
#pragma unroll 1
                for (uint32_t q = lane; q < rows; q += kWave) {
                    fin(rb + q, static_cast<double>(rs[0]));
#pragma unroll
                    for (int i = 0; i + 1 < WavePre::R; ++i) rs[i] = rs[i + 1];
                }
            }
            fin.tile_end(tid);
            wave_sync();  // the LDS tile is reused by this wavefront's next tile
        }
        if (PIPE) cur = nxt;
    }
}

// The hot-slice sweeps (relabelled graphs): one workgroup of kHotWaves
// wavefronts per compute unit (two per SIMD: with 12 or 16 per CU and the
// smaller slice that leaves, the R-MAT 22 sweeps were slower) whose LDS
// holds the wave tiles and the first `nhot` entries of the gathered vector
// (the most gathered vertices), copied in at the start of the sweep.  nhot
// is a multiple of 4, at most hot_max<V>().
constexpr int kHotWaves = 8;
constexpr int kHotBlock = kHotWaves * kWave;
constexpr uint32_t kHotMax = 32704;    // FP32: 127.75 KB; with the tiles, 163,744 B of 163,840
constexpr uint32_t kHotMax64 = 12224;  // FP64: 95.5 KB; with the tiles, 163,520 B
static_assert(sizeof(WaveLds<float, kHotWaves, false>) + 4 * kHotMax <= 163840, "LDS of one CU");
static_assert(sizeof(WaveLds<double, kHotWaves, false>) + 8 * kHotMax64 <= 163840,
              "LDS of one CU");
template <typename V>
constexpr uint32_t hot_max() {
    return sizeof(V) == 4 ? kHotMax : kHotMax64;
}

// Entries below `nhot` (a multiple of 4, at most hot_max<T>(); `src` 16-byte
// aligned) into LDS by a kHotBlock-thread workgroup, then a workgroup
// barrier.  Every thread issues all its 16-byte loads before the first LDS
// write, so the fill costs one memory round trip, not one per step.
template <typename T>
__device__ __forceinline__ void fill_hot(const T *src, uint32_t nhot, T *hot) {
    constexpr uint32_t per = 16 / sizeof(T);
    constexpr int kSteps = (int)((hot_max<T>() / per + kHotBlock - 1) / kHotBlock);
    // No branches: a step past nhot reads entries 0..3 again and writes them
    // back unchanged, so all loads are in flight before the first wait.
    float4 v[kSteps];
#pragma unroll
    for (int r = 0; r < kSteps; ++r) {
        const uint32_t i = per * (threadIdx.x + (uint32_t)r * kHotBlock);
        v[r] = *reinterpret_cast<const float4 *>(src + (i < nhot ? i : 0u));
    }
#pragma unroll
    for (int r = 0; r < kSteps; ++r) {
        const uint32_t i = per * (threadIdx.x + (uint32_t)r * kHotBlock);
        *reinterpret_cast<float4 *>(hot + (i < nhot ? i : 0u)) = v[r];
    }
    __syncthreads();
}

// The hot-slice length for a tuning value in KB (<= 0: off), n vertices and
// value type V.
template <typename V>
inline uint32_t hot_len(int kb, uint32_t n) {
    if (kb <= 0) return 0;
    const uint64_t want = (uint64_t)kb * 1024 / sizeof(V);
    return (uint32_t)std::min<uint64_t>(std::min<uint64_t>(want, hot_max<V>()), n) & ~3u;
}

// The sliced sweep's extra arguments (see Sliced): piece tiles with their
// offsets, slots and packed stream; the segments of each piece workgroup;
// the light rows' tiles and stream.
struct SlicedArgs {
    const Tile *htiles;
    const uint32_t *htbase;
    const uint32_t *hoff;
    const uint32_t *hpacked;
    const SliceSeg *seg;
    const uint32_t *wseg;
    uint32_t heavy_wgs;
    uint32_t nltiles;
    const Tile *ltiles;
    const uint32_t *ltbase;
    const uint32_t *lpacked;
    uint32_t lbits;
    uint32_t nmtiles;
    const Tile *mtiles;
    const uint32_t *mtbase;
    const uint32_t *moff;
    const uint32_t *mpacked;
    uint32_t mbits;
    uint32_t mpart;
};

// A piece's sum into the partial of the same index (coalesced: a tile's
// pieces are consecutive); the fold finds a row's partials by index.
struct PieceFin : TileAll {
    double *long_part;  // already offset to the stream's first partial
    __device__ __forceinline__ void operator()(uint32_t piece, double sum) const {
        long_part[piece] = sum;
    }
};

// The piece workgroup's share of a sliced sweep: for each of its segments,
// the slice of `src` into LDS (`hot`), then its wavefronts walk the
// segment's tiles gathering from LDS only.  TT is the tile's element type.
template <typename V, typename TT, typename S>
__device__ __forceinline__ void sliced_pieces(const SlicedArgs &sa, const S *src,
                                              double *long_part, uint32_t short_row,
                                              WaveLds<TT, kHotWaves, false> &L, S *hot) {
    const uint32_t wid = threadIdx.x / kWave;
    PieceFin fin;
    fin.long_part = long_part;
    for (uint32_t q = sa.wseg[blockIdx.x]; q < sa.wseg[blockIdx.x + 1]; ++q) {
        const SliceSeg sg = sa.seg[q];
        __syncthreads();  // every wavefront is done with the previous slice
        fill_hot(src + sg.base, sg.len, hot);
        spmv_waves<V, false, true, true>(sa.htiles + sg.t0, sg.t1 - sg.t0, sa.hoff, sa.hpacked,
                                         kWaveSlotBits, sa.htbase + sg.t0, src + sg.base,
                                         static_cast<const V *>(nullptr), long_part, short_row,
                                         L, fin, static_cast<const S *>(hot), sg.len, nullptr,
                                         wid, (uint32_t)kHotWaves);
    }
}

// The light workgroups' share of a sliced sweep (hot slice already in LDS):
// the light rows with their finisher, then the heavy rows' main pieces.
// This workgroup is light workgroup b of nb.
template <typename V, typename TT, typename S, typename Fin>
__device__ __forceinline__ void sliced_light(const SlicedArgs &sa, const uint32_t *off,
                                             const S *src, double *long_part, uint32_t short_row,
                                             WaveLds<TT, kHotWaves, false> &L, Fin &fin,
                                             const S *hot, uint32_t nhot, uint32_t b,
                                             uint32_t nb) {
    const uint32_t wid = threadIdx.x / kWave;
    spmv_waves<V, false, true>(sa.ltiles, sa.nltiles, off, sa.lpacked, sa.lbits, sa.ltbase, src,
                               static_cast<const V *>(nullptr), long_part, short_row, L, fin, hot,
                               nhot, nullptr, b * kHotWaves + wid, nb * kHotWaves);
    PieceFin pf;
    pf.long_part = long_part + sa.mpart;
    spmv_waves<V, false, true>(sa.mtiles, sa.nmtiles, sa.moff, sa.mpacked, sa.mbits, sa.mtbase,
                               src, static_cast<const V *>(nullptr), long_part, short_row, L, pf,
                               hot, nhot, nullptr, b * kHotWaves + wid, nb * kHotWaves);
}

// Sum of a split row's partials through an index list: long_part[long_idx[q]]
// for q over the row's slots, in slot order, 16 loads in flight.
__device__ __forceinline__ double long_row_sum_idx(const double *long_part,
                                                   const uint32_t *long_slot,
                                                   const uint32_t *long_idx, uint32_t k) {
    const uint32_t e = long_slot[k + 1];
    uint32_t q = long_slot[k];
    double acc = 0.0;
    for (; q + 16 <= e; q += 16) {
        uint32_t ix[16];
        double v[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) ix[i] = long_idx[q + i];
#pragma unroll
        for (int i = 0; i < 16; ++i) v[i] = long_part[ix[i]];
#pragma unroll
        for (int i = 0; i < 16; ++i) acc += v[i];
    }
    for (; q < e; ++q) acc += long_part[long_idx[q]];
    return acc;
}

// Sum of a split row's partial slots, in chunk order.
__device__ __forceinline__ double long_row_sum(const double *long_part, const uint32_t *long_slot,
                                               uint32_t k) {
    // Slot order, 32 loads in flight, then eight: a hub has hundreds of slots
    // and sets the fold's critical path (the sum's order is the slot order).
    const uint32_t e = long_slot[k + 1];
    uint32_t q = long_slot[k];
    double acc = 0.0;
    for (; q + 32 <= e; q += 32) {
        double v[32];
#pragma unroll
        for (int i = 0; i < 32; ++i) v[i] = long_part[q + i];
#pragma unroll
        for (int i = 0; i < 32; ++i) acc += v[i];
    }
    for (; q + 8 <= e; q += 8) {
        double v[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) v[i] = long_part[q + i];
#pragma unroll
        for (int i = 0; i < 8; ++i) acc += v[i];
    }
    for (; q < e; ++q) acc += long_part[q];
    return acc;
}

}  // namespace gt
