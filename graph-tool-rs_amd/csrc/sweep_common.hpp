// The pull-SpMV engine shared by the PageRank, eigenvector and adjoint sweeps.
//
// Work is cut into wave tiles of <= kWaveTileNnz in-entries and <= kWaveTileRows
// rows; a row longer than kWaveTileNnz is cut into kWaveTileNnz-entry chunks
// whose partial sums go to slots of long_part.  Inside a tile the entries are
// stored sorted by source, each packed as (source << bits) | slot, where slot
// is the entry's position in row order: the 64 lanes of every gather
// instruction read ascending addresses, so lanes that touch the same cache
// line share one request, and the gathered value is dropped into LDS at
// `slot`, which restores row order for the row sums.  The packing costs no
// bytes over a plain 4-byte index stream.  With bits == 0 (vertex ids too wide
// to pack) the entries are the plain sources in row order and an entry's slot
// is its position.
//
// Every wavefront walks its own tiles (wave w of the grid takes tiles w,
// w + W, ...) with no workgroup barrier:
//   phase 1  kWaveTileNnz / 64 packed words per lane in flight (coalesced),
//            as many gathers, each value times the packed-order weight when
//            weighted, scattered into the wavefront's LDS tile;
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
// the pair for finishers that need neither.
#pragma once

#include "device_common.hpp"
#include "gt_internal.hpp"

namespace gt {

static_assert(kWavesPerBlock * kWave == kBlock, "workgroups are whole wavefronts");

template <typename V>
struct WaveLds {
    V vals[kWavesPerBlock][kWaveTileNnz];
    V rowsum[kWavesPerBlock][kWaveTileRows];
    double red[8];
};

// Default tile hooks: every tile is processed, nothing is done at its end.
struct TileAll {
    __device__ __forceinline__ bool tile(uint32_t) const { return true; }
    __device__ __forceinline__ void tile_end(uint32_t) const {}
};

// Orders this wavefront's LDS accesses: writes before the call are visible to
// its lanes' reads after it.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// V is the value type of the sums (and of the LDS tile); the gathered source
// vector may be narrower (S), each value widened as it arrives.
template <typename V, bool WEIGHTED, typename S, typename Fin>
__device__ __forceinline__ void spmv_waves(const Tile *__restrict__ tiles, uint32_t ntiles,
                                           const uint32_t *__restrict__ off,
                                           const uint32_t *__restrict__ packed, uint32_t bits,
                                           const S *__restrict__ src, const V *__restrict__ w,
                                           double *__restrict__ long_part, uint32_t short_row,
                                           WaveLds<V> &L, Fin &fin) {
    constexpr int K = kWaveTileNnz / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const uint32_t slot_mask = (1u << bits) - 1u;
    V *vals = L.vals[wid];
    V *rowsum = L.rowsum[wid];
    const uint32_t nwaves = gridDim.x * kWavesPerBlock;
    for (uint32_t ti = blockIdx.x * kWavesPerBlock + wid; ti < ntiles; ti += nwaves) {
        if (!fin.tile(ti)) continue;
        const Tile t = tiles[ti];
        const uint32_t nb = t.nnz_begin, cnt = t.nnz_end - t.nnz_begin;
        uint32_t p[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const uint32_t j = lane + k * kWave;
            p[k] = j < cnt ? __builtin_nontemporal_load(packed + nb + j) : 0u;
        }
        V v[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const uint32_t j = lane + k * kWave;
            v[k] = j < cnt ? static_cast<V>(src[p[k] >> bits]) : V(0);
        }
        if (WEIGHTED) {
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const uint32_t j = lane + k * kWave;
                if (j < cnt) v[k] *= __builtin_nontemporal_load(w + nb + j);
            }
        }
        if (t.row_end == kSplit) {
            V acc = V(0);
#pragma unroll
            for (int k = 0; k < K; ++k) acc += v[k];
            const double tot = wave_sum(static_cast<double>(acc));
            if (lane == 0) long_part[t.row_begin] = tot;
            fin.tile_end(ti);
            continue;
        }
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const uint32_t j = lane + k * kWave;
            if (j < cnt) vals[bits ? (p[k] & slot_mask) : j] = v[k];
        }
        wave_sync();
        const uint32_t rb = t.row_begin, rows = t.row_end - t.row_begin;
        for (uint32_t q0 = 0; q0 < rows; q0 += kWave) {
            const uint32_t q = q0 + lane;
            uint32_t lo = 0, hi = 0;
            bool longrow = false;
            if (q < rows) {
                lo = off[rb + q] - nb;
                hi = off[rb + q + 1] - nb;
                if (hi - lo < short_row) {
                    V acc = V(0);
                    for (uint32_t k = lo; k < hi; ++k) acc += vals[k];
                    rowsum[q] = acc;
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
                for (uint32_t k = rlo + lane; k < rhi; k += kWave) acc += vals[k];
                acc = wave_sum(acc);
                if (lane == 0) rowsum[q0 + r] = acc;
            }
        }
        wave_sync();
        for (uint32_t q = lane; q < rows; q += kWave)
            fin(rb + q, static_cast<double>(rowsum[q]));
        fin.tile_end(ti);
        wave_sync();  // the LDS tile is reused by this wavefront's next tile
    }
}

// Sum of a split row's partial slots, in chunk order.
__device__ __forceinline__ double long_row_sum(const double *long_part, const uint32_t *long_slot,
                                               uint32_t k) {
    // Slot order, eight loads in flight: a hub has hundreds of slots.
    const uint32_t e = long_slot[k + 1];
    uint32_t q = long_slot[k];
    double acc = 0.0;
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
