// The pull-SpMV tile engine shared by the PageRank, eigenvector and adjoint
// sweeps.
//
// A workgroup of 256 threads walks tiles blockIdx.x, blockIdx.x + gridDim.x,
// ...  For each tile:
//   phase 1  the tile's <= kTileNnz in-entries are streamed (coalesced u32
//            index loads, eight per thread in flight) and the gathered source
//            values -- times the pull-ordered weight when weighted -- land in
//            LDS;
//   phase 2  row sums out of LDS into an LDS row buffer: one thread per row
//            when the tile's rows average fewer than 16 entries, one
//            wavefront per row with a shuffle-tree reduction otherwise; a
//            split-row chunk is reduced by the whole workgroup into its
//            partial slot instead;
//   phase 3  `fin(row, sum)` for every row of the tile, one thread per row,
//            so the rows' global reads and writes go out together and
//            coalesced.
// Three barriers per tile: after phase 1; after the short rows of phase 2
// (the long-row queue is complete); after the long rows (the LDS tile is free
// for the next tile's phase 1, and the row buffer is not written again before
// the next tile's first barrier).
#pragma once

#include "device_common.hpp"
#include "gt_internal.hpp"

namespace gt {

template <typename V, bool WEIGHTED>
__device__ __forceinline__ void gather_tile(const uint32_t *__restrict__ idx,
                                            const V *__restrict__ src,
                                            const V *__restrict__ w, uint32_t nb,
                                            uint32_t cnt, V *lds) {
    constexpr int K = kTileNnz / kBlock;
    uint32_t s[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const uint32_t j = threadIdx.x + k * kBlock;
        s[k] = j < cnt ? __builtin_nontemporal_load(idx + nb + j) : 0u;
    }
    V v[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const uint32_t j = threadIdx.x + k * kBlock;
        v[k] = j < cnt ? src[s[k]] : V(0);
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const uint32_t j = threadIdx.x + k * kBlock;
        if (j < cnt) {
            V x = v[k];
            if (WEIGHTED) x *= __builtin_nontemporal_load(w + nb + j);
            lds[j] = x;
        }
    }
}

// Rows of at least this many entries are summed by a wavefront.
constexpr uint32_t kShortRow = 32;

// LDS a kernel using spmv_tiles must provide.  nlongq must be zero when the
// tile loop starts (spmv_tiles resets it after each tile).
template <typename V>
struct TileLds {
    V vals[kTileNnz];
    double rowsum[kTileRows];
    double red[8];
    uint16_t longq[kTileNnz / kShortRow];
    uint32_t nlongq;
};

template <typename V, bool WEIGHTED, typename Fin>
__device__ __forceinline__ void spmv_tiles(const Tile *__restrict__ tiles, uint32_t ntiles,
                                           const uint32_t *__restrict__ off,
                                           const uint32_t *__restrict__ idx,
                                           const V *__restrict__ src, const V *__restrict__ w,
                                           double *__restrict__ long_part, TileLds<V> &L,
                                           Fin &fin) {
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wid = tid / kWave;
    if (tid == 0) L.nlongq = 0;  // published by the first tile's barrier
    for (uint32_t ti = blockIdx.x; ti < ntiles; ti += gridDim.x) {
        const Tile t = tiles[ti];
        const uint32_t nb = t.nnz_begin, cnt = t.nnz_end - t.nnz_begin;
        gather_tile<V, WEIGHTED>(idx, src, w, nb, cnt, L.vals);
        __syncthreads();
        if (t.row_end == kSplit) {
            double v = 0.0;
            for (uint32_t k = tid; k < cnt; k += kBlock) v += static_cast<double>(L.vals[k]);
            v = block_sum(v, L.red);  // ends with a barrier: LDS is free again
            if (tid == 0) long_part[t.row_begin] = v;
            continue;
        }
        // Rows shorter than kShortRow: one thread each.  Longer rows (at most
        // kTileNnz / kShortRow of them) are queued and summed by a wavefront
        // each, so no thread walks a long row alone whatever the tile's mix.
        const uint32_t rb = t.row_begin, rows = t.row_end - t.row_begin;
        for (uint32_t q = tid; q < rows; q += kBlock) {
            const uint32_t lo = off[rb + q] - nb, hi = off[rb + q + 1] - nb;
            if (hi - lo < kShortRow) {
                V acc = V(0);
                for (uint32_t k = lo; k < hi; ++k) acc += L.vals[k];
                L.rowsum[q] = static_cast<double>(acc);
            } else {
                L.longq[atomicAdd(&L.nlongq, 1u)] = (uint16_t)q;
            }
        }
        __syncthreads();
        const uint32_t nq = L.nlongq;
        for (uint32_t i = wid; i < nq; i += kBlock / kWave) {
            const uint32_t q = L.longq[i];
            const uint32_t lo = off[rb + q] - nb, hi = off[rb + q + 1] - nb;
            V acc = V(0);
            for (uint32_t k = lo + lane; k < hi; k += kWave) acc += L.vals[k];
            acc = wave_sum(acc);
            if (lane == 0) L.rowsum[q] = static_cast<double>(acc);
        }
        __syncthreads();
        if (tid == 0) L.nlongq = 0;
        for (uint32_t q = tid; q < rows; q += kBlock) fin(rb + q, L.rowsum[q]);
    }
    __syncthreads();
}

// Sum of a split row's partial slots, in chunk order.
__device__ __forceinline__ double long_row_sum(const double *long_part, const uint32_t *long_slot,
                                               uint32_t k) {
    double acc = 0.0;
    for (uint32_t q = long_slot[k]; q < long_slot[k + 1]; ++q) acc += long_part[q];
    return acc;
}

}  // namespace gt
