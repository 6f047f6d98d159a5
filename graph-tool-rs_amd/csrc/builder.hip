// Device builder: the adjacency AdjList::add_edge(src[i], tgt[i]), i = 0..m,
// produces (crates/gt-core/src/adj/list.rs:258-280, 471-500;
// adj/block.rs:171-209), built on the GPU with integer work only.
//
//   1. degree histogram      out_deg[src[i]]++, in_deg[tgt[i]]++ (atomics:
//                            the counts do not depend on their order)
//   2. exclusive scan        -> out_off, in_off (n+1 each)
//   3. stable scatter        edge ids grouped by source (the out-halves) and
//                            by target, each group in id order: an LSD radix
//                            sort, 8 bits per pass, whose scatter ranks
//                            equal digits by wavefront ballots, so it is
//                            stable and the result is unique
//   4. push-swap replay      per vertex v, the in-half is a queue: an
//                            in-insert appends, an out-insert while the queue
//                            is non-empty moves its front to the back; a
//                            self-loop's out-insert comes first.  Replayed
//                            over v's two id-sorted lists merged in id order,
//                            in a ring buffer of capacity in_degree(v)
//   5. gathers               in_other = src[in_eid], out_other = tgt[out_eid]
// The result equals the host adjacency word for word.
#include <algorithm>
#include <deque>

#include "device_common.hpp"
#include "gt_internal.hpp"

namespace gt {

constexpr int kScanItems = 16;                     // per thread
constexpr uint32_t kScanTile = kBlock * kScanItems;  // 4096 per workgroup

// ------------------------------------------------------------------ scan

// Exclusive sum over one workgroup's values (blocked: thread t owns
// items t*16 .. t*16+15 of the tile); returns the tile total.
__device__ __forceinline__ uint32_t block_exclusive(uint32_t (&v)[kScanItems], uint32_t *lds) {
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        uint32_t x = v[k];
        v[k] = sum;
        sum += x;
    }
    // Scan the per-thread sums across the workgroup (Hillis-Steele in LDS).
    lds[threadIdx.x] = sum;
    __syncthreads();
    for (int o = 1; o < kBlock; o <<= 1) {
        uint32_t t = threadIdx.x >= (unsigned)o ? lds[threadIdx.x - o] : 0u;
        __syncthreads();
        lds[threadIdx.x] += t;
        __syncthreads();
    }
    const uint32_t incl = lds[threadIdx.x];
    const uint32_t total = lds[kBlock - 1];
    __syncthreads();
    const uint32_t excl = incl - sum;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) v[k] += excl;
    return total;
}

__global__ __launch_bounds__(kBlock) void scan_tile_sums(const uint32_t *__restrict__ in,
                                                         uint64_t len, uint32_t *__restrict__ sums) {
    __shared__ uint32_t lds[kBlock];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
        if (base + k < len) s += in[base + k];
    lds[threadIdx.x] = s;
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
        if (threadIdx.x < (unsigned)o) lds[threadIdx.x] += lds[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) sums[blockIdx.x] = lds[0];
}

__global__ __launch_bounds__(kBlock) void scan_tiles(const uint32_t *__restrict__ in, uint64_t len,
                                                     const uint32_t *__restrict__ tile_base,
                                                     uint32_t *__restrict__ out) {
    __shared__ uint32_t lds[kBlock];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) v[k] = base + k < len ? in[base + k] : 0u;
    block_exclusive(v, lds);
    const uint32_t add = tile_base ? tile_base[blockIdx.x] : 0u;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
        if (base + k < len) out[base + k] = v[k] + add;
}

__global__ void scan_total(const uint32_t *in, const uint32_t *out, uint64_t len,
                           uint32_t *total) {
    if (threadIdx.x == 0) total[0] = out[len - 1] + in[len - 1];
}

// out[i] = in[0] + ... + in[i-1] for i < len, and out[len] = total when
// `with_total`.  Sums are modulo 2^32; callers keep totals below that.
gt_status exclusive_scan_u32(const uint32_t *in, uint64_t len, uint32_t *out, bool with_total,
                                hipStream_t s, std::deque<DevBuf> &tmp) {
    const uint64_t tiles = (len + kScanTile - 1) / kScanTile;
    if (len == 0) {
        if (with_total) GT_HIP(hipMemsetAsync(out, 0, sizeof(uint32_t), s));
        return GT_OK;
    }
    if (tiles == 1) {
        hipLaunchKernelGGL(scan_tiles, dim3(1), dim3(kBlock), 0, s, in, len, nullptr, out);
    } else {
        tmp.emplace_back();
        DevBuf &sums = tmp.back();
        GT_TRY(sums.reserve(sizeof(uint32_t) * (tiles + 1)));
        hipLaunchKernelGGL(scan_tile_sums, dim3(tiles), dim3(kBlock), 0, s, in, len,
                           sums.as<uint32_t>());
        tmp.emplace_back();
        DevBuf &bases = tmp.back();
        GT_TRY(bases.reserve(sizeof(uint32_t) * (tiles + 1)));
        GT_TRY(exclusive_scan_u32(sums.as<uint32_t>(), tiles, bases.as<uint32_t>(), false, s, tmp));
        hipLaunchKernelGGL(scan_tiles, dim3(tiles), dim3(kBlock), 0, s, in, len,
                           bases.as<uint32_t>(), out);
    }
    GT_HIP(hipGetLastError());
    if (with_total) {
        // total = out[len-1] + in[len-1]
        hipLaunchKernelGGL(scan_total, dim3(1), dim3(64), 0, s, in, out, len, out + len);
        GT_HIP(hipGetLastError());
    }
    return GT_OK;
}

// ------------------------------------------------------------ histogram

__global__ __launch_bounds__(kBlock) void check_endpoints(const uint32_t *__restrict__ src,
                                                          const uint32_t *__restrict__ tgt,
                                                          uint64_t m, uint32_t n,
                                                          uint32_t *__restrict__ bad) {
    uint32_t b = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < m;
         i += (uint64_t)gridDim.x * kBlock)
        b |= (src[i] >= n) | (tgt[i] >= n);
    if (b) atomicOr(bad, 1u);
}

__global__ __launch_bounds__(kBlock) void degree_hist(const uint32_t *__restrict__ src,
                                                      const uint32_t *__restrict__ tgt, uint64_t m,
                                                      uint32_t *__restrict__ outd,
                                                      uint32_t *__restrict__ ind) {
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < m;
         i += (uint64_t)gridDim.x * kBlock) {
        atomicAdd(&outd[src[i]], 1u);
        atomicAdd(&ind[tgt[i]], 1u);
    }
}

// ------------------------------------------------------- LSD radix sort

// Pass over 8 bits of the key starting at `shift`.  Tiles of kScanTile
// items; hist is digit-major: hist[d * ntiles + tile].
__global__ __launch_bounds__(kBlock) void radix_hist(const uint32_t *__restrict__ keys,
                                                     uint64_t m, int shift, uint32_t ntiles,
                                                     uint32_t *__restrict__ hist) {
    __shared__ uint32_t cnt[256];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
    for (uint32_t k = threadIdx.x; k < kScanTile; k += kBlock) {
        const uint64_t i = base + k;
        if (i < m) atomicAdd(&cnt[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(uint64_t)threadIdx.x * ntiles + blockIdx.x] = cnt[threadIdx.x];
}

// Stable scatter.  Items are taken in 16 rounds of 256 consecutive items;
// inside a round, a lane's rank among equal digits is the number of lower
// lanes of its wavefront with that digit (eight ballots) plus the counts of
// lower wavefronts, plus the running count of earlier rounds.  `vals` may be
// null: the value is then the item's own index (an edge id).
__global__ __launch_bounds__(kBlock) void radix_scatter(const uint32_t *__restrict__ keys,
                                                        const uint32_t *__restrict__ vals,
                                                        uint64_t m, int shift, uint32_t ntiles,
                                                        const uint32_t *__restrict__ scanned,
                                                        uint32_t *__restrict__ keys_out,
                                                        uint32_t *__restrict__ vals_out) {
    __shared__ uint32_t base[256];
    __shared__ uint32_t wcnt[kBlock / kWave][256];
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wid = tid / kWave;
    base[tid] = scanned[(uint64_t)tid * ntiles + blockIdx.x];
#pragma unroll
    for (int w = 0; w < kBlock / kWave; ++w) wcnt[w][tid] = 0;
    __syncthreads();
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const uint64_t tile0 = (uint64_t)blockIdx.x * kScanTile;
    for (int r = 0; r < kScanItems; ++r) {
        const uint64_t i = tile0 + (uint64_t)r * kBlock + tid;
        const bool valid = i < m;
        const uint32_t key = valid ? keys[i] : 0u;
        const uint32_t d = (key >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const uint64_t bal = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? bal : ~bal;
        }
        const uint32_t below = __popcll(peers & lt_mask);
        const bool leader = valid && below == 0;
        if (leader) wcnt[wid][d] = __popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t pos = base[d] + below;
            for (int w = 0; w < wid; ++w) pos += wcnt[w][d];
            keys_out[pos] = key;
            vals_out[pos] = vals ? vals[i] : (uint32_t)i;
        }
        __syncthreads();
        uint32_t add = 0;
#pragma unroll
        for (int w = 0; w < kBlock / kWave; ++w) {
            add += wcnt[w][tid];
            wcnt[w][tid] = 0;
        }
        base[tid] += add;
        __syncthreads();
    }
}

// Items 0..m-1 sorted stably by key[i] (bits below `bits` significant);
// ids_out receives vals[i] (or i when vals is null) in sorted order.
gt_status stable_sort_ids(const uint32_t *key, const uint32_t *vals, uint64_t m, int bits,
                          DevBuf &ids_out, hipStream_t s, std::deque<DevBuf> &tmp) {
    const uint32_t ntiles = (uint32_t)((m + kScanTile - 1) / kScanTile);
    DevBuf kA, kB, vA;
    GT_TRY(kA.reserve(sizeof(uint32_t) * m));
    GT_TRY(kB.reserve(sizeof(uint32_t) * m));
    GT_TRY(vA.reserve(sizeof(uint32_t) * m));
    GT_TRY(ids_out.reserve(sizeof(uint32_t) * m));
    DevBuf hist, scanned;
    GT_TRY(hist.reserve(sizeof(uint32_t) * 256 * (size_t)ntiles));
    GT_TRY(scanned.reserve(sizeof(uint32_t) * 256 * (size_t)ntiles));
    const int passes = std::max(1, (bits + 7) / 8);
    const uint32_t *kin = key;
    const uint32_t *vin = vals;  // null: the first pass takes the item's own index
    // Ping-pong so the last pass lands in ids_out.
    uint32_t *kbuf[2] = {kA.as<uint32_t>(), kB.as<uint32_t>()};
    uint32_t *vbuf[2];
    if (passes % 2 == 1) {
        vbuf[0] = ids_out.as<uint32_t>();
        vbuf[1] = vA.as<uint32_t>();
    } else {
        vbuf[0] = vA.as<uint32_t>();
        vbuf[1] = ids_out.as<uint32_t>();
    }
    for (int p = 0; p < passes; ++p) {
        const int shift = 8 * p;
        hipLaunchKernelGGL(radix_hist, dim3(ntiles), dim3(kBlock), 0, s, kin, m, shift, ntiles,
                           hist.as<uint32_t>());
        GT_TRY(exclusive_scan_u32(hist.as<uint32_t>(), 256ull * ntiles, scanned.as<uint32_t>(), false,
                              s, tmp));
        hipLaunchKernelGGL(radix_scatter, dim3(ntiles), dim3(kBlock), 0, s, kin, vin, m, shift,
                           ntiles, scanned.as<uint32_t>(), kbuf[p & 1], vbuf[p & 1]);
        GT_HIP(hipGetLastError());
        kin = kbuf[p & 1];
        vin = vbuf[p & 1];
    }
    // The temporaries must outlive the queued kernels.
    GT_HIP(hipStreamSynchronize(s));
    return GT_OK;
}

// ---------------------------------------------------------------- replay

// One thread per vertex.  A = v's in-entries (edge ids, ascending), B = its
// out-entries (ascending).  Only out-inserts between the first and the last
// in-insert move entries; earlier ones meet an empty queue, later ones a full
// ring and only advance its head.
__global__ __launch_bounds__(kBlock) void push_swap_replay(
    const uint32_t *__restrict__ in_off, const uint32_t *__restrict__ pre,
    const uint32_t *__restrict__ out_off, const uint32_t *__restrict__ out_eid, uint32_t n,
    uint32_t *__restrict__ ring, uint32_t *__restrict__ in_eid) {
    for (uint32_t v = blockIdx.x * kBlock + threadIdx.x; v < n; v += gridDim.x * kBlock) {
        const uint32_t a0 = in_off[v], k = in_off[v + 1] - a0;
        if (k == 0) continue;
        const uint32_t *A = pre + a0;
        uint32_t *R = ring + a0;
        const uint32_t b0 = out_off[v], o = out_off[v + 1] - b0;
        const uint32_t *B = out_eid + b0;
        // Skip out-inserts up to and including A[0]'s id (queue still empty;
        // a self-loop's out-insert precedes its own in-insert).
        uint32_t lo = 0, hi = o;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (B[mid] <= A[0]) lo = mid + 1; else hi = mid;
        }
        uint32_t ib = lo, ia = 0, head = 0, size = 0;
        while (ia < k) {
            if (ib < o && B[ib] <= A[ia]) {
                // out-insert on a non-empty queue (size >= 1 here)
                if (size < k) {
                    uint32_t back = head + size;
                    if (back >= k) back -= k;
                    R[back] = R[head];
                }
                head = (head + 1 == k) ? 0 : head + 1;
                ++ib;
            } else {
                uint32_t back = head + size;
                if (back >= k) back -= k;
                R[back] = A[ia];
                ++size;
                ++ia;
            }
        }
        // Remaining out-inserts rotate a full ring.
        head = (uint32_t)(((uint64_t)head + (o - ib)) % k);
        for (uint32_t q = 0; q < k; ++q) {
            uint32_t p = head + q;
            if (p >= k) p -= k;
            in_eid[a0 + q] = R[p];
        }
    }
}

__global__ __launch_bounds__(kBlock) void gather_other(const uint32_t *__restrict__ eid,
                                                       const uint32_t *__restrict__ end, uint64_t m,
                                                       uint32_t *__restrict__ other) {
    for (uint64_t j = (uint64_t)blockIdx.x * kBlock + threadIdx.x; j < m;
         j += (uint64_t)gridDim.x * kBlock)
        other[j] = end[eid[j]];
}

static unsigned grid_of(uint64_t work) {
    uint64_t b = (work + kBlock - 1) / kBlock;
    if (b > 8192) b = 8192;
    return b ? (unsigned)b : 1u;
}

static int bits_for(uint32_t n) {
    int b = 0;
    while (b < 32 && (n - 1) >> b) ++b;
    return b == 0 ? 1 : b;
}

gt_status build_edges_impl(uint64_t graph_id, uint32_t n, uint64_t m, const uint32_t *src,
                           const uint32_t *tgt, int32_t loc, gt_graph **out) {
    clear_error();
    if (!out) return fail(GT_ERR_INVALID_ARG, "null output");
    *out = nullptr;
    if (m > 0xffffffffull) return fail(GT_ERR_INVALID_ARG, "edge count exceeds the index width");
    if (n == 0xffffffffu) return fail(GT_ERR_INVALID_ARG, "vertex count exceeds the index width");
    if (m && (!src || !tgt)) return fail(GT_ERR_INVALID_ARG, "null edge list");
    if (loc != GT_HOST && loc != GT_DEVICE) return fail(GT_ERR_INVALID_ARG, "bad location");
    if (m && n == 0) return fail(GT_ERR_NO_SUCH_VERTEX, "no such vertex: the graph has none");
    gt_graph *g = nullptr;
    GT_TRY(graph_new(graph_id, n, m, m, &g));
    struct Guard {
        gt_graph *g;
        bool keep = false;
        ~Guard() {
            if (!keep) gt_graph_free(g);
        }
    } guard{g};
    hipStream_t s = g->stream;
    std::deque<DevBuf> tmp;
    DevBuf dsrc, dtgt;
    const uint32_t *S = src, *T = tgt;
    if (loc == GT_HOST && m) {
        GT_TRY(dsrc.reserve(sizeof(uint32_t) * m));
        GT_TRY(dtgt.reserve(sizeof(uint32_t) * m));
        GT_HIP(hipMemcpyAsync(dsrc.ptr, src, sizeof(uint32_t) * m, hipMemcpyHostToDevice, s));
        GT_HIP(hipMemcpyAsync(dtgt.ptr, tgt, sizeof(uint32_t) * m, hipMemcpyHostToDevice, s));
        S = dsrc.as<uint32_t>();
        T = dtgt.as<uint32_t>();
    }
    Orient &I = g->in, &O = g->out;
    GT_TRY(I.off.reserve(sizeof(uint32_t) * ((size_t)n + 1)));
    GT_TRY(O.off.reserve(sizeof(uint32_t) * ((size_t)n + 1)));
    GT_TRY(I.idx.reserve(sizeof(uint32_t) * std::max<uint64_t>(1, m)));
    GT_TRY(I.eid.reserve(sizeof(uint32_t) * std::max<uint64_t>(1, m)));
    GT_TRY(O.idx.reserve(sizeof(uint32_t) * std::max<uint64_t>(1, m)));
    if (m) {
        DevBuf bad;
        GT_TRY(bad.reserve(sizeof(uint32_t)));
        GT_HIP(hipMemsetAsync(bad.ptr, 0, sizeof(uint32_t), s));
        hipLaunchKernelGGL(check_endpoints, dim3(grid_of(m)), dim3(kBlock), 0, s, S, T, m, n,
                           bad.as<uint32_t>());
        GT_HIP(hipGetLastError());
        uint32_t hbad = 0;
        GT_HIP(hipMemcpyAsync(&hbad, bad.ptr, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        GT_HIP(hipStreamSynchronize(s));
        if (hbad) return fail(GT_ERR_NO_SUCH_VERTEX, "no such vertex: an endpoint is >= n");
    }
    // 1-2: degrees and offsets
    DevBuf outd, ind;
    GT_TRY(outd.reserve(sizeof(uint32_t) * std::max<uint32_t>(1, n)));
    GT_TRY(ind.reserve(sizeof(uint32_t) * std::max<uint32_t>(1, n)));
    GT_HIP(hipMemsetAsync(outd.ptr, 0, sizeof(uint32_t) * n, s));
    GT_HIP(hipMemsetAsync(ind.ptr, 0, sizeof(uint32_t) * n, s));
    if (m)
        hipLaunchKernelGGL(degree_hist, dim3(grid_of(m)), dim3(kBlock), 0, s, S, T, m,
                           outd.as<uint32_t>(), ind.as<uint32_t>());
    GT_HIP(hipGetLastError());
    GT_TRY(exclusive_scan_u32(outd.as<uint32_t>(), n, O.off.as<uint32_t>(), true, s, tmp));
    GT_TRY(exclusive_scan_u32(ind.as<uint32_t>(), n, I.off.as<uint32_t>(), true, s, tmp));
    if (m) {
        // 3: stable grouping by source (final out-halves) and by target.
        const int bits = bits_for(n);
        GT_TRY(stable_sort_ids(S, nullptr, m, bits, O.eid, s, tmp));
        DevBuf pre;
        GT_TRY(stable_sort_ids(T, nullptr, m, bits, pre, s, tmp));
        // 4: push-swap replay into the in-halves.
        DevBuf ring;
        GT_TRY(ring.reserve(sizeof(uint32_t) * m));
        hipLaunchKernelGGL(push_swap_replay, dim3(grid_of(n)), dim3(kBlock), 0, s,
                           I.off.as<uint32_t>(), pre.as<uint32_t>(), O.off.as<uint32_t>(),
                           O.eid.as<uint32_t>(), n, ring.as<uint32_t>(), I.eid.as<uint32_t>());
        GT_HIP(hipGetLastError());
        // 5: the other endpoint of every entry.
        hipLaunchKernelGGL(gather_other, dim3(grid_of(m)), dim3(kBlock), 0, s,
                           I.eid.as<uint32_t>(), S, m, I.idx.as<uint32_t>());
        hipLaunchKernelGGL(gather_other, dim3(grid_of(m)), dim3(kBlock), 0, s,
                           O.eid.as<uint32_t>(), T, m, O.idx.as<uint32_t>());
        GT_HIP(hipGetLastError());
        GT_HIP(hipStreamSynchronize(s));
    } else {
        GT_TRY(O.eid.reserve(sizeof(uint32_t)));
    }
    // Host copies of the offsets for the SpMV schedule.
    I.h_off.resize((size_t)n + 1);
    O.h_off.resize((size_t)n + 1);
    GT_HIP(hipMemcpyAsync(I.h_off.data(), I.off.ptr, sizeof(uint32_t) * ((size_t)n + 1),
                          hipMemcpyDeviceToHost, s));
    GT_HIP(hipMemcpyAsync(O.h_off.data(), O.off.ptr, sizeof(uint32_t) * ((size_t)n + 1),
                          hipMemcpyDeviceToHost, s));
    GT_HIP(hipStreamSynchronize(s));
    guard.keep = true;
    *out = g;
    return GT_OK;
}

}  // namespace gt

extern "C" gt_status gt_graph_build_edges(uint64_t graph_id, uint32_t n, uint64_t m,
                                          const uint32_t *src, const uint32_t *tgt, int32_t loc,
                                          gt_graph **out) {
    return gt::build_edges_impl(graph_id, n, m, src, tgt, loc, out);
}
