// Locality-friendly relabelling of a pull orientation (built once per graph
// and orientation, integer work only).
//
// The sweeps' cost is the gather of the source vector, one random read per
// in-entry.  Renumbering vertices by descending gather frequency packs the
// frequently read entries of that vector into few cache lines (on R-MAT the
// top few percent of vertices carry most in-entries), and sorting each row's
// entries by the new label makes a wavefront's gathers ascend.  The reference
// sums a row in its storage order; this order differs only in rounding.
//
//   iperm  = vertex ids sorted stably by descending frequency (LSD radix sort
//            on the complement of the frequency)
//   perm   = inverse of iperm
//   entries: stable sort by new entry label, then stable sort by new row
//            label (LSD on the pair), starting from storage order
#include <algorithm>
#include <deque>

#include "device_common.hpp"
#include "gt_internal.hpp"

namespace gt {

gt_status stable_sort_ids(const uint32_t *key, const uint32_t *vals, uint64_t m, int bits,
                          DevBuf &ids_out, hipStream_t s, std::deque<DevBuf> &tmp);
gt_status exclusive_scan_u32(const uint32_t *in, uint64_t len, uint32_t *out, bool with_total,
                             hipStream_t s, std::deque<DevBuf> &tmp);

__global__ __launch_bounds__(kBlock) void freq_key(const uint32_t *__restrict__ off, uint32_t n,
                                                   uint32_t *__restrict__ key) {
    for (uint32_t v = blockIdx.x * kBlock + threadIdx.x; v < n; v += gridDim.x * kBlock)
        key[v] = 0xffffffffu - (off[v + 1] - off[v]);
}

__global__ __launch_bounds__(kBlock) void invert_perm(const uint32_t *__restrict__ iperm, uint32_t n,
                                                      uint32_t *__restrict__ perm) {
    for (uint32_t k = blockIdx.x * kBlock + threadIdx.x; k < n; k += gridDim.x * kBlock)
        perm[iperm[k]] = k;
}

// One wavefront per row: row id of each entry.
__global__ __launch_bounds__(kBlock) void expand_rows(const uint32_t *__restrict__ off, uint32_t n,
                                                     uint32_t *__restrict__ row) {
    const int lane = threadIdx.x & (kWave - 1);
    const uint32_t waves = gridDim.x * (kBlock / kWave);
    for (uint32_t r = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave; r < n; r += waves)
        for (uint32_t j = off[r] + lane; j < off[r + 1]; j += kWave) row[j] = r;
}

// out[j] = map[in[j]]
__global__ __launch_bounds__(kBlock) void map_u32(const uint32_t *__restrict__ in,
                                                  const uint32_t *__restrict__ map, uint64_t m,
                                                  uint32_t *__restrict__ out) {
    for (uint64_t j = (uint64_t)blockIdx.x * kBlock + threadIdx.x; j < m;
         j += (uint64_t)gridDim.x * kBlock)
        out[j] = map[in[j]];
}

// out[j] = map[in[pos[j]]]  and  eout[j] = eid[pos[j]]
__global__ __launch_bounds__(kBlock) void place_entries(const uint32_t *__restrict__ pos,
                                                        const uint32_t *__restrict__ idx,
                                                        const uint32_t *__restrict__ eid,
                                                        const uint32_t *__restrict__ map,
                                                        uint64_t m, uint32_t *__restrict__ iout,
                                                        uint32_t *__restrict__ eout) {
    for (uint64_t j = (uint64_t)blockIdx.x * kBlock + threadIdx.x; j < m;
         j += (uint64_t)gridDim.x * kBlock) {
        const uint32_t p = pos[j];
        iout[j] = map[idx[p]];
        eout[j] = eid[p];
    }
}

// len[k] = row length of old row iperm[k]
__global__ __launch_bounds__(kBlock) void new_lengths(const uint32_t *__restrict__ off,
                                                      const uint32_t *__restrict__ iperm,
                                                      uint32_t n, uint32_t *__restrict__ len) {
    for (uint32_t k = blockIdx.x * kBlock + threadIdx.x; k < n; k += gridDim.x * kBlock) {
        const uint32_t v = iperm[k];
        len[k] = off[v + 1] - off[v];
    }
}

static unsigned ggrid(uint64_t work) {
    uint64_t b = (work + kBlock - 1) / kBlock;
    if (b > 8192) b = 8192;
    return b ? (unsigned)b : 1u;
}

static int bits_of(uint32_t n) {
    int b = 0;
    while (b < 32 && ((uint64_t)(n - 1) >> b)) ++b;
    return b == 0 ? 1 : b;
}

gt_status ensure_relabeled(gt_graph *g, int which, Relabeled **out) {
    Relabeled &R = g->rel[which];
    *out = &R;
    if (R.ready) return GT_OK;
    const uint32_t n = g->n;
    const uint64_t m = g->m;
    Orient &pull = which ? g->out : g->in;
    Orient &push = which ? g->in : g->out;
    hipStream_t s = g->stream;
    std::deque<DevBuf> tmp;
    GT_TRY(R.perm.reserve(sizeof(uint32_t) * std::max<uint32_t>(1, n)));
    GT_TRY(R.pull.off.reserve(sizeof(uint32_t) * ((size_t)n + 1)));
    GT_TRY(R.push_off.reserve(sizeof(uint32_t) * ((size_t)n + 1)));
    GT_TRY(R.pull.idx.reserve(sizeof(uint32_t) * std::max<uint64_t>(1, m)));
    GT_TRY(R.pull.eid.reserve(sizeof(uint32_t) * std::max<uint64_t>(1, m)));
    if (n == 0) {
        GT_HIP(hipMemsetAsync(R.pull.off.ptr, 0, sizeof(uint32_t), s));
        GT_HIP(hipMemsetAsync(R.push_off.ptr, 0, sizeof(uint32_t), s));
        R.pull.h_off.assign(1, 0);
        GT_HIP(hipStreamSynchronize(s));
        R.ready = true;
        return GT_OK;
    }
    // Vertex order: descending frequency (push-row length), ties by id.
    {
        tmp.emplace_back();
        DevBuf &key = tmp.back();
        GT_TRY(key.reserve(sizeof(uint32_t) * n));
        hipLaunchKernelGGL(freq_key, dim3(ggrid(n)), dim3(kBlock), 0, s, push.off.as<uint32_t>(),
                           n, key.as<uint32_t>());
        GT_TRY(stable_sort_ids(key.as<uint32_t>(), nullptr, n, 32, R.iperm, s, tmp));
        hipLaunchKernelGGL(invert_perm, dim3(ggrid(n)), dim3(kBlock), 0, s,
                           R.iperm.as<uint32_t>(), n, R.perm.as<uint32_t>());
        GT_HIP(hipGetLastError());
    }
    // Offsets in the new order.
    {
        tmp.emplace_back();
        DevBuf &len = tmp.back();
        GT_TRY(len.reserve(sizeof(uint32_t) * n));
        hipLaunchKernelGGL(new_lengths, dim3(ggrid(n)), dim3(kBlock), 0, s,
                           pull.off.as<uint32_t>(), R.iperm.as<uint32_t>(), n, len.as<uint32_t>());
        GT_TRY(exclusive_scan_u32(len.as<uint32_t>(), n, R.pull.off.as<uint32_t>(), true, s, tmp));
        hipLaunchKernelGGL(new_lengths, dim3(ggrid(n)), dim3(kBlock), 0, s,
                           push.off.as<uint32_t>(), R.iperm.as<uint32_t>(), n, len.as<uint32_t>());
        GT_TRY(exclusive_scan_u32(len.as<uint32_t>(), n, R.push_off.as<uint32_t>(), true, s, tmp));
        GT_HIP(hipGetLastError());
    }
    if (m) {
        const int bits = bits_of(n);
        DevBuf row, k1, k2, ids1, ids2;
        GT_TRY(row.reserve(sizeof(uint32_t) * m));
        GT_TRY(k1.reserve(sizeof(uint32_t) * m));
        GT_TRY(k2.reserve(sizeof(uint32_t) * m));
        hipLaunchKernelGGL(expand_rows, dim3(ggrid((uint64_t)n * kWave)), dim3(kBlock), 0, s,
                           pull.off.as<uint32_t>(), n, row.as<uint32_t>());
        // Secondary key first: the new label of each entry.
        hipLaunchKernelGGL(map_u32, dim3(ggrid(m)), dim3(kBlock), 0, s, pull.idx.as<uint32_t>(),
                           R.perm.as<uint32_t>(), m, k1.as<uint32_t>());
        GT_HIP(hipGetLastError());
        GT_TRY(stable_sort_ids(k1.as<uint32_t>(), nullptr, m, bits, ids1, s, tmp));
        // Primary key: the new label of each entry's row, in that order.
        hipLaunchKernelGGL(map_u32, dim3(ggrid(m)), dim3(kBlock), 0, s, ids1.as<uint32_t>(),
                           row.as<uint32_t>(), m, k2.as<uint32_t>());
        hipLaunchKernelGGL(map_u32, dim3(ggrid(m)), dim3(kBlock), 0, s, k2.as<uint32_t>(),
                           R.perm.as<uint32_t>(), m, k1.as<uint32_t>());
        GT_HIP(hipGetLastError());
        GT_TRY(stable_sort_ids(k1.as<uint32_t>(), ids1.as<uint32_t>(), m, bits, ids2, s, tmp));
        hipLaunchKernelGGL(place_entries, dim3(ggrid(m)), dim3(kBlock), 0, s,
                           ids2.as<uint32_t>(), pull.idx.as<uint32_t>(), pull.eid.as<uint32_t>(),
                           R.perm.as<uint32_t>(), m, R.pull.idx.as<uint32_t>(),
                           R.pull.eid.as<uint32_t>());
        GT_HIP(hipGetLastError());
        GT_HIP(hipStreamSynchronize(s));
    }
    R.pull.h_off.resize((size_t)n + 1);
    GT_HIP(hipMemcpyAsync(R.pull.h_off.data(), R.pull.off.ptr, sizeof(uint32_t) * ((size_t)n + 1),
                          hipMemcpyDeviceToHost, s));
    GT_HIP(hipStreamSynchronize(s));
    R.pull.scheduled = false;
    R.pull.mm.ready = false;
    R.ready = true;
    return GT_OK;
}

}  // namespace gt
