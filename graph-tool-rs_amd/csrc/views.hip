// Mask-filtered views on the device (Filtered<&AdjList, MaskFilter> and
// Filtered<Rev<&AdjList>, MaskFilter>, crates/gt-core/src/view/filtered.rs).
//
// An incidence of a kept vertex survives iff keep_edge(e) && keep_vertex(other)
// (keeps_incident, filtered.rs:122-124).  Per call, both orientations are
// compacted to their surviving entries, in storage order; rows of vertices
// the filter drops are emptied (the sweeps carry those slots unchanged, as
// centrality.rs:345-351 does).  The filtered out-degree is then the
// compacted push orientation's row length.
#include <algorithm>
#include <deque>

#include "device_common.hpp"
#include "gt_internal.hpp"

namespace gt {

gt_status exclusive_scan_u32(const uint32_t *in, uint64_t len, uint32_t *out, bool with_total,
                             hipStream_t s, std::deque<DevBuf> &tmp);

__device__ __forceinline__ bool keeps(const uint8_t *vmask, const uint8_t *emask, uint32_t other,
                                      uint32_t e) {
    return emask[e] != 0 && vmask[other] != 0;
}

__global__ __launch_bounds__(kBlock) void count_kept(const uint32_t *__restrict__ off,
                                                     const uint32_t *__restrict__ idx,
                                                     const uint32_t *__restrict__ eid,
                                                     const uint8_t *__restrict__ vmask,
                                                     const uint8_t *__restrict__ emask, uint32_t n,
                                                     uint32_t *__restrict__ cnt) {
    for (uint32_t v = blockIdx.x * kBlock + threadIdx.x; v < n; v += gridDim.x * kBlock) {
        uint32_t c = 0;
        if (vmask[v])
            for (uint32_t j = off[v]; j < off[v + 1]; ++j) c += keeps(vmask, emask, idx[j], eid[j]);
        cnt[v] = c;
    }
}

__global__ __launch_bounds__(kBlock) void scatter_kept(
    const uint32_t *__restrict__ off, const uint32_t *__restrict__ idx,
    const uint32_t *__restrict__ eid, const uint8_t *__restrict__ vmask,
    const uint8_t *__restrict__ emask, uint32_t n, const uint32_t *__restrict__ noff,
    uint32_t *__restrict__ nidx, uint32_t *__restrict__ neid) {
    for (uint32_t v = blockIdx.x * kBlock + threadIdx.x; v < n; v += gridDim.x * kBlock) {
        if (!vmask[v]) continue;
        uint32_t w = noff[v];
        for (uint32_t j = off[v]; j < off[v + 1]; ++j) {
            if (keeps(vmask, emask, idx[j], eid[j])) {
                nidx[w] = idx[j];
                neid[w] = eid[j];
                ++w;
            }
        }
    }
}

static unsigned rows_grid(uint32_t n) {
    uint32_t b = (n + kBlock - 1) / kBlock;
    if (b > 4096) b = 4096;
    return b ? b : 1u;
}

static gt_status compact(gt_graph *g, const Orient &base, Orient &out, const uint8_t *vmask,
                         const uint8_t *emask, std::deque<DevBuf> &tmp) {
    const uint32_t n = g->n;
    hipStream_t s = g->stream;
    tmp.emplace_back();
    DevBuf &cnt = tmp.back();
    GT_TRY(cnt.reserve(sizeof(uint32_t) * std::max<uint32_t>(1, n)));
    GT_TRY(out.off.reserve(sizeof(uint32_t) * ((size_t)n + 1)));
    hipLaunchKernelGGL(count_kept, dim3(rows_grid(n)), dim3(kBlock), 0, s,
                       base.off.as<uint32_t>(), base.idx.as<uint32_t>(), base.eid.as<uint32_t>(),
                       vmask, emask, n, cnt.as<uint32_t>());
    GT_HIP(hipGetLastError());
    GT_TRY(exclusive_scan_u32(cnt.as<uint32_t>(), n, out.off.as<uint32_t>(), true, s, tmp));
    out.h_off.resize((size_t)n + 1);
    GT_HIP(hipMemcpyAsync(out.h_off.data(), out.off.ptr, sizeof(uint32_t) * ((size_t)n + 1),
                          hipMemcpyDeviceToHost, s));
    GT_HIP(hipStreamSynchronize(s));
    const uint64_t mk = out.h_off[n];
    GT_TRY(out.idx.reserve(sizeof(uint32_t) * std::max<uint64_t>(1, mk)));
    GT_TRY(out.eid.reserve(sizeof(uint32_t) * std::max<uint64_t>(1, mk)));
    hipLaunchKernelGGL(scatter_kept, dim3(rows_grid(n)), dim3(kBlock), 0, s,
                       base.off.as<uint32_t>(), base.idx.as<uint32_t>(), base.eid.as<uint32_t>(),
                       vmask, emask, n, out.off.as<uint32_t>(), out.idx.as<uint32_t>(),
                       out.eid.as<uint32_t>());
    GT_HIP(hipGetLastError());
    out.scheduled = false;
    out.mm.ready = false;
    out.heavy_ready = false;
    out.bcs.ready = false;
    return GT_OK;
}

// Resolve a filtered view: stage the masks, compact both orientations into
// the graph's filtered workspace, count the kept vertices.
gt_status filtered_view(gt_graph *g, const gt_view *view, bool reversed, Orient **pull,
                        Orient **push, const uint8_t **active, uint32_t *nactive) {
    const uint32_t n = g->n;
    const uint64_t eb = g->edge_bound;
    if (!view->vmask || !view->emask) return fail(GT_ERR_INVALID_ARG, "filter: null mask");
    if (view->vmask_len < n) return fail(GT_ERR_INVALID_ARG, "filter: vertex mask is shorter than the vertex bound");
    if (view->emask_len < eb) return fail(GT_ERR_INVALID_ARG, "filter: edge mask is shorter than the edge bound");
    hipStream_t s = g->stream;
    Workspace &ws = g->ws;
    GT_TRY(ws.masks.reserve((size_t)n + eb + 16));
    uint8_t *vm = ws.masks.as<uint8_t>();
    uint8_t *em = vm + n;
    if (n) GT_HIP(hipMemcpyAsync(vm, view->vmask, n, hipMemcpyHostToDevice, s));
    if (eb) GT_HIP(hipMemcpyAsync(em, view->emask, eb, hipMemcpyHostToDevice, s));
    uint32_t cnt = 0;
    for (uint32_t v = 0; v < n; ++v) cnt += view->vmask[v] != 0;
    std::deque<DevBuf> tmp;
    const Orient &bpull = reversed ? g->out : g->in;
    const Orient &bpush = reversed ? g->in : g->out;
    GT_TRY(compact(g, bpull, ws.fpull, vm, em, tmp));
    GT_TRY(compact(g, bpush, ws.fpush, vm, em, tmp));
    GT_HIP(hipStreamSynchronize(s));
    *pull = &ws.fpull;
    *push = &ws.fpush;
    *active = vm;
    *nactive = cnt;
    return GT_OK;
}

}  // namespace gt
