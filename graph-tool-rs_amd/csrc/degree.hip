// Degree kernels on the device: crates/gt-algo/src/degree.rs:122-237.
//
//   weighted_out_degree / weighted_in_degree / weighted_degree (:122-185):
//     Unity      -> the degree itself;
//     Constant c -> degree * c, one rounding, 0 when there is no incident edge
//                   (constant_weight, :85-100);
//     otherwise  -> the sum of w over the neighbourhood in storage order
//                   (sum_weights, :60-75); the total degree walks the block
//                   as the out-half followed by the in-half, so its sum
//                   continues the out-half's (a self-loop counts twice).
//   out_degrees (:199-216): an int64 map over the whole vertex bound, cleared,
//     then the view's out-degrees.
//   degree_histogram (:230-247): hist[d] = number of the view's vertices of
//     total degree d, unit bins from 0, length max degree + 1.
//
// Views: Directed, Reversed (halves swapped), the filtered kinds (compacted
// orientations; vertices the filter drops read 0 and are not counted by the
// histogram), and Undirected, whose out-edges and all-edges are both the
// whole block (view/undirected.rs:106-116): its out-degree is the total
// degree and an in-degree is refused.  The integer degrees and the histogram
// are exact; weighted sums follow the storage order except in rows of more
// than 4096 entries (a fixed tree order there).
#include <algorithm>
#include <string>

#include "device_common.hpp"
#include "gt_internal.hpp"
#include "props.hpp"

namespace gt {

enum { kDegOut = 0, kDegIn = 1, kDegTotal = 2 };

// d(v) = |row v of a| + |row v of b| (b may be null); inactive vertices 0.
__device__ __forceinline__ uint32_t row_deg(const uint32_t *a, const uint32_t *b,
                                            const uint8_t *active, uint32_t v) {
    if (active && !active[v]) return 0;
    uint32_t d = a[v + 1] - a[v];
    if (b) d += b[v + 1] - b[v];
    return d;
}

// Unity (constant == false) or Constant c (constant == true) degrees as f64.
__global__ __launch_bounds__(kBlock) void deg_plain(const uint32_t *a, const uint32_t *b,
                                                    const uint8_t *active, uint32_t n,
                                                    bool constant, double c, double *out) {
    for (uint32_t v = blockIdx.x * kBlock + threadIdx.x; v < n; v += gridDim.x * kBlock) {
        const uint32_t d = row_deg(a, b, active, v);
        out[v] = constant ? (d ? (double)d * c : 0.0) : (double)d;
    }
}

__global__ __launch_bounds__(kBlock) void deg_int64(const uint32_t *a, const uint32_t *b,
                                                    const uint8_t *active, uint32_t n,
                                                    int64_t *out) {
    for (uint32_t v = blockIdx.x * kBlock + threadIdx.x; v < n; v += gridDim.x * kBlock)
        out[v] = (int64_t)row_deg(a, b, active, v);
}

// Inactive vertices of a filtered view read 0 (their compacted rows are
// already empty; this also clears what an accumulation left).
__global__ __launch_bounds__(kBlock) void deg_mask(const uint8_t *active, uint32_t n,
                                                   double *out) {
    for (uint32_t v = blockIdx.x * kBlock + threadIdx.x; v < n; v += gridDim.x * kBlock)
        if (!active[v]) out[v] = 0.0;
}

__global__ __launch_bounds__(kBlock) void deg_max(const uint32_t *a, const uint32_t *b,
                                                  const uint8_t *active, uint32_t n,
                                                  uint32_t *mx) {
    uint32_t m = 0;
    for (uint32_t v = blockIdx.x * kBlock + threadIdx.x; v < n; v += gridDim.x * kBlock)
        if (!active || active[v]) m = max(m, row_deg(a, b, nullptr, v));
    if (m) atomicMax(mx, m);
}

__global__ __launch_bounds__(kBlock) void deg_hist(const uint32_t *a, const uint32_t *b,
                                                   const uint8_t *active, uint32_t n,
                                                   unsigned long long *hist) {
    for (uint32_t v = blockIdx.x * kBlock + threadIdx.x; v < n; v += gridDim.x * kBlock)
        if (!active || active[v]) atomicAdd(&hist[row_deg(a, b, nullptr, v)], 1ull);
}

namespace {

unsigned dgrid(uint32_t n) {
    uint32_t b = (n + kBlock - 1) / kBlock;
    if (b > 4096) b = 4096;
    return b ? b : 1u;
}

// The orientations whose rows are the view's out- and in-halves.  For an
// undirected view `in` is null and `out` covers the whole block as the pair
// (out, extra).
struct DegView {
    Orient *out = nullptr, *in = nullptr, *extra = nullptr;
    const uint8_t *active = nullptr;
    uint32_t nactive = 0;
};

gt_status degree_view(gt_graph *g, const gt_view *view, bool need_in, DegView *dv) {
    const int kind = view ? view->kind : GT_VIEW_DIRECTED;
    if (kind == GT_VIEW_UNDIRECTED) {
        if (need_in) {
            Orient *p, *q;
            bool r;
            const uint8_t *act;
            uint32_t na;
            return resolve_view(g, view, &p, &q, &r, &act, &na);  // the NotDirected error
        }
        dv->out = &g->out;
        dv->extra = &g->in;
        dv->nactive = g->n;
        return GT_OK;
    }
    if (kind == GT_VIEW_UNDIRECTED_FILTERED)
        return fail(GT_ERR_UNSUPPORTED, "degrees of an undirected filtered view");
    Orient *pull = nullptr, *push = nullptr;
    bool reversed = false;
    GT_TRY(resolve_view(g, view, &pull, &push, &reversed, &dv->active, &dv->nactive));
    dv->out = push;
    dv->in = pull;
    return GT_OK;
}

}  // namespace

gt_status degrees_impl(gt_graph *g, const gt_view *view, int32_t which, const gt_prop *weight,
                       double *out, int32_t out_loc) {
    clear_error();
    if (!g) return fail(GT_ERR_INVALID_ARG, "null graph");
    if (which < kDegOut || which > kDegTotal) return fail(GT_ERR_INVALID_ARG, "unknown degree kind");
    if (!out && g->n) return fail(GT_ERR_INVALID_ARG, "null output");
    if (out_loc != GT_HOST && out_loc != GT_DEVICE) return fail(GT_ERR_INVALID_ARG, "bad location");
    DegView dv;
    GT_TRY(degree_view(g, view, which == kDegIn, &dv));
    const uint32_t n = g->n;
    if (n == 0) return GT_OK;
    hipStream_t s = g->stream;
    Workspace &ws = g->ws;
    double *dst = out;
    if (out_loc == GT_HOST) {
        GT_TRY(ws.y3.reserve(sizeof(double) * n));
        dst = ws.y3.as<double>();
    }
    // The halves walked, in block order.
    Orient *first = nullptr, *second = nullptr;
    if (dv.extra) {
        first = dv.out;
        second = dv.extra;  // undirected: out- and all-edges are the whole block
    } else if (which == kDegOut) {
        first = dv.out;
    } else if (which == kDegIn) {
        first = dv.in;
    } else {
        first = dv.out;
        second = dv.in;
    }
    const int kind = weight ? weight->kind : GT_MAP_UNITY;
    if (kind == GT_MAP_UNITY || kind == GT_MAP_CONSTANT) {
        hipLaunchKernelGGL(deg_plain, dim3(dgrid(n)), dim3(kBlock), 0, s,
                           first->off.as<uint32_t>(),
                           second ? second->off.as<uint32_t>() : nullptr, dv.active, n,
                           kind == GT_MAP_CONSTANT, weight ? weight->value : 1.0, dst);
    } else {
        const void *src = nullptr;
        int dt = GT_F64;
        GT_TRY(stage_edge_source(g, weight, &src, &dt));
        GT_TRY(stage_degree_into(g, weight, *first, src, dt, dst, false));
        if (second) GT_TRY(stage_degree_into(g, weight, *second, src, dt, dst, true));
        if (dv.active)
            hipLaunchKernelGGL(deg_mask, dim3(dgrid(n)), dim3(kBlock), 0, s, dv.active, n, dst);
    }
    GT_HIP(hipGetLastError());
    if (out_loc == GT_HOST)
        GT_HIP(hipMemcpyAsync(out, dst, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    GT_HIP(hipStreamSynchronize(s));
    return GT_OK;
}

gt_status out_degrees_impl(gt_graph *g, const gt_view *view, int64_t *out, int32_t out_loc) {
    clear_error();
    if (!g) return fail(GT_ERR_INVALID_ARG, "null graph");
    if (!out && g->n) return fail(GT_ERR_INVALID_ARG, "null output");
    if (out_loc != GT_HOST && out_loc != GT_DEVICE) return fail(GT_ERR_INVALID_ARG, "bad location");
    DegView dv;
    GT_TRY(degree_view(g, view, false, &dv));
    const uint32_t n = g->n;
    if (n == 0) return GT_OK;
    hipStream_t s = g->stream;
    int64_t *dst = out;
    if (out_loc == GT_HOST) {
        GT_TRY(g->ws.y3.reserve(sizeof(int64_t) * n));
        dst = g->ws.y3.as<int64_t>();
    }
    // An undirected view's out-degree is the whole block (both halves).
    hipLaunchKernelGGL(deg_int64, dim3(dgrid(n)), dim3(kBlock), 0, s, dv.out->off.as<uint32_t>(),
                       dv.extra ? dv.extra->off.as<uint32_t>() : nullptr, dv.active, n, dst);
    GT_HIP(hipGetLastError());
    if (out_loc == GT_HOST)
        GT_HIP(hipMemcpyAsync(out, dst, sizeof(int64_t) * n, hipMemcpyDeviceToHost, s));
    GT_HIP(hipStreamSynchronize(s));
    return GT_OK;
}

gt_status degree_histogram_impl(gt_graph *g, const gt_view *view, uint64_t *hist, uint64_t cap,
                                uint64_t *len) {
    clear_error();
    if (!g || !len) return fail(GT_ERR_INVALID_ARG, "null argument");
    DegView dv;
    GT_TRY(degree_view(g, view, false, &dv));
    const uint32_t n = g->n;
    *len = 0;
    if (n == 0 || dv.nactive == 0) return GT_OK;
    hipStream_t s = g->stream;
    const uint32_t *a = dv.out->off.as<uint32_t>();
    const uint32_t *b = (dv.extra ? dv.extra : dv.in)->off.as<uint32_t>();
    Workspace &ws = g->ws;
    GT_TRY(ws.state.reserve(std::max<size_t>(sizeof(SweepState), 64)));
    uint32_t *mx = ws.state.as<uint32_t>();
    GT_HIP(hipMemsetAsync(mx, 0, sizeof(uint32_t), s));
    hipLaunchKernelGGL(deg_max, dim3(dgrid(n)), dim3(kBlock), 0, s, a, b, dv.active, n, mx);
    GT_HIP(hipGetLastError());
    uint32_t hm = 0;
    GT_HIP(hipMemcpyAsync(&g->h_state->counter, mx, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    GT_HIP(hipStreamSynchronize(s));
    hm = g->h_state->counter;
    *len = (uint64_t)hm + 1;
    if (!hist || cap < *len) return GT_OK;
    GT_TRY(ws.y3.reserve(sizeof(unsigned long long) * (*len)));
    unsigned long long *dh = ws.y3.as<unsigned long long>();
    GT_HIP(hipMemsetAsync(dh, 0, sizeof(unsigned long long) * (*len), s));
    hipLaunchKernelGGL(deg_hist, dim3(dgrid(n)), dim3(kBlock), 0, s, a, b, dv.active, n, dh);
    GT_HIP(hipGetLastError());
    GT_HIP(hipMemcpyAsync(hist, dh, sizeof(uint64_t) * (*len), hipMemcpyDeviceToHost, s));
    GT_HIP(hipStreamSynchronize(s));
    return GT_OK;
}

}  // namespace gt

using namespace gt;

extern "C" {

gt_status gt_degrees(gt_graph *g, const gt_view *view, int32_t which, const gt_prop *weight,
                     double *out, int32_t out_loc) {
    return degrees_impl(g, view, which, weight, out, out_loc);
}

gt_status gt_out_degrees(gt_graph *g, const gt_view *view, int64_t *out, int32_t out_loc) {
    return out_degrees_impl(g, view, out, out_loc);
}

gt_status gt_degree_histogram(gt_graph *g, const gt_view *view, uint64_t *hist, uint64_t cap,
                              uint64_t *len) {
    return degree_histogram_impl(g, view, hist, cap, len);
}

}  // extern "C"
