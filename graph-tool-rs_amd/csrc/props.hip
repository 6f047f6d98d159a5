// Staging of the property maps a sweep reads: edge weights permuted into the
// pull orientation's storage order, the weighted out-degree, and vertex maps
// (personalisation) as f64 device arrays.
//
// The weighted out-degree restates crates/gt-algo/src/centrality.rs:281-297:
// deg(v) = sum of w(e) over v's out-entries in storage order, starting from
// 0.0.  Rows up to kSeqDeg entries are summed in that order -- rows shorter
// than kWaveDeg by one thread, longer ones by a wavefront that loads 64
// weights at a time and folds them lane by lane -- so the value is the
// reference's bit for bit; longer rows use a fixed shuffle-tree order
// (deterministic, within rounding of the sequential sum).
#include "device_common.hpp"
#include "gt_internal.hpp"
#include "props.hpp"

namespace gt {

constexpr uint32_t kWaveDeg = 32;
constexpr uint32_t kSeqDeg = 4096;

template <typename S>
__device__ __forceinline__ double read_w(const void *w, uint32_t e, int kind, double value) {
    if (kind == GT_MAP_CONSTANT) return value;
    return static_cast<double>(static_cast<const S *>(w)[e]);
}

// One thread per row shorter than kWaveDeg, added in storage order.  The
// row's entries go in batches of eight, every load of a batch issued before
// the first add (a short row's tail is one batch with its lanes past the row
// masked), so a row costs one round trip per eight entries.
template <typename S>
__global__ __launch_bounds__(kBlock) void deg_short(const uint32_t *__restrict__ off,
                                                    const uint32_t *__restrict__ eid,
                                                    const void *w, int kind, double value,
                                                    uint32_t n, bool accumulate,
                                                    double *__restrict__ deg) {
    // Each XCD takes a contiguous eighth of the rows, so the weights its rows
    // read stay in its L2 (edge ids of nearby rows are near each other when
    // edges are numbered graph by graph or row by row).  (Non-temporal
    // edge-id loads measured slower here: 2.66 against 2.18 ms at 312 M
    // entries.)
    const XcdSpan sp = xcd_span(n);
    for (uint32_t v = sp.begin + sp.index * kBlock + threadIdx.x; v < sp.begin + sp.count;
         v += sp.groups * kBlock) {
        const uint32_t lo = off[v], hi = off[v + 1];
        if (hi - lo >= kWaveDeg) continue;
        double acc = accumulate ? deg[v] : 0.0;
        for (uint32_t j = lo; j < hi; j += 8) {
            uint32_t e[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) e[u] = j + u < hi ? eid[j + u] : 0u;
            double x[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) x[u] = j + u < hi ? read_w<S>(w, e[u], kind, value) : 0.0;
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (j + u < hi) acc += x[u];
        }
        deg[v] = acc;
    }
}

// One wavefront per listed row (kWaveDeg <= length <= kSeqDeg): 64 weights
// loaded at once, then added in storage order by a lane-by-lane fold that
// every lane repeats.
template <typename S>
__global__ __launch_bounds__(kBlock) void deg_mid(const uint32_t *__restrict__ off,
                                                  const uint32_t *__restrict__ eid,
                                                  const void *w, int kind, double value,
                                                  const uint32_t *__restrict__ rows,
                                                  uint32_t nrows, bool accumulate,
                                                  double *__restrict__ deg) {
    const int lane = threadIdx.x & (kWave - 1);
    // (Contiguous eighths of the row list per XCD, as deg_short walks its
    // rows, measured slower at R-MAT 22: 1.34 against 1.01 ms; the list's
    // rows vary in length, so the eighths are uneven work.)
    const uint32_t waves = gridDim.x * (kBlock / kWave);
    for (uint32_t k = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave; k < nrows;
         k += waves) {
        const uint32_t v = rows[k];
        const uint32_t lo = off[v], hi = off[v + 1];
        double acc = accumulate ? deg[v] : 0.0;
        // Storage order, 64 weights at a time: the next 64 are loaded before
        // this 64 is folded, and the fold reads each lane's weight with
        // readlane (no LDS round trip in the chain of adds).
        double x =
            lo + lane < hi ? read_w<S>(w, __builtin_nontemporal_load(eid + lo + lane), kind, value)
                           : 0.0;
        for (uint32_t j0 = lo; j0 < hi; j0 += kWave) {
            const uint32_t jn = j0 + kWave + lane;
            const double xn =
                jn < hi ? read_w<S>(w, __builtin_nontemporal_load(eid + jn), kind, value) : 0.0;
            const uint32_t cnt = hi - j0 < (uint32_t)kWave ? hi - j0 : (uint32_t)kWave;
            const uint64_t xb = __builtin_bit_cast(uint64_t, x);
            const int xlo = (int)(uint32_t)xb, xhi = (int)(uint32_t)(xb >> 32);
            for (uint32_t u = 0; u < cnt; ++u) {
                const uint32_t l = (uint32_t)__builtin_amdgcn_readlane(xlo, (int)u);
                const uint32_t h = (uint32_t)__builtin_amdgcn_readlane(xhi, (int)u);
                acc += __builtin_bit_cast(double, ((uint64_t)h << 32) | l);
            }
            x = xn;
        }
        if (lane == 0) deg[v] = acc;
    }
}

// One workgroup per row longer than kSeqDeg (list computed on the host).
template <typename S>
__global__ __launch_bounds__(kBlock) void deg_long(const uint32_t *__restrict__ off,
                                                   const uint32_t *__restrict__ eid,
                                                   const void *w, int kind, double value,
                                                   const uint32_t *__restrict__ rows,
                                                   bool accumulate, double *__restrict__ deg) {
    __shared__ double red[8];
    const uint32_t v = rows[blockIdx.x];
    const uint32_t lo = off[v], hi = off[v + 1];
    // Each thread's entries j, j + kBlock, ... in order, eight in flight (the
    // hub rows are tens of thousands of entries long).
    double acc = 0.0;
    uint32_t j = lo + threadIdx.x;
    for (; j + 7 * kBlock < hi; j += 8 * kBlock) {
        uint32_t e[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) e[i] = eid[j + i * kBlock];
        double x[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) x[i] = read_w<S>(w, e[i], kind, value);
#pragma unroll
        for (int i = 0; i < 8; ++i) acc += x[i];
    }
    for (; j < hi; j += kBlock) acc += read_w<S>(w, eid[j], kind, value);
    acc = block_sum(acc, red);
    if (threadIdx.x == 0) deg[v] = accumulate ? deg[v] + acc : acc;
}

// wcsc[j] = w(eid[j]) in the value type of the sweep.  Each thread takes
// eight entries a stride apart per round (coalesced), all eight index loads
// and then all eight gathers in flight before the stores.
template <typename S, typename V>
__global__ __launch_bounds__(kBlock) void permute_w(const uint32_t *__restrict__ eid,
                                                    const void *w, int kind, double value,
                                                    uint64_t m, V *__restrict__ out) {
    // Each XCD takes a contiguous eighth of the entries (see deg_short); the
    // edge ids and the output stream past L2.
    uint64_t begin = 0, end = m, first = (uint64_t)blockIdx.x * kBlock + threadIdx.x,
             stride = (uint64_t)gridDim.x * kBlock;
    if (gridDim.x % kXcds == 0) {
        const uint32_t x = blockIdx.x % kXcds;
        begin = m * x / kXcds;
        end = m * (x + 1) / kXcds;
        first = begin + (uint64_t)(blockIdx.x / kXcds) * kBlock + threadIdx.x;
        stride = (uint64_t)(gridDim.x / kXcds) * kBlock;
    }
    for (uint64_t j0 = first; j0 < end; j0 += 8 * stride) {
        uint32_t e[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const uint64_t j = j0 + u * stride;
            e[u] = j < end ? __builtin_nontemporal_load(eid + j) : 0u;
        }
        double x[8];
#pragma unroll
        for (int u = 0; u < 8; ++u)
            x[u] = j0 + u * stride < end ? read_w<S>(w, e[u], kind, value) : 0.0;
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const uint64_t j = j0 + u * stride;
            if (j < end) __builtin_nontemporal_store(static_cast<V>(x[u]), out + j);
        }
    }
}

__global__ __launch_bounds__(kBlock) void f32_to_f64(const float *__restrict__ a, uint64_t n,
                                                     double *__restrict__ b) {
    for (uint64_t j = (uint64_t)blockIdx.x * kBlock + threadIdx.x; j < n;
         j += (uint64_t)gridDim.x * kBlock)
        b[j] = static_cast<double>(a[j]);
}

static unsigned grid_for(uint64_t n) {
    uint64_t b = (n + kBlock - 1) / kBlock;
    if (b > 4096) b = 4096;
    return b ? (unsigned)b : 1u;
}

// Long rows of an orientation (> kSeqDeg entries), cached.
// Host lists of the rows deg_mid and deg_long take, built once per orientation.
static gt_status heavy_rows(Orient &o, uint32_t n) {
    if (o.heavy_ready) return GT_OK;
    std::vector<uint32_t> mid, heavy;
    for (uint32_t v = 0; v < n; ++v) {
        const uint32_t len = o.h_off[v + 1] - o.h_off[v];
        if (len > kSeqDeg) heavy.push_back(v);
        else if (len >= kWaveDeg) mid.push_back(v);
    }
    o.nheavy = (uint32_t)heavy.size();
    o.nmid = (uint32_t)mid.size();
    GT_TRY(o.heavy.reserve(sizeof(uint32_t) * std::max<size_t>(1, heavy.size())));
    GT_TRY(o.mid.reserve(sizeof(uint32_t) * std::max<size_t>(1, mid.size())));
    if (!heavy.empty())
        GT_HIP(hipMemcpy(o.heavy.ptr, heavy.data(), sizeof(uint32_t) * heavy.size(),
                         hipMemcpyHostToDevice));
    if (!mid.empty())
        GT_HIP(hipMemcpy(o.mid.ptr, mid.data(), sizeof(uint32_t) * mid.size(),
                         hipMemcpyHostToDevice));
    o.heavy_ready = true;
    return GT_OK;
}

gt_status stage_edge_source(gt_graph *g, const gt_prop *w, const void **src, int *dtype) {
    if (w->kind == GT_MAP_CONSTANT) {
        *src = nullptr;
        *dtype = GT_F64;
        return GT_OK;
    }
    if (w->kind != GT_MAP_DENSE) return fail(GT_ERR_INVALID_ARG, "unknown weight map kind");
    if (!w->data && g->edge_bound) return fail(GT_ERR_INVALID_ARG, "weight map data is null");
    if (w->len < g->edge_bound)
        return fail(GT_ERR_INVALID_ARG, "weight map is shorter than the edge bound");
    if (w->graph_id && w->graph_id != g->id)
        return fail(GT_ERR_WRONG_GRAPH, "weight map does not belong to this graph");
    if (w->dtype != GT_F64 && w->dtype != GT_F32)
        return fail(GT_ERR_INVALID_ARG, "weight map dtype must be f64 or f32");
    *dtype = w->dtype;
    if (w->loc == GT_DEVICE) {
        *src = w->data;
        return GT_OK;
    }
    const size_t bytes = (w->dtype == GT_F64 ? 8 : 4) * (size_t)g->edge_bound;
    GT_TRY(g->ws.wsrc.reserve(bytes));
    if (bytes) GT_HIP(hipMemcpyAsync(g->ws.wsrc.ptr, w->data, bytes, hipMemcpyHostToDevice, g->stream));
    *src = g->ws.wsrc.ptr;
    return GT_OK;
}

template <typename S>
static gt_status launch_deg(gt_graph *g, Orient &push, const void *src, int kind, double value,
                            double *deg, bool accumulate) {
    hipStream_t s = g->stream;
    const uint32_t n = g->n;
    hipLaunchKernelGGL((deg_short<S>), dim3(grid_for(n)), dim3(kBlock), 0, s,
                       push.off.as<uint32_t>(), push.eid.as<uint32_t>(), src, kind, value, n,
                       accumulate, deg);
    GT_TRY(heavy_rows(push, n));
    if (push.nmid) {
        const unsigned mg = (unsigned)std::min<uint32_t>(8192, (push.nmid + 3) / 4);
        hipLaunchKernelGGL((deg_mid<S>), dim3(mg), dim3(kBlock), 0, s, push.off.as<uint32_t>(),
                           push.eid.as<uint32_t>(), src, kind, value, push.mid.as<uint32_t>(),
                           push.nmid, accumulate, deg);
    }
    if (push.nheavy)
        hipLaunchKernelGGL((deg_long<S>), dim3(push.nheavy), dim3(kBlock), 0, s,
                           push.off.as<uint32_t>(), push.eid.as<uint32_t>(), src, kind, value,
                           push.heavy.as<uint32_t>(), accumulate, deg);
    GT_HIP(hipGetLastError());
    return GT_OK;
}

// out[j] = w(eid[j]) for the `m` entries of an orientation.
template <typename S, typename V>
static gt_status launch_permute(gt_graph *g, const uint32_t *eid, uint64_t m, const void *src,
                                int kind, double value, V *out) {
    if (m == 0) return GT_OK;
    hipLaunchKernelGGL((permute_w<S, V>), dim3(grid_for(m)), dim3(kBlock), 0, g->stream, eid, src,
                       kind, value, m, out);
    GT_HIP(hipGetLastError());
    return GT_OK;
}

gt_status stage_degree(gt_graph *g, const gt_prop *w, Orient &push, const void *src, int dtype,
                       DevBuf &out) {
    GT_TRY(out.reserve(sizeof(double) * std::max<uint32_t>(1, g->n)));
    if (dtype == GT_F32)
        return launch_deg<float>(g, push, src, w->kind, w->value, out.as<double>(), false);
    return launch_deg<double>(g, push, src, w->kind, w->value, out.as<double>(), false);
}

gt_status stage_degree_into(gt_graph *g, const gt_prop *w, Orient &o, const void *src, int dtype,
                            double *out, bool accumulate) {
    if (g->n == 0) return GT_OK;
    if (dtype == GT_F32)
        return launch_deg<float>(g, o, src, w->kind, w->value, out, accumulate);
    return launch_deg<double>(g, o, src, w->kind, w->value, out, accumulate);
}

template <typename V>
gt_status stage_csc_weights(gt_graph *g, const gt_prop *w, Orient &pull, const void *src,
                            int dtype, DevBuf &out) {
    GT_TRY(out.reserve(sizeof(V) * std::max<uint64_t>(1, g->m)));
    const uint32_t *eid = pull.eid.as<uint32_t>();
    const uint64_t m = pull.h_off.empty() ? 0 : pull.h_off.back();
    if (dtype == GT_F32)
        return launch_permute<float, V>(g, eid, m, src, w->kind, w->value, out.as<V>());
    return launch_permute<double, V>(g, eid, m, src, w->kind, w->value, out.as<V>());
}
template gt_status stage_csc_weights<double>(gt_graph *, const gt_prop *, Orient &, const void *,
                                             int, DevBuf &);
template gt_status stage_csc_weights<float>(gt_graph *, const gt_prop *, Orient &, const void *,
                                            int, DevBuf &);

template <typename V>
gt_status stage_pull_weights(gt_graph *g, const gt_prop *w, Orient &pull, const void *src,
                             int dtype, DevBuf &out) {
    if (!pull.scheduled) return fail(GT_ERR_INVALID_ARG, "internal: pull weights before schedule");
    GT_TRY(out.reserve(sizeof(V) * std::max<uint64_t>(1, g->m)));
    const uint32_t *eid = pull.peid.as<uint32_t>();
    const uint64_t m = pull.h_off.empty() ? 0 : pull.h_off.back();
    if (dtype == GT_F32)
        return launch_permute<float, V>(g, eid, m, src, w->kind, w->value, out.as<V>());
    return launch_permute<double, V>(g, eid, m, src, w->kind, w->value, out.as<V>());
}
template gt_status stage_pull_weights<double>(gt_graph *, const gt_prop *, Orient &, const void *,
                                              int, DevBuf &);

gt_status widen_pull_weights(gt_graph *g, const Orient &pull, const DevBuf &f32, DevBuf &out) {
    const uint64_t m = pull.h_off.empty() ? 0 : pull.h_off.back();
    GT_TRY(out.reserve(sizeof(double) * std::max<uint64_t>(1, g->m)));
    if (m == 0) return GT_OK;
    hipLaunchKernelGGL(f32_to_f64, dim3(grid_for(m)), dim3(kBlock), 0, g->stream, f32.as<float>(), m,
                       out.as<double>());
    GT_HIP(hipGetLastError());
    return GT_OK;
}
template gt_status stage_pull_weights<float>(gt_graph *, const gt_prop *, Orient &, const void *,
                                             int, DevBuf &);

gt_status stage_vertex_prop(gt_graph *g, const gt_prop *pm, DevBuf &buf, const double **dev,
                            double *scalar, bool *dense, double default_value) {
    *dev = nullptr;
    *dense = false;
    if (!pm) {
        *scalar = default_value;
        return GT_OK;
    }
    if (pm->kind == GT_MAP_UNITY) {
        *scalar = 1.0;
        return GT_OK;
    }
    if (pm->kind == GT_MAP_CONSTANT) {
        *scalar = pm->value;
        return GT_OK;
    }
    if (pm->kind != GT_MAP_DENSE) return fail(GT_ERR_INVALID_ARG, "unknown vertex map kind");
    const uint32_t n = g->n;
    if (pm->len < n) return fail(GT_ERR_INVALID_ARG, "vertex map is shorter than the vertex bound");
    if (!pm->data && n) return fail(GT_ERR_INVALID_ARG, "vertex map data is null");
    if (pm->graph_id && pm->graph_id != g->id)
        return fail(GT_ERR_WRONG_GRAPH, "vertex map does not belong to this graph");
    *dense = true;
    if (pm->dtype == GT_F64 && pm->loc == GT_DEVICE) {
        *dev = static_cast<const double *>(pm->data);
        return GT_OK;
    }
    GT_TRY(buf.reserve(sizeof(double) * std::max<uint32_t>(1, n)));
    if (pm->dtype == GT_F64) {
        if (n) GT_HIP(hipMemcpyAsync(buf.ptr, pm->data, sizeof(double) * n, hipMemcpyHostToDevice,
                                     g->stream));
    } else if (pm->dtype == GT_F32) {
        const float *f = static_cast<const float *>(pm->data);
        if (pm->loc == GT_HOST) {
            std::vector<double> tmp(f, f + n);
            GT_HIP(hipMemcpyAsync(buf.ptr, tmp.data(), sizeof(double) * n, hipMemcpyHostToDevice,
                                  g->stream));
            GT_HIP(hipStreamSynchronize(g->stream));
        } else {
            hipLaunchKernelGGL(f32_to_f64, dim3(grid_for(n)), dim3(kBlock), 0, g->stream, f,
                               (uint64_t)n, buf.as<double>());
            GT_HIP(hipGetLastError());
        }
    } else {
        return fail(GT_ERR_INVALID_ARG, "vertex map dtype must be f64 or f32");
    }
    *dev = buf.as<double>();
    return GT_OK;
}

}  // namespace gt
