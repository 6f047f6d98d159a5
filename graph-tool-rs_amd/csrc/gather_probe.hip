// libgt_probe.so: the random-gather rate behind the pull sweeps' roofline
// (see include/gt_probe.h).  Measurement only: nothing in libgt_hip.so calls
// it.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>

#include "gt_probe.h"

namespace {

thread_local std::string g_err;

int probe_fail(const char *what, hipError_t e) {
    g_err = std::string(what) + ": " + hipGetErrorString(e);
    return -1;
}

constexpr int kWaves = 8;
constexpr int kThreads = kWaves * 64;
constexpr uint32_t kTile = 1024;

// Label of entry e: a 32-bit mix of e (the low bits of a multiply-xorshift
// hash), reduced to the window.
__device__ __forceinline__ uint32_t label_of(uint64_t e, uint32_t window) {
    uint64_t x = e * 0x9E3779B97F4A7C15ull;
    x ^= x >> 29;
    x *= 0xBF58476D1CE4E5B9ull;
    x ^= x >> 32;
    return (uint32_t)(x % window);
}

__global__ __launch_bounds__(kThreads) void fill_labels(uint32_t *idx, uint64_t n, uint32_t window) {
    for (uint64_t e = (uint64_t)blockIdx.x * kThreads + threadIdx.x; e < n;
         e += (uint64_t)gridDim.x * kThreads)
        idx[e] = label_of(e, window);
}

__global__ __launch_bounds__(kThreads) void fill_src(float *src, uint32_t window) {
    for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < window; i += gridDim.x * kThreads)
        src[i] = (float)(i & 7);
}

// One workgroup per compute unit; wavefront w of the grid takes tiles w,
// w + W, ...; per tile, 16 coalesced label loads, then 16 gathers for the
// lanes whose entry is below `active` (exec-masked), all in flight before the
// first add.
__global__ __launch_bounds__(kThreads) void gather_walk(const uint32_t *__restrict__ idx,
                                                        const float *__restrict__ src,
                                                        uint32_t tiles, uint32_t active,
                                                        float *__restrict__ out) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wave = blockIdx.x * kWaves + threadIdx.x / 64;
    const uint32_t waves = gridDim.x * kWaves;
    float acc = 0.f;
    for (uint32_t t = wave; t < tiles; t += waves) {
        const uint32_t *p = idx + (size_t)t * kTile;
        uint32_t q[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) q[k] = __builtin_nontemporal_load(p + lane + 64 * k);
        float g[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            g[k] = 0.f;
            if (lane + 64u * k < active) g[k] = src[q[k]];
        }
#pragma unroll
        for (int k = 0; k < 16; ++k) acc += g[k];
    }
    out[(size_t)blockIdx.x * kThreads + threadIdx.x] = acc;
}

}  // namespace

extern "C" {

const char *gt_probe_error(void) { return g_err.c_str(); }

int gt_probe_gather(uint32_t window, uint32_t tiles, uint32_t active, uint32_t reps,
                    double *us_per_launch, uint32_t *compute_units) {
    g_err.clear();
    if (!us_per_launch || !compute_units) {
        g_err = "null output";
        return -1;
    }
    if (window == 0 || tiles == 0 || active > kTile || reps == 0) {
        g_err = "window and tiles must be positive, active at most 1024, reps positive";
        return -1;
    }
    int dev = 0, cus = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return probe_fail("device query", e);
    const uint64_t n = (uint64_t)tiles * kTile;
    uint32_t *idx = nullptr;
    float *src = nullptr, *out = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t a = nullptr, b = nullptr;
    int rc = 0;
    float ms = 0.f;
    if ((e = hipMalloc(&idx, n * sizeof(uint32_t))) != hipSuccess) { rc = probe_fail("hipMalloc", e); goto done; }
    if ((e = hipMalloc(&src, (size_t)window * sizeof(float))) != hipSuccess) { rc = probe_fail("hipMalloc", e); goto done; }
    if ((e = hipMalloc(&out, (size_t)cus * kThreads * sizeof(float))) != hipSuccess) { rc = probe_fail("hipMalloc", e); goto done; }
    if ((e = hipStreamCreate(&s)) != hipSuccess) { rc = probe_fail("hipStreamCreate", e); goto done; }
    if ((e = hipEventCreate(&a)) != hipSuccess || (e = hipEventCreate(&b)) != hipSuccess) {
        rc = probe_fail("hipEventCreate", e);
        goto done;
    }
    hipLaunchKernelGGL(fill_labels, dim3(4096), dim3(kThreads), 0, s, idx, n, window);
    hipLaunchKernelGGL(fill_src, dim3(1024), dim3(kThreads), 0, s, src, window);
    for (int r = 0; r < 2; ++r)
        hipLaunchKernelGGL(gather_walk, dim3(cus), dim3(kThreads), 0, s, idx, src, tiles, active, out);
    if ((e = hipGetLastError()) != hipSuccess) { rc = probe_fail("launch", e); goto done; }
    hipEventRecord(a, s);
    for (uint32_t r = 0; r < reps; ++r)
        hipLaunchKernelGGL(gather_walk, dim3(cus), dim3(kThreads), 0, s, idx, src, tiles, active, out);
    hipEventRecord(b, s);
    if ((e = hipEventSynchronize(b)) != hipSuccess) { rc = probe_fail("hipEventSynchronize", e); goto done; }
    if ((e = hipEventElapsedTime(&ms, a, b)) != hipSuccess) { rc = probe_fail("hipEventElapsedTime", e); goto done; }
    *us_per_launch = (double)ms * 1e3 / reps;
    *compute_units = (uint32_t)cus;
done:
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    if (s) (void)hipStreamDestroy(s);
    if (out) (void)hipFree(out);
    if (src) (void)hipFree(src);
    if (idx) (void)hipFree(idx);
    return rc;
}

}  // extern "C"
