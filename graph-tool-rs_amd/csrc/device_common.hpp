// Device helpers shared by the sweep kernels: deterministic wave / workgroup
// reductions (fixed shuffle tree, so a result depends only on the inputs) and
// the last-workgroup ticket that folds per-workgroup partials in index order.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace gt {

constexpr int kWave = 64;

// Most workgroups of a kernel that ends in a last-workgroup ticket: each
// arrival is an atomic on one counter, and thousands of them serialise.
constexpr unsigned kTicketGrid = 512;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// Sum over a 256-thread workgroup; every thread gets the total.  `red` must
// hold 4 doubles of LDS.  Contains two barriers.
__device__ __forceinline__ double block_sum(double v, double *red) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    v = wave_sum(v);
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double t = (red[0] + red[1]) + (red[2] + red[3]);
    __syncthreads();
    return t;
}

// Two sums at once (one barrier pair instead of two).
__device__ __forceinline__ void block_sum2(double &a, double &b, double *red) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    a = wave_sum(a);
    b = wave_sum(b);
    if (lane == 0) {
        red[wid] = a;
        red[4 + wid] = b;
    }
    __syncthreads();
    a = (red[0] + red[1]) + (red[2] + red[3]);
    b = (red[4] + red[5]) + (red[6] + red[7]);
    __syncthreads();
}

// After a workgroup has published its partials: returns true in exactly one
// workgroup of the grid -- the last to arrive -- with every other
// workgroup's writes visible to it.  `counter` must be zero at launch; the
// caller of the true branch resets it.
//
// One thread per workgroup fences: an agent-scope release (L2 write-back, no
// invalidate) before taking the ticket, and an acquire (which invalidates
// the non-coherent caches) only in the last workgroup.  The workgroup barrier
// before the ticket makes the other threads' stores part of what the release
// publishes; the one after it extends the acquire to the whole workgroup.  A
// fence in every thread, or an acquire in every workgroup, would invalidate
// the L2 under the gathers of the workgroups still running.
__device__ __forceinline__ bool last_to_arrive(uint32_t *counter, uint32_t *s_flag) {
    __syncthreads();
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const uint32_t t = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_AGENT);
        const bool last = t == gridDim.x - 1;
        if (last) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        *s_flag = last ? 1u : 0u;
    }
    __syncthreads();
    return *s_flag != 0;
}


// The share of `units` consecutive work items this workgroup's XCD walks when
// the grid is dealt to the eight XCDs in turn (workgroup b on XCD b % 8):
// XCD x takes items [begin, begin + count), the x-th eighth, and its
// workgroups walk them with `first` / `stride` (in wavefronts of NW per
// workgroup).  A grid that is not a multiple of eight walks all items as one.
constexpr uint32_t kXcds = 8;
struct XcdSpan {
    uint32_t begin, count, index, groups;
    __device__ uint32_t first(uint32_t nw) const {
        return index * nw + (uint32_t)__builtin_amdgcn_readfirstlane(threadIdx.x / 64);
    }
    __device__ uint32_t stride(uint32_t nw) const { return groups * nw; }
};
__device__ __forceinline__ XcdSpan xcd_span(uint32_t units) {
    XcdSpan s;
    if (gridDim.x % kXcds != 0) {
        s.begin = 0;
        s.count = units;
        s.index = blockIdx.x;
        s.groups = gridDim.x;
        return s;
    }
    const uint32_t x = blockIdx.x % kXcds;
    s.begin = (uint32_t)((uint64_t)units * x / kXcds);
    s.count = (uint32_t)((uint64_t)units * (x + 1) / kXcds) - s.begin;
    s.index = blockIdx.x / kXcds;
    s.groups = gridDim.x / kXcds;
    return s;
}

}  // namespace gt
