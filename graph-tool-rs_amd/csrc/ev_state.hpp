// Eigenvector loop state and the end-of-launch decision, shared by the wave
// engine's fold (eigenvector.hip) and the block-tile fold (blocks.hip).
//
// Launch L writes the unscaled y_L = A x_{L-1} (gathered as (A y_{L-1}) /
// scale_{L-1}) and, in the same pass, the delta of sweep L-1, so a sweep's
// convergence is learnt one launch after it (centrality.rs:405-497 checks it
// at the end of the sweep itself; the extra launch changes no result).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "gt_internal.hpp"

namespace gt {

// Laid over SweepState's storage: delta, iter, done and bail sit at
// SweepState's offsets so the shared driver can read them.
struct EvState {
    double sc_cur;      // scale of y_{L-1}
    double delta;       // delta of the last completed check
    double sc_prev;     // scale of y_{L-2}
    double norm_last;   // norm_{L-1}
    double eig;         // norm of the final iterate
    uint64_t iter;      // reference sweep count reached
    uint32_t done;
    uint32_t counter;
    uint32_t launches;  // launches that did work (= L)
    uint32_t bail;      // set by a block sweep that cannot represent its sums
    double final_scale; // scale of the final iterate
    uint64_t final_iter;
};
static_assert(offsetof(EvState, delta) == offsetof(SweepState, delta), "layout");
static_assert(offsetof(EvState, iter) == offsetof(SweepState, iter), "layout");
static_assert(offsetof(EvState, done) == offsetof(SweepState, done), "layout");
static_assert(offsetof(EvState, bail) == offsetof(SweepState, bail), "layout");
static_assert(sizeof(EvState) >= sizeof(SweepState), "layout");

// Closes launch L = launches + 1 given its squared norm s2 = |y_L|^2 and d,
// the delta of sweep L-1 (meaningless for L = 1): sweep L-1 is the answer
// when d < epsilon; sweep L is when it reaches max_iter.  Then the scales
// move on.  One thread of the last workgroup calls it.
__device__ __forceinline__ void ev_close(EvState *st, double s2, double d, double epsilon,
                                         uint64_t max_iter) {
    const uint32_t L = st->launches + 1;
    const double norm = sqrt(s2);
    bool done = false;
    if (L >= 2) {
        st->delta = d;
        if (!(d >= epsilon)) {
            done = true;
            st->final_iter = L - 1;
            st->final_scale = st->sc_cur;
            st->eig = st->norm_last;
            st->iter = L - 1;
        }
    }
    if (!done) {
        st->iter = L;
        if (max_iter > 0 && L >= max_iter) {
            done = true;
            st->final_iter = L;
            st->final_scale = norm == 0.0 ? 1.0 : norm;
            st->eig = norm;
        }
    }
    st->sc_prev = st->sc_cur;
    st->sc_cur = norm == 0.0 ? 1.0 : norm;
    st->norm_last = norm;
    st->launches = L;
    st->done = done ? 1u : 0u;
    st->counter = 0;
}

}  // namespace gt
