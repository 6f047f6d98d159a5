// Host-side loop driver shared by the power iterations.
#pragma once

#include <cmath>

#include "gt_internal.hpp"

namespace gt {

// Profiling: HIP events around every sweep launch, on the stream it runs on.
gt_status prof_prepare(gt_graph *g);
gt_status prof_begin(gt_graph *g, uint64_t k);
gt_status prof_end(gt_graph *g, uint64_t k);
gt_status prof_flush(gt_graph *g, uint64_t launched);

// Launch sweeps until the device reports done.  The first batch is the
// sweep count the same kind of loop (`hint_slot`) took last time on this
// graph, so a repeated solve waits on the device once; later batches are
// sized from the observed geometric decay of delta, so the host waits a few
// times per solve and few launches find the loop already finished.
// `sweep(k)` issues the SpMV kernel of sweep k of this run (k = 0, 1, ...),
// timed when profiling; `fold(k)` then issues what completes the sweep (the
// split rows and the convergence decision).
template <typename Sweep, typename Fold>
gt_status drive(gt_graph *g, SweepState *dst, double eps_pred, uint64_t cap, Sweep sweep,
                Fold fold, uint64_t *launched_out, int hint_slot) {
    hipStream_t s = g->stream;
    uint64_t launched = 0;
    const uint64_t hint = g->sweep_hint[hint_slot];
    uint64_t batch = (hint >= 2 && hint <= 64) ? hint : 2;
    double prev_delta = -1.0;
    uint64_t prev_iter = 0;
    for (;;) {
        if (cap > 0 && batch > cap - launched) batch = cap - launched;
        for (uint64_t k = 0; k < batch; ++k, ++launched) {
            GT_TRY(prof_begin(g, launched));
            sweep(launched);
            GT_TRY(prof_end(g, launched));
            fold(launched);
        }
        GT_HIP(hipGetLastError());
        GT_HIP(hipMemcpyAsync(g->h_state, dst, sizeof(SweepState), hipMemcpyDeviceToHost, s));
        GT_HIP(hipStreamSynchronize(s));
        const SweepState h = *g->h_state;
        if (h.done || (cap > 0 && launched >= cap)) {
            g->sweep_hint[hint_slot] = h.launches;
            break;
        }
        uint64_t next = 4;
        if (prev_delta > 0.0 && h.delta > 0.0 && h.iter > prev_iter && h.delta < prev_delta) {
            const double rho =
                std::pow(h.delta / prev_delta, 1.0 / (double)(h.iter - prev_iter));
            const double need = std::log(eps_pred / h.delta) / std::log(rho);
            next = (std::isfinite(need) && need > 1.0) ? (uint64_t)std::ceil(need) : 1;
            if (next > 64) next = 64;
        } else if (prev_delta > 0.0) {
            next = 8;
        }
        prev_delta = h.delta;
        prev_iter = h.iter;
        batch = next;
    }
    GT_TRY(prof_flush(g, launched));
    *launched_out = launched;
    return GT_OK;
}


}  // namespace gt
