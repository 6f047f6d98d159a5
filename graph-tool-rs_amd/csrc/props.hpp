// Property-map staging used by the centrality entry points.
#pragma once

#include "gt_internal.hpp"

namespace gt {

// Resolve a weight map to a device source indexed by edge id.  CONSTANT maps
// yield src == nullptr (the kernels read w->value instead).
gt_status stage_edge_source(gt_graph *g, const gt_prop *w, const void **src, int *dtype);

// deg(v) = sum of w over the out-entries of `push` row v, in storage order.
gt_status stage_degree(gt_graph *g, const gt_prop *w, Orient &push, const void *src, int dtype,
                       DevBuf &out);

// out[v] (+)= sum of w over row v of `o` in storage order; with `accumulate`
// the sum continues from out[v] (a second half of the same block).
gt_status stage_degree_into(gt_graph *g, const gt_prop *w, Orient &o, const void *src, int dtype,
                            double *out, bool accumulate);

// Weights permuted into the storage order of `pull`, converted to V.
template <typename V>
gt_status stage_csc_weights(gt_graph *g, const gt_prop *w, Orient &pull, const void *src,
                            int dtype, DevBuf &out);
// Weights in the packed order of `pull`'s wave tiles (ensure_schedule first),
// converted to V.
template <typename V>
gt_status stage_pull_weights(gt_graph *g, const gt_prop *w, Orient &pull, const void *src,
                             int dtype, DevBuf &out);
// The f32 pull weights (packed order, as stage_pull_weights<float> wrote them)
// widened to f64: the same values as staging an f32 map in f64 again, without
// the gather through the edge ids.
gt_status widen_pull_weights(gt_graph *g, const Orient &pull, const DevBuf &f32, DevBuf &out);

// A vertex map as a scalar (UNITY / CONSTANT / absent -> default_value) or
// an f64 device array.
gt_status stage_vertex_prop(gt_graph *g, const gt_prop *pm, DevBuf &buf, const double **dev,
                            double *scalar, bool *dense, double default_value);

// Shared validation of the centrality entry points.
// The four directed views: pull / push orientations, and for a filtered view
// the device vertex mask (`active`, else null) and the kept-vertex count.
gt_status resolve_view(gt_graph *g, const gt_view *view, Orient **pull, Orient **push,
                       bool *reversed, const uint8_t **active, uint32_t *nactive);
gt_status filtered_view(gt_graph *g, const gt_view *view, bool reversed, Orient **pull,
                        Orient **push, const uint8_t **active, uint32_t *nactive);
gt_status check_vertex_out(gt_graph *g, const gt_vertex_out *r, const char *who);
gt_status check_params(gt_graph *g, const gt_power_iteration *p);

}  // namespace gt
