/*
 * gt_hip.h -- C ABI of the MI355X (gfx950) PageRank / eigenvector-centrality
 * engine.  Plain pointers and sizes only; every entry point returns a
 * gt_status and leaves a message for gt_last_error() on failure.
 *
 * Reference interface replaced (graph-tool-rs, paths relative to its tree):
 *   gt_pagerank_f64 ....... crates/gt-algo/src/centrality.rs:234-248 (pagerank)
 *                           and :261-384 (pagerank_with)
 *   gt_eigenvector_f64 .... crates/gt-algo/src/centrality.rs:405-497 (eigenvector)
 *   gt_betweenness ........ crates/gt-algo/src/centrality.rs:500-689 (Brandes, betweenness)
 *   gt_degrees, gt_out_degrees, gt_degree_histogram
 *                           crates/gt-algo/src/degree.rs:122-247
 *   gt_read_gt ............ crates/gt-io/src/gt.rs:555 (read)
 *   gt_read_edge_list ..... crates/gt-io/src/csv.rs:129 (read_edge_list)
 *   gt_power_iteration .... crates/gt-algo/src/centrality.rs:79-99 (PowerIteration)
 *   gt_plan ............... crates/gt-core/src/par/plan.rs:38-101 (Plan{n, chunks})
 *   gt_prop ............... crates/gt-core/src/prop/map.rs:28-56 (ReadProp with
 *                           IS_UNITY / IS_CONSTANT), prop/dense.rs:373-460
 *                           (Unity, Constant), prop/dense.rs:88-127 (DenseProp)
 *   gt_view ............... crates/gt-py/src/dispatch.rs:17-60 (ViewKind; the four
 *                           directed arms dispatch_bidi accepts, :315-336)
 *   gt_graph_upload ....... the in-halves / out-halves an &AdjList yields through
 *                           Bidirectional::in_edges / GraphRef::out_edges
 *                           (crates/gt-core/src/graph.rs:71-107, adj/list.rs:156-161)
 *   gt_graph_build_edges .. AdjList::add_edge in edge-id order
 *                           (crates/gt-core/src/adj/list.rs:258-280, 471-500) and
 *                           ParBuilder::build (adj/builder.rs:178-197)
 *   error codes ........... the panics of centrality.rs:145-149 (plan cover),
 *                           :300-302 / :422-424 (wrong graph), dispatch.rs:176-179
 *                           (NotDirected)
 * Entry points marked "extension" have no counterpart in the reference.
 */
#ifndef GT_HIP_H
#define GT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum gt_status {
    GT_OK = 0,
    GT_ERR_INVALID_ARG = 1,   /* malformed argument (null pointer, short map, ...) */
    GT_ERR_PLAN_COVER = 2,    /* "sweep: the plan must cover the output exactly" */
    GT_ERR_WRONG_GRAPH = 3,   /* "...map does not belong to this graph" */
    GT_ERR_NOT_DIRECTED = 4,  /* undirected view passed to a Bidirectional kernel */
    GT_ERR_NO_SUCH_VERTEX = 5,/* edge endpoint >= vertex count (AdjList::add_edge) */
    GT_ERR_HIP = 6,           /* a HIP runtime call failed */
    GT_ERR_OOM = 7,           /* device allocation failed */
    GT_ERR_UNSUPPORTED = 8    /* valid request the library does not serve */
} gt_status;

/* Where a buffer argument lives. */
typedef enum gt_loc { GT_HOST = 0, GT_DEVICE = 1 } gt_loc;
/* Element type of a dense buffer. */
typedef enum gt_dtype { GT_F64 = 0, GT_F32 = 1 } gt_dtype;
/* Property-map kinds the path distinguishes (ReadProp::IS_UNITY / IS_CONSTANT). */
typedef enum gt_map_kind { GT_MAP_UNITY = 0, GT_MAP_CONSTANT = 1, GT_MAP_DENSE = 2 } gt_map_kind;

/* A read-only property map: edge weights (indexed by edge id, len >= edge_bound)
 * or a vertex map such as the personalisation vector (len >= vertex_bound). */
typedef struct gt_prop {
    int32_t kind;        /* gt_map_kind */
    int32_t dtype;       /* gt_dtype of data (DENSE only) */
    int32_t loc;         /* gt_loc of data (DENSE only) */
    int32_t reserved;
    double value;        /* CONSTANT */
    const void *data;    /* DENSE */
    uint64_t len;        /* DENSE: element count */
    uint64_t graph_id;   /* owner graph of a DENSE map; 0 = not checked */
} gt_prop;

/* Plan{n, chunks}: fixes the CPU fold order in the reference.  The device
 * path checks the cover (len == vertex_bound) and otherwise ignores it. */
typedef struct gt_plan {
    uint64_t len;
    uint64_t chunks;
} gt_plan;

typedef struct gt_power_iteration {
    double epsilon;      /* L1 change per sweep that stops the loop */
    uint64_t max_iter;   /* 0 = no cap */
    gt_plan plan;
} gt_power_iteration;

/* A writable vertex map that is both the initial iterate and the result
 * (DenseProp<f64, VertexTag> after sized_for): data[0..vertex_bound) is used,
 * slots [had, vertex_bound) are seeded by the kernel, len >= vertex_bound. */
typedef struct gt_vertex_out {
    double *data;
    uint64_t len;
    uint64_t had;
    int32_t loc;         /* gt_loc */
    int32_t reserved;
    uint64_t graph_id;   /* owner graph; 0 = not checked */
} gt_vertex_out;

typedef enum gt_view_kind {
    GT_VIEW_DIRECTED = 0,
    GT_VIEW_REVERSED = 1,
    GT_VIEW_DIRECTED_FILTERED = 2,
    GT_VIEW_REVERSED_FILTERED = 3,
    GT_VIEW_UNDIRECTED = 4,
    GT_VIEW_UNDIRECTED_FILTERED = 5
} gt_view_kind;

/* The view a call runs on.  Masks are uint8 (non-zero keeps), indexed by the
 * unfiltered vertex / edge index, host memory, len >= the bound
 * (Filtered::masked, crates/gt-core/src/view/filtered.rs:170-200). */
typedef struct gt_view {
    int32_t kind;        /* gt_view_kind */
    int32_t reserved;
    const uint8_t *vmask;
    uint64_t vmask_len;
    const uint8_t *emask;
    uint64_t emask_len;
} gt_view;

/* Opaque device-resident graph: CSC of the in-halves and CSR of the
 * out-halves, both in AdjList storage order, plus the SpMV row schedule. */
typedef struct gt_graph gt_graph;

/* Message of the last failed call on this thread ("" if none). */
const char *gt_last_error(void);

/* Number of visible devices; 0 when no GPU is present. */
gt_status gt_device_count(int32_t *count);

/* Extension: a cap on the device memory the library's own buffers may hold
 * (graphs, schedules, workspaces), 0 for none (the default).  A call whose
 * allocation would pass it fails with GT_ERR_OOM, the sizes in
 * gt_last_error(), exactly as when the device itself runs out.
 * gt_device_memory_in_use reports the bytes held now. */
gt_status gt_set_device_memory_limit(uint64_t bytes);
gt_status gt_device_memory_in_use(uint64_t *bytes);

/* ---------------------------------------------------------------- graphs */

/* Upload an adjacency that already exists on the host, in storage order:
 * in_off[n+1] / in_other[m] / in_eid[m] are the in-halves (sources and edge ids
 * as in_edges(v) yields them), out_off / out_other / out_eid the out-halves.
 * Edge ids must be < edge_bound.  All pointers are host memory. */
gt_status gt_graph_upload(uint64_t graph_id, uint32_t n, uint64_t edge_bound,
                          const uint32_t *in_off, const uint32_t *in_other,
                          const uint32_t *in_eid, const uint32_t *out_off,
                          const uint32_t *out_other, const uint32_t *out_eid,
                          gt_graph **out);

/* Build on the device from an edge list, exactly as add_edge(src[i], tgt[i])
 * for i = 0..m would: edge i gets id i, out-halves in id order, in-halves in
 * the push-swap order of the reference's Block::insert_out.  src / tgt are in
 * `loc` memory.  Integer-only; the result equals gt_graph_upload of the
 * host-built adjacency word for word. */
gt_status gt_graph_build_edges(uint64_t graph_id, uint32_t n, uint64_t m,
                               const uint32_t *src, const uint32_t *tgt,
                               int32_t loc, gt_graph **out);

gt_status gt_graph_free(gt_graph *g);

/* Sizes: vertex_bound, live edge count, edge_bound. */
gt_status gt_graph_info(const gt_graph *g, uint32_t *n, uint64_t *m,
                        uint64_t *edge_bound, uint64_t *graph_id);

/* Copy the adjacency back to host buffers (same layout as gt_graph_upload);
 * any pointer may be NULL to skip that array. */
gt_status gt_graph_download(const gt_graph *g, uint32_t *in_off,
                            uint32_t *in_other, uint32_t *in_eid,
                            uint32_t *out_off, uint32_t *out_other,
                            uint32_t *out_eid);

/* Use this HIP stream (a hipStream_t, e.g. torch's current stream) for every
 * later call on g; NULL restores the graph's own stream. */
gt_status gt_graph_set_stream(gt_graph *g, void *hip_stream);

/* ------------------------------------------------------------- centrality */

/* PageRank in FP64: the drop-in for pagerank / pagerank_with.
 * pers == NULL means the uniform 1/num_vertices of `pagerank`.
 * On return *iters is the sweep count the reference returns.  Unfiltered
 * solves, unweighted or with finite non-negative weights, sum their rows on
 * block tiles in two-word fixed point (to about 2^-100 of the iterate's L1
 * mass, the same bits on every run); tuning block_f64 = -1 selects the wave
 * engine's FP64 row sums. */
gt_status gt_pagerank_f64(gt_graph *g, const gt_view *view,
                          const gt_prop *weight, const gt_prop *pers,
                          double damping, const gt_power_iteration *params,
                          gt_vertex_out *rank, uint64_t *iters);

/* Extension.  Same call, computed with FP32 values, weights and SpMV
 * accumulation while the dangling-mass sum, the teleport update and the L1
 * convergence test stay FP64, followed by one FP64 sweep over the last FP32
 * contributions (FP64 sums, teleport update and output); rank is FP64.
 * The FP32 sweeps stop when delta < max(epsilon, 2^-22 * sum|x|): below
 * that the L1 change of an FP32 iterate is rounding noise, so an epsilon
 * under about 2.4e-7 (for a unit-mass rank) is not reached -- at R-MAT 22
 * with epsilon 1e-9 the result lies within L1 7e-8 of the FP64 fixed point.
 * No refinement sweep follows a run that hit max_iter or whose delta turned
 * NaN (the FP32 iterate is returned, widened).
 * *iters counts every sweep performed (FP32 and FP64). */
gt_status gt_pagerank_mixed(gt_graph *g, const gt_view *view,
                            const gt_prop *weight, const gt_prop *pers,
                            double damping, const gt_power_iteration *params,
                            gt_vertex_out *rank, uint64_t *iters);

/* No counterpart in the reference: gt_pagerank_f64 for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  It calls no hipMalloc / hipFree: its
 * iterates, contributions, partials and loop state live in a library-owned
 * 64 MiB static device array, one per device, which limits it to about
 * 1.6 M vertices (GT_ERR_UNSUPPORTED beyond, with the size in the message).
 * Serves Directed and Reversed views, Unity (or null) weights, a null,
 * constant or FP64 device personalisation and a device rank map; the graph's
 * sweep schedule for that view must already exist (one gt_pagerank_f64 call
 * on it builds it).  It runs the wave engine: results equal those of
 * gt_pagerank_f64 with tuning block_f64 = -1 bit for bit, and its default
 * block-tile solve's within the drop-in's bar.  Calls that overlap on one
 * device share the array and race. */
gt_status gt_pagerank_f64_noalloc(gt_graph *g, const gt_view *view,
                                  const gt_prop *weight, const gt_prop *pers,
                                  double damping, const gt_power_iteration *params,
                                  gt_vertex_out *rank, uint64_t *iters);

/* Extension.  k personalised PageRank vectors solved together as one SpMM per
 * sweep: column c is pagerank_with with the one-hot personalisation
 * e_{seeds[c]}, started from it, stopped on its own L1 change (< epsilon) and
 * frozen after, so iters[c] is the sweep count of that single-vector run.
 * FP32 state and SpMM accumulation, FP64 teleport update, per-column dangling
 * mass and L1 test.  x is [n][ldx] row-major FP32 in x_loc memory; columns
 * [0, k) are written.  Directed and reversed views. */
gt_status gt_pagerank_batched(gt_graph *g, const gt_view *view, const gt_prop *weight,
                              const uint32_t *seeds, uint32_t k, double damping,
                              const gt_power_iteration *params, float *x, uint64_t ldx,
                              int32_t x_loc, uint64_t *iters);

/* Extension.  Gradient of a loss through the PageRank fixed point by implicit
 * differentiation: solves lambda = dl_dx + d M^T lambda + d (pers . lambda) q
 * (M[t,s] = w_e / deg_s, q = sink indicator) with the same power iteration on
 * the opposite orientation, stores no forward iterate, then
 *   grad_w[e]     = d * x_s / deg_s * (lambda_t - (M^T lambda)_s)   (0 if deg_s = 0)
 *   grad_pers[v]  = ((1 - d) + d * sum_{sinks} x) * lambda_v
 *   grad_damping  = lambda . (x - pers) / d.
 * x is the forward solution and dl_dx = dL/dx, both DENSE vertex maps.
 * precision GT_F64 or GT_F32 (FP32 values with FP64 L1 test and sums).
 * grad_w: [edge_bound] of grad_w_dtype, grad_pers: [n] f64, both in grad_loc
 * memory and optional (NULL); grad_damping is a host scalar (optional).
 * Directed and reversed views. */
gt_status gt_pagerank_backward(gt_graph *g, const gt_view *view, const gt_prop *weight,
                               const gt_prop *pers, double damping,
                               const gt_power_iteration *params, const gt_prop *x,
                               const gt_prop *dl_dx, int32_t precision, void *grad_w,
                               int32_t grad_w_dtype, int32_t grad_loc, double *grad_pers,
                               double *grad_damping, uint64_t *iters);

/* Extension (no counterpart in the reference): an uneven batch of independent
 * graphs solved together.  g holds them as the vertex ranges
 * [vertex_offsets[k], vertex_offsets[k+1]) of one block-diagonal graph (an edge
 * joining two ranges is an error); vertex_offsets has num_graphs + 1 entries
 * from 0 to the vertex bound.  Graph k runs pagerank (pers = 1 / its vertex
 * count) with damping[k], in FP32 values, weights and SpMV accumulation with
 * FP64 dangling mass, teleport update and L1 test; it stops when its own L1
 * change drops below max(epsilon, 2^-22) (the FP32 floor of a unit-mass
 * iterate) or at max_iter sweeps, and is frozen after.  params->plan.len must
 * equal the vertex bound.  rank: [n] f64 in rank_loc memory (overwritten,
 * cold start); iters: [num_graphs] host.  The sweeps run as replays of a
 * captured HIP graph of 8 sweeps each, on block tiles cut at the graphs'
 * bounds (each graph's row sums in 64-bit fixed point at its own scale)
 * unless the tuning turns them off or a weight is negative or not finite,
 * then on the wave engine. */
gt_status gt_pagerank_uneven(gt_graph *g, uint32_t num_graphs, const uint32_t *vertex_offsets,
                             const double *damping, const gt_prop *weight,
                             const gt_power_iteration *params, double *rank, int32_t rank_loc,
                             uint64_t *iters);

/* Extension.  The adjoint of gt_pagerank_uneven for losses L = sum_k c_k . x_k:
 * per graph, the transposed system of gt_pagerank_backward with damping[k],
 * FP32 values, stopped when its L1 change drops below
 * max(epsilon, 2^-22) * ||dl_dx restricted to the graph||_1.  x (the forward
 * result) and dl_dx are [n] f64, grad_w [edge_bound] f32 and grad_pers [n] f64
 * (both optional), all in `loc` memory; grad_damping [num_graphs] and iters
 * [num_graphs] are host arrays. */
gt_status gt_pagerank_uneven_backward(gt_graph *g, uint32_t num_graphs,
                                      const uint32_t *vertex_offsets, const double *damping,
                                      const gt_prop *weight, const gt_power_iteration *params,
                                      const double *x, const double *dl_dx, int32_t loc,
                                      float *grad_w, double *grad_pers, double *grad_damping,
                                      uint64_t *iters);

/* Eigenvector centrality, FP64: the drop-in for eigenvector.  *eig is the
 * last norm (the eigenvalue estimate), *iters the sweep count. */
gt_status gt_eigenvector_f64(gt_graph *g, const gt_view *view,
                             const gt_prop *weight,
                             const gt_power_iteration *params,
                             gt_vertex_out *x, double *eig, uint64_t *iters);

/* ------------------------------------------------------------- betweenness */

/* Output maps of gt_betweenness: vertex_bc [vertex_bound] and edge_bc
 * [edge_bound] f64 in `loc` memory (DenseProp<f64, VertexTag / EdgeTag> after
 * sized_for).  edge == NULL skips the edge map (extension). */
typedef struct gt_bc_out {
    double *vertex;
    uint64_t vertex_len;
    double *edge;
    uint64_t edge_len;
    int32_t loc;             /* gt_loc */
    int32_t reserved;
    uint64_t vertex_graph_id;  /* owner graph of each map; 0 = not checked */
    uint64_t edge_graph_id;
} gt_bc_out;

/* Extension: what one gt_betweenness call did. */
typedef struct gt_bc_stats {
    uint64_t sources;          /* pivots with at least one out-edge (the others add nothing) */
    uint64_t batches;          /* groups of 64 sources run together */
    uint64_t batches_rerun;    /* batches whose BFS went deeper than the levels launched
                                  for them, run again with the host in the loop */
    uint64_t levels;           /* BFS levels over all batches */
    uint64_t forward_entries;  /* adjacency entries the forward level kernels read */
    uint64_t forward_hits;     /* (entry, source) pairs that added a path count */
    uint64_t forward_hit_entries;   /* entries with at least one such source */
    uint64_t backward_entries; /* adjacency entries the backward level kernels read */
    uint64_t backward_hits;    /* (entry, source) pairs that added a dependency */
    uint64_t backward_hit_entries;  /* entries with at least one such source */
    double forward_ms;         /* device time of the forward launches (profiling only) */
    double backward_ms;        /* device time of the backward launches (profiling only) */
    uint64_t streams;          /* streams the batches ran on (1 when profiling) */
} gt_bc_stats;

/* betweenness (crates/gt-algo/src/centrality.rs:614-689): Brandes over the
 * view's out_edges (any of the six view kinds; undirected = the whole block),
 * unweighted, unnormalised, both maps zero-filled first and halved on an
 * undirected view; errors "betweenness: vertex/edge map does not belong to
 * this graph" as the reference panics (:644-651).  pivots == NULL (npivots 0)
 * means every vertex of the view, the reference's only mode; a pivot list is
 * an extension that restricts the sources (each must be a vertex of the
 * view).  stats may be NULL. */
gt_status gt_betweenness(gt_graph *g, const gt_view *view, const uint32_t *pivots,
                         uint64_t npivots, const gt_bc_out *out, gt_bc_stats *stats);

/* ------------------------------------------------------------- degrees */

/* The degree kernels of crates/gt-algo/src/degree.rs:122-237, on the device.
 * which: 0 = weighted_out_degree, 1 = weighted_in_degree, 2 = weighted_degree
 * (in + out).  weight UNITY: the degree; CONSTANT c: degree * c (0 without
 * incident edges); DENSE: the sum in storage order (the total walks the
 * out-half, then the in-half).  out: [n] f64 in out_loc memory; vertices a
 * filtered view drops get 0.  Views: the four directed kinds and Undirected
 * (whose out-degree is the whole block; its in-degree is refused). */
gt_status gt_degrees(gt_graph *g, const gt_view *view, int32_t which, const gt_prop *weight,
                     double *out, int32_t out_loc);
/* out_degrees (degree.rs:199-216): [n] int64, every slot written. */
gt_status gt_out_degrees(gt_graph *g, const gt_view *view, int64_t *out, int32_t out_loc);
/* degree_histogram (degree.rs:230-247): hist[d] = number of the view's
 * vertices of total degree d.  *len = max degree + 1 (0 without vertices);
 * with hist == NULL or cap < *len only *len is written. */
gt_status gt_degree_histogram(gt_graph *g, const gt_view *view, uint64_t *hist, uint64_t cap,
                              uint64_t *len);

/* ------------------------------------------------------------- input formats */

/* A parsed graph file: the edge list in edge-id order, property columns and
 * (for symbolic edge lists) vertex names.  Build the device graph from it
 * with gt_graph_from_doc. */
typedef struct gt_doc gt_doc;

/* graph-tool's binary .gt format (crates/gt-io/src/gt.rs:555 `read`, layout
 * :1-58): the adjacency in file order (edge i = the i-th entry), every
 * property column (domain 0 graph / 1 vertex / 2 edge, value-type index of
 * graph-tool's value_types: 0 bool, 1 int16, 2 int32, 3 int64, 4 double,
 * 5 long double, 6 string, 7..13 vectors of those, 14 python object).
 * Truncated or malformed input is an error with the byte offset. */
gt_status gt_read_gt(const void *bytes, uint64_t len, gt_doc **out);

/* A delimited edge list (crates/gt-io/src/csv.rs:129 `read_edge_list`):
 * Python's default csv dialect, fields stripped, blank records skipped,
 * `header` drops the first record; `symbolic` interns endpoint names in
 * first-seen order (source before target), otherwise endpoints are indices
 * and the graph has max + 1 vertices.  Record i is edge i.  Errors name the
 * line a record started on. */
gt_status gt_read_edge_list(const void *bytes, uint64_t len, uint8_t delimiter, int32_t header,
                            int32_t symbolic, gt_doc **out);

gt_status gt_doc_free(gt_doc *d);
gt_status gt_doc_info(const gt_doc *d, uint64_t *n, uint64_t *m, int32_t *directed,
                      uint32_t *nprops, uint64_t *nnames);
/* src / tgt: m entries each, in edge-id order. */
gt_status gt_doc_edges(const gt_doc *d, uint32_t *src, uint32_t *tgt);
gt_status gt_doc_comment(const gt_doc *d, const char **s, uint64_t *len);
/* Property k: name (not NUL-terminated), domain, value-type index, element count. */
gt_status gt_doc_prop_info(const gt_doc *d, uint32_t k, const char **name, uint64_t *name_len,
                           int32_t *domain, int32_t *kind, uint64_t *count);
/* A numeric scalar column (bool, int16, int32, int64, double) as f64. */
gt_status gt_doc_prop_f64(const gt_doc *d, uint32_t k, double *out);
/* Element i of any other column: strings and python objects as their bytes,
 * long doubles as 16 bytes, numeric vectors as packed f64, vectors of long
 * double as packed 16-byte values, vectors of strings as a u64 count then
 * (u64 length, bytes) per string. */
gt_status gt_doc_prop_cell(const gt_doc *d, uint32_t k, uint64_t i, const void **data,
                           uint64_t *len);
gt_status gt_doc_vertex_name(const gt_doc *d, uint64_t v, const char **s, uint64_t *len);
/* The device graph of the document's edge list (gt_graph_build_edges). */
gt_status gt_graph_from_doc(const gt_doc *d, uint64_t graph_id, gt_graph **out);

/* ------------------------------------------------------------- tuning */

/* Extension: launch and schedule parameters of one graph, for an autotuner.
 * A zero field means the built-in default.  Changing tile_rows or mm_tile_nnz
 * rebuilds the affected schedules on the next call. */
typedef struct gt_tuning {
    int32_t sweep_blocks_per_cu; /* resident workgroups per CU, pull sweeps: 1..8 (3) */
    int32_t tile_rows;           /* rows per wave tile: 16..256 (256) */
    int32_t short_row;           /* rows with fewer entries are summed by one lane: 1..1024 (64) */
    int32_t mm_tile_nnz;         /* batched SpMM: entries per tile, 256..8192 (1024) */
    int32_t mm_blocks_per_cu;    /* batched SpMM: resident workgroups per CU: 1..8 (8) */
    int32_t sweep_hot_kb;        /* FP32 PageRank sweeps on a relabelled graph: KB of the most
                                    gathered contributions held in LDS, one workgroup per CU;
                                    -1 turns it off: -1, 1..128 (128) */
    int32_t sliced_min_len;      /* unweighted PageRank sweeps on a relabelled graph: rows with
                                    at least this many entries gather their entries in the
                                    "warm" slices (the sweep_hot_kb-sized slices after the hot
                                    one) from LDS copies held by dedicated workgroups: 2..1024,
                                    or 0 / -1 for off (off) */
    int32_t sliced_share;        /* percent of the sliced sweep's workgroups given to the warm
                                    slices: 1..99 (0: from a cost model) */
    int32_t block_tile_nnz;      /* PageRank, adjoint and eigenvector sweeps on a relabelled,
                                    unfiltered graph, and the uneven batch's sweeps, unweighted
                                    or with finite non-negative weights: entries per block tile
                                    (rows sorted together by source, summed in 64-bit fixed
                                    point); -1 turns the block sweeps off and the wave engine
                                    runs: -1, 1024..1048576 (131072) */
    int32_t block_long_min;      /* block tiles: rows with more entries than this are summed in
                                    lane-private slots (at most 32 per tile): 16..65536 (2048) */
    int32_t block_per_cu;        /* FP32 block sweeps: resident workgroups per CU: 1..3 (2) */
    int32_t block_f64;           /* the FP64 PageRank drop-in's sweeps on block tiles (two-word
                                    fixed point): 1 on, -1 off (1) */
} gt_tuning;

gt_status gt_graph_set_tuning(gt_graph *g, const gt_tuning *t);
/* The values in effect (defaults filled in). */
gt_status gt_graph_get_tuning(const gt_graph *g, gt_tuning *t);

/* ------------------------------------------------------------- profiling */

/* Sweep kernels, for gt_sweep_stats. */
typedef enum gt_sweep_kind {
    GT_SWEEP_PAGERANK_F64 = 0,
    GT_SWEEP_PAGERANK_F32 = 1,
    GT_SWEEP_EIGENVECTOR = 2,
    GT_SWEEP_OTHER = 3,
    GT_SWEEP_BATCHED = 4,   /* personalised SpMM sweep */
    GT_SWEEP_ADJOINT = 5,   /* transposed (adjoint) sweep */
    GT_SWEEP_BC_FORWARD = 6,  /* betweenness: forward (path-count) level launches */
    GT_SWEEP_BC_BACKWARD = 7  /* betweenness: backward (dependency) level launches */
} gt_sweep_kind;
#define GT_SWEEP_KINDS 8

/* Extension.  While enabled, every sweep launch on g is bracketed by HIP
 * events recorded on the stream it runs on; gt_sweep_stats reports, for one
 * kernel kind, the summed device time and the launch count (launches that
 * found the loop finished included), and the sweep count of the last call.
 * Enabling resets the sums. */
gt_status gt_graph_set_profiling(gt_graph *g, int32_t enable);
gt_status gt_sweep_stats(const gt_graph *g, int32_t kind, double *total_ms,
                         uint64_t *launches, uint64_t *last_sweeps);

/* Extension.  How the last pull-sweep solve on g (PageRank, eigenvector,
 * adjoint or uneven batch) laid out the orientation it swept.  slot_bits is
 * 10 when every wave tile's entries are packed as (source - base) << 10 |
 * slot (each tile's sources span fewer than 2^22 labels: always below
 * 2^22 vertices), 0 when they are plain 32-bit sources in row order.
 * sliced_rows / sliced_entries / slices describe the rows whose entries are
 * gathered from on-chip slices of the source vector (0 when none were).
 * block_tiles / block_long_rows / block_groups describe the block tiles the
 * FP32 sweeps ran on (0 when they ran on the wave engine). */
typedef struct gt_sweep_layout {
    uint32_t slot_bits;
    uint32_t tiles;
    uint32_t split_rows;
    uint32_t slices;
    uint64_t sliced_rows;
    uint64_t sliced_entries;
    uint32_t block_tiles;
    uint32_t block_long_rows;
    uint64_t block_groups;
} gt_sweep_layout;
gt_status gt_last_layout(const gt_graph *g, gt_sweep_layout *out);

/* ------------------------------------------------------------- host side */

/* Extension: seeded workload generators (counter-based SplitMix64, the
 * generator of crates/gt-algo/tests/u16_centrality.rs:58-72).
 * gt_gen_er: per edge source = next % n, then target = next % n
 * (crates/gt-algo/benches/kernels.rs:68-70).
 * gt_gen_rmat: one 53-bit uniform per bit level, most significant first,
 * quadrants (a, b, c, 1-a-b-c) over (source bit, target bit). */
gt_status gt_gen_er(uint32_t n, uint64_t m, uint64_t seed, uint32_t *src, uint32_t *tgt);
gt_status gt_gen_rmat(uint32_t scale, uint64_t m, double a, double b, double c,
                      uint64_t seed, uint32_t *src, uint32_t *tgt);

/* A batch of directed Chung-Lu graphs (uneven-batch workload).  Graph k draws
 * from the stream seeded with mix(seed + k): n_k = round(exp(U[ln n_min,
 * ln n_max])), damping_k = d_lo + (d_hi - d_lo) U, m_k = round(mean_degree *
 * n_k) edges whose source and target are drawn independently with
 * probability proportional to (v + 1)^(-1 / (exponent - 1)) over the graph's
 * vertices, weight 1 - U in (0, 1].  Vertex and edge ids are global (graph k
 * after graphs 0..k-1).  With src == NULL only n_g, m_g and damping are
 * written; then call again with sum(m_g)-entry src, tgt and weight. */
gt_status gt_gen_chung_lu_batch(uint32_t num_graphs, uint64_t seed, uint32_t n_min,
                                uint32_t n_max, double mean_degree, double exponent,
                                double d_lo, double d_hi, uint32_t *n_g, uint64_t *m_g,
                                double *damping, uint32_t *src, uint32_t *tgt, float *weight);

/* The host adjacency AdjList::add_edge(src[i], tgt[i]) for i = 0..m builds
 * (crates/gt-core/src/adj/list.rs:258-280, 471-500), flattened into the
 * arrays gt_graph_upload takes (in_off / out_off hold n+1 entries, the rest m). */
gt_status gt_host_adjacency(uint32_t n, uint64_t m, const uint32_t *src,
                            const uint32_t *tgt, uint32_t *in_off, uint32_t *in_other,
                            uint32_t *in_eid, uint32_t *out_off, uint32_t *out_other,
                            uint32_t *out_eid);

#ifdef __cplusplus
}
#endif

#endif /* GT_HIP_H */
