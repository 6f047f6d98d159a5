// Block-tile sweeps for the FP32 PageRank phase (relabelled, unfiltered,
// unweighted).  The layout is Blocks (gt_internal.hpp); the kernels and
// their launch helpers are in blocks.hip.
//
// A tile's entries are sorted by source across all its rows, so the 64 lanes
// of a gather instruction read 64 consecutive entries of that order: with
// 65,536 entries per tile the sources of one instruction are dense enough
// that many lanes share a 128-byte line (R-MAT scale 22: about 17 M line
// requests per sweep, against 29 M for 1,024-entry wave tiles with the
// 128 KB hot slice).  Row order is lost inside a tile, so a row's entries
// are added into its LDS accumulator as they come.  The accumulators are
// 64-bit fixed point: a contribution c is rounded to c * 2^K (K from the
// mass of the gathered iterate, so that no partial sum can reach 2^62) and
// added as an integer.  Integer addition does not depend on order, so the
// row sums -- and the sweep -- are the same bits on every run whatever the
// order the workgroup's wavefronts add in, and they carry 2^-61 of the
// iterate's L1 mass absolute precision, finer than the FP32 row sums of the
// wave engine.
#pragma once

#include "ev_state.hpp"
#include "gt_internal.hpp"

namespace gt {

// The sweeps on block tiles: the FP32 form's FP32 sweep, its FP64
// refinement (FP32 contributions gathered, FP64 output), the FP64 PageRank
// sweep (FP64 contributions, each term held in two 64-bit words: round(c
// 2^K) and the rest at 2^(K + 38), so the sums keep about 100 bits of the
// iterate's mass), the eigenvector sweep (the FP64 iterate gathered, in two
// words), the adjoint's transposed sweep in FP32 (one word) and FP64
// (two words), and the uneven batch's forward and adjoint sweeps in FP32
// (one word, per-graph damping, scale and loop state).
enum : int {
    kBlkF32 = 0,
    kBlkRefine = 1,
    kBlkF64 = 2,
    kBlkEv = 3,
    kBlkAdj32 = 4,
    kBlkAdj64 = 5,
    kBlkUFwd = 6,
    kBlkUAdj = 7
};
// The two-word sweeps' low words stay below 2^62 for rows of at most this
// many entries; a graph with a longer row runs them on the wave engine.
constexpr uint32_t kBlkF64MaxRow = 1u << 23;

// Device state of a block-sweep loop.
struct BlkState {
    double mass;         // bound on the sum of |terms| a gathered value feeds (see blk_scale)
    uint32_t next_tile;  // tile counter of the running sweep
    uint32_t pad;        // set by blk_stage_weights on a weight it cannot use
    unsigned long long wmax;  // largest staged weight, as its bits (adjoint's mass bound)
};

// A mass above this makes the sweep give up (SweepState::bail): the loop is
// run again on the wave engine, whose FP32 sums follow the values to
// overflow as the reference's would.
constexpr double kBlkMassMax = 0x1p900;

// A batch of graphs on tiles cut at the graphs' bounds (each tile in one
// graph): per-graph parameters and loop state, and the list of the tiles of
// the graphs still iterating.  Sweep k takes list entries through ctr and
// clears ctr_next and list_n_next for sweep k + 1; the fold after it lists
// the tiles of each graph that goes on.
struct BlkUneven {
    const uint32_t *tile_graph;  // [ntiles]
    const uint32_t *voff;        // [G + 1]
    const double *damp;          // [G]
    double *mass;                // [G]: bound on the terms of the graph's next sweep
    double *ps;                  // [G]: dangling mass (forward) / p . lambda (adjoint)
    double *delta;               // [G]
    const double *scale;         // [G]: adjoint, ||dL/dx_g||_1
    uint64_t *iter;              // [G]
    uint32_t *done;              // [G]
    uint32_t *active;            // graphs still iterating; bail right after it
    const uint32_t *tb, *lb;     // [G + 1]: first tile and first long row of each graph
    uint32_t *list;
    const uint32_t *list_n;
    uint32_t *list_n_next;
    uint32_t *ctr, *ctr_next;
    uint32_t ngraphs;
};

struct BlkArgs {
    const BlkTile *tiles;
    uint32_t ntiles;
    uint32_t nlong;
    uint32_t qlast;             // last block of groups (loads ahead of a tile's end clamp to it)
    uint32_t n;                 // rows
    const uint32_t *words;
    const uint32_t *gbase;
    const uint32_t *ltab;
    const uint32_t *long_rows;
    const uint32_t *odeg;       // out-degree, sweep labels
    unsigned long long *slots;  // [nlong] int64 sums of the long rows
    unsigned long long *slots_lo;  // [nlong] their low words (FP64 sweep), else null
    double *part;               // [3 * ntiles]: delta, dangling, mass per tile
    double *fold_part;          // [3 * fold grid]
    BlkState *bs;
    SweepState *st;             // PageRank state; the eigenvector's EvState (same offsets
                                // of done, bail and counter) for the eigenvector sweep
    EvState *ev;                // eigenvector sweep: its state, else null
    const double *pers;         // dense personalisation, sweep labels (DPERS)
    double pval;
    double damping;
    double eps_abs, eps_rel;    // stop when delta < max(eps_abs, eps_rel * initial mass)
    uint64_t max_iter;
    const float *x_cur;
    float *x_next;
    const float *c_cur;
    float *c_next;
    double *x_next64;           // the refinement's and the FP64 sweeps' output
    const double *x_cur64;      // FP64 sweeps: iterate and contributions (for the
    const double *c_cur64;      // eigenvector both y_{L-1})
    double *c_next64;
    const double *x_prev64;     // eigenvector: y_{L-2} (y_{L-1} on the first launch)
    // Weighted PageRank sweeps: the weights in block order (FP32 for the FP32
    // sweep, FP64 for the refinement and the FP64 sweep), null when
    // unweighted; the weighted out-degree in sweep labels.
    const void *wblk;
    const double *wdeg;
    // The adjoint's sweep: dL/dx (sweep labels), the rows' offsets (their
    // unweighted degree), and, for the pass that computes mu = M^T lambda
    // instead of iterating, its output (null otherwise).
    const double *cvec;
    const uint32_t *row_off;
    double *mu_out;
    BlkUneven u;                // the uneven batch's sweeps, else zero
};

// Build (once per tile size and long-row threshold) the block tiles of R's
// pull orientation.  Waits for the stream.
gt_status ensure_blocks(gt_graph *g, Relabeled &R, uint32_t tile_nnz, uint32_t long_min,
                        Blocks **out);
// The block tiles of orientation o cut at `bounds` as well (no tile holds
// rows on both sides of a bound), built once per tile size, threshold and
// bounds into o.blk.  `gathered_off` (device, n + 1) gives the times each
// vertex is gathered as its row length.  Rows need not be sorted.  Waits
// for the stream.
gt_status ensure_blocks_cut(gt_graph *g, Orient &o, const uint32_t *gathered_off,
                            uint32_t tile_nnz, uint32_t long_min,
                            const std::vector<uint32_t> &bounds, Blocks **out);
// A weighted solve's weights in block order into `out` (FP64 when `f64`,
// else FP32), from the map's device source (stage_edge_source).  Sets
// *usable to whether every weight is finite and non-negative: the block
// sweeps bound their fixed-point sums by the iterate's mass, which other
// weights break, so such a solve runs on the wave engine.  Waits for the
// stream.
gt_status blk_stage_weights(gt_graph *g, const Blocks &B, const gt_prop *w, const void *src,
                            int dtype, bool f64, DevBuf &out, bool *usable);
// The FP32 block-order weights of an f32 map (as blk_stage_weights wrote
// them) widened to FP64: the same values as staging the map in FP64 again,
// without the gather through the edge ids.
gt_status blk_widen_weights(gt_graph *g, const Blocks &B, const DevBuf &f32, DevBuf &out);
// Sizes the per-call buffers of a block loop over B into the workspace.
gt_status blk_reserve(gt_graph *g, const Blocks &B);
// The layout and workspace fields of the sweep arguments over B (after
// blk_reserve); the caller adds the loop's state, vectors and parameters.
// Also reports B in g->layout.
BlkArgs blk_layout_args(gt_graph *g, const Blocks &B);
// After the init kernel: zero the long-row sums, take the initial mass,
// clear the bail flag.  The eigenvector's and the adjoint's mass (sum of
// |v| * the times v is gathered, over the start vector v = b.x_cur64, times
// the largest weight when weighted) is a reduction in a fixed order.
void blk_launch_start(const BlkArgs &b, hipStream_t s);
void blk_launch_start_ev(const BlkArgs &b, hipStream_t s);
// One sweep of `mode` (workgroups: `per_cu` per compute unit for the FP32
// and refinement sweeps, one for the two-word sweeps; at most one per tile).
// The refinement runs whatever the loop state says.
void blk_launch_sweep(const BlkArgs &b, bool dpers, int per_cu, int mode, hipStream_t s);
// Completes a sweep: long rows, partials in tile order, convergence.
void blk_launch_fold(const BlkArgs &b, bool dpers, int mode, hipStream_t s);
// The uneven batch (mode kBlkUFwd or kBlkUAdj): each graph's first mass
// from its start iterate x0 (sum of |x0| times the gathered count when
// `gathered`, times the largest weight when b.wblk), long-row sums cleared;
// then one sweep, and the per-graph fold (one wavefront a graph).
void blk_launch_ustart(const BlkArgs &b, const float *x0, bool gathered, hipStream_t s);
void blk_launch_usweep(const BlkArgs &b, int per_cu, int mode, hipStream_t s);
void blk_launch_ufold(const BlkArgs &b, int mode, hipStream_t s);

}  // namespace gt
