// Internal declarations shared by the translation units of libgt_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "gt_hip.h"

namespace gt {

// ---------------------------------------------------------------- errors

// Per-thread message behind gt_last_error().
void set_error(const std::string &msg);
void clear_error();

struct Status {
    gt_status code = GT_OK;
};

// Early-return helpers.  GT_TRY propagates a gt_status; GT_HIP converts a
// failed HIP call into GT_ERR_HIP (GT_ERR_OOM for allocation failures).
#define GT_TRY(expr)                          \
    do {                                      \
        gt_status _st = (expr);               \
        if (_st != GT_OK) return _st;         \
    } while (0)

gt_status hip_fail(hipError_t e, const char *what, const char *file, int line);

#define GT_HIP(call)                                                   \
    do {                                                               \
        hipError_t _e = (call);                                        \
        if (_e != hipSuccess) return ::gt::hip_fail(_e, #call, __FILE__, __LINE__); \
    } while (0)

gt_status fail(gt_status code, const std::string &msg);

// ---------------------------------------------------------------- memory

// Device bytes the library holds (every DevBuf), for the memory limit.
void devmem_released(size_t bytes);

// Owning device buffer (hipMalloc / hipFree).
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (ptr) {
            (void)hipFree(ptr);
            devmem_released(bytes);
        }
        ptr = nullptr;
        bytes = 0;
    }
    // Grow to at least `b` bytes (contents are not preserved).
    gt_status reserve(size_t b);
    template <typename T> T *as() const { return static_cast<T *>(ptr); }
};

// ---------------------------------------------------------------- schedule

// One unit of SpMV work for a wavefront (pull sweeps) or a workgroup (SpMM).
// Either a run of whole rows [row_begin, row_end) whose in-entries are
// [nnz_begin, nnz_end), or -- when row_end == kSplit -- one chunk
// [nnz_begin, nnz_end) of a long row whose partial sum goes to slot
// row_begin.
struct Tile {
    uint32_t row_begin;
    uint32_t row_end;
    uint32_t nnz_begin;
    uint32_t nnz_end;
};
constexpr uint32_t kSplit = 0xffffffffu;
// Edge-list slot of an edge id with no edge.
constexpr uint32_t kNoEdge = 0xffffffffu;

// Pull-sweep wave tiles: at most kWaveTileNnz entries (16 per lane) and
// kWaveTileRows rows (4 per lane); longer rows are split into chunks of
// kWaveTileNnz entries.  An entry is packed as ((source - base) <<
// kWaveSlotBits) | slot, base being the tile's smallest source, when every
// tile's sources span fewer than 2^(32 - kWaveSlotBits) labels.
constexpr uint32_t kWaveTileNnz = 1024;
constexpr uint32_t kWaveTileRows = 256;
constexpr uint32_t kWaveSlotBits = 10;
constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / 64;  // wavefronts of 64 lanes

// A tiling of an orientation's rows for the pull kernels.
struct Schedule {
    DevBuf tiles;            // Tile[ntiles]
    DevBuf long_rows;        // uint32[nlong]: row of each split row
    DevBuf long_slot;        // uint32[nlong+1]: first partial slot of each
    uint32_t ntiles = 0, nlong = 0, nslots = 0;
    uint32_t tile_nnz = 0;   // entry bound the schedule was built with
    uint32_t tile_rows = 0;  // row bound (SpMM schedules)
    bool ready = false;
};

// SpMM tiles: each in-entry gathers a 64-column row (256 bytes), so tiles are
// smaller and long rows split sooner.  Default entries per tile (at most a
// quarter as many rows).
constexpr uint32_t kMmTileNnz = 1024;

// Row tiles of the betweenness level kernels: pieces of the rows longer than
// tile_nnz entries first (row, kSplit, first, end; pieces of one row adjacent),
// then runs of at most 64 whole rows and tile_nnz entries (row_begin,
// row_end, first, end).
struct BcSchedule {
    DevBuf tiles;            // uint4[ntiles]
    DevBuf long_rows;        // uint32[nlong]
    DevBuf long_first;       // uint32[nlong + 1]: first piece of each long row
    uint32_t ntiles = 0, npieces = 0, nlong = 0;
    uint32_t tile_nnz = 0;
    bool ready = false;
};

// Block tiles of an orientation (the block sweeps: a relabelled pull
// orientation, or an orientation cut at the bounds of a batch of graphs):
// each tile is a run of rows [r0, r1) with at most tile_nnz entries, its
// entries sorted by source across all its rows and cut into groups of 64
// (one gather instruction), padded to whole blocks of kBlkQuad groups; a
// block's words are stored lane-major (lane l's words of its groups side by
// side).  A word is ((source - gbase[group]) << 13) |
// at, where `at` is the LDS accumulator the entry adds into: the row's,
// r - r0 < kBlkRows, for a short row; for one of the tile's long rows (more
// than long_min entries, at most kBlkLongMax per tile, the j-th) the slot of
// the entry's lane in that row's lane slots, kBlkAcc + 64 j + (lane + j) %
// 64; kBlkNull (a dummy) for a padding entry.  A row longer than tile_nnz is
// cut into chunk tiles of its own (r0 == r1).  Long rows are finished by the
// fold from int64 slots (long_rows[k], slot k; ltab lists each tile's).
constexpr uint32_t kBlkCodeBits = 13;
constexpr uint32_t kBlkRows = 4032;      // rows in a tile's range
constexpr uint32_t kBlkAcc = 4096;       // row accumulators and dummies, then lane slots
constexpr uint32_t kBlkNull = 4095;      // accumulator of the padding entries (a dummy)
constexpr uint32_t kBlkLongMax = 32;     // long rows per tile
constexpr uint32_t kBlkSpan = 1u << (32 - kBlkCodeBits);  // source span of a group
constexpr uint32_t kBlkQuad = 4;         // groups per block (one 16-byte load a lane)
constexpr uint32_t kBlkTileNnzDefault = 131072;
constexpr uint32_t kBlkLongMinDefault = 2048;

struct BlkTile {
    uint32_t r0, r1;      // rows finished by the tile (long rows excepted)
    uint32_t g0, g1;      // groups of 64 words
    uint32_t lfirst;      // first entry of ltab
    uint32_t nlong;       // long rows
    uint32_t pad[2];
};

struct Blocks {
    bool ready = false;
    uint32_t tile_nnz = 0, long_min = 0;
    DevBuf tiles;        // BlkTile[ntiles]
    DevBuf words;        // uint32[ngroups * 64]
    DevBuf gbase;        // uint32[ngroups]
    DevBuf ltab;         // uint32[]: long-row slot of each tile's long code
    DevBuf long_rows;    // uint32[nlong]
    DevBuf odeg;         // uint32[n]: times each vertex is gathered (a pull
                         // orientation's out-degree), in sweep labels
    DevBuf eids;         // uint32[ngroups * 64]: edge id of each word (0 for padding)
    uint32_t ntiles = 0, nlong = 0;
    uint32_t max_row = 0;  // entries of the longest row
    uint64_t ngroups = 0;
    // Tiles cut at these row bounds as well (a batch of graphs), with the
    // host copies the batch needs: each tile's first row (a chunk tile's
    // row) and the long rows.
    std::vector<uint32_t> bounds;
    std::vector<uint32_t> h_tile_row, h_long_rows;
    // A batch's per-tile graph and per-graph first tile / first long row
    // (uint32 each), uploaded for the bounds in batch_bounds.
    DevBuf batch[3];
    std::vector<uint32_t> batch_bounds;
};

// One orientation of the adjacency: row r lists the entries a kernel pulls
// from (in-halves for the CSC, out-halves for the CSR), in storage order.
struct Orient {
    DevBuf off;   // uint32[n+1]
    DevBuf idx;   // uint32[m]: the other endpoint
    DevBuf eid;   // uint32[m]: edge id
    std::vector<uint32_t> h_off;  // host copy of off, for the schedule
    // Pull-sweep schedule over this orientation's rows (wave tiles).
    DevBuf tiles;            // Tile[ntiles]
    DevBuf long_rows;        // uint32[nlong]: row of each split row
    DevBuf long_slot;        // uint32[nlong+1]: first partial slot of each
    DevBuf packed;           // uint32[m]: entries in tile-sorted packed order
    DevBuf peid;             // uint32[m]: edge id of each packed entry
    DevBuf tbase;            // uint32[ntiles]: source label of packed value 0
    uint32_t ntiles = 0, nlong = 0, nslots = 0;
    uint32_t bits = 0;       // slot bits of `packed` (0: plain sources, row order)
    uint32_t sched_rows = 0; // rows-per-tile bound the schedule was built with
    bool scheduled = false;
    // Set when tiles were cut at these row bounds (a batch of graphs): the
    // host copies of the tile list and the split rows.
    std::vector<uint32_t> bounds;
    std::vector<Tile> h_tiles;
    std::vector<uint32_t> h_long_rows;
    // A batch's per-tile graph and per-graph first tile / first split row
    // over the wave tiles (uint32 each), uploaded for the bounds in
    // batch_bounds; cleared when the schedule is rebuilt.
    DevBuf batch[3];
    std::vector<uint32_t> batch_bounds;
    Schedule mm;             // SpMM schedule (batched personalisation)
    Blocks blk;              // block tiles cut at graph bounds (uneven batch)
    // Rows for the wavefront (mid) and workgroup (heavy) degree sums.
    DevBuf heavy, mid;
    uint32_t nheavy = 0, nmid = 0;
    bool heavy_ready = false;
    BcSchedule bcs;          // betweenness row tiles
    // Set on an orientation that merges an out-half and an in-half per row
    // (out-half first): the out-half length of each row.
    DevBuf first_len;
};

// One stretch of a workgroup's share of the sliced pass: tiles [t0, t1) of
// one slice, whose sources are the labels [base, base + len).
struct SliceSeg {
    uint32_t base, len, t0, t1;
};

// A relabelled pull orientation split for the sliced sweep.  Sources are
// cut into slices of `width` labels: slice 0 is the hot slice every light
// workgroup holds in LDS; the "warm" slices after it are each held by the
// workgroups that sum the warm pieces.  A row with at least min_len entries
// ("heavy") is cut into one warm piece per warm slice it has entries in and
// one main piece with the rest (a piece longer than kWaveTileNnz into
// chunks), each summed into its own partial.  A heavy row's slots -- main
// piece first, then warm pieces in slice order -- are the fold's split-row
// slots, read through long_idx.  Light rows and main pieces are swept by the
// launch's other workgroups, with the hot slice and global gathers.
struct Sliced {
    bool ready = false;
    bool active = false;     // some row is heavy (else the plain hot sweep runs)
    uint32_t width = 0, min_len = 0, sched_rows = 0, share = 0, grid = 0;
    // Light rows: wave tiles over runs of light rows, packed like an
    // orientation's stream at the orientation's entry positions.
    DevBuf ltiles, ltbase, lpacked, lpeid;
    uint32_t nltiles = 0, lbits = 0;
    uint64_t light_entries = 0;
    // Heavy rows, their slot ranges, and the partial each slot reads
    // (long_idx[slot]; nout partials in all).
    DevBuf long_rows, long_slot, long_idx;
    uint32_t nlong = 0, nslots = 0, nout = 0;
    // Warm pieces: wave tiles whose rows are pieces (hoff: piece offsets
    // into hpacked; a piece's sum goes to the partial of its index) or one
    // chunk (row_end == kSplit, row_begin = its partial); sources in hpacked
    // are relative to the slice.  Warm slices are 1 .. warm_end - 1.
    DevBuf htiles, htbase, hoff, hpacked, hpeid;
    uint32_t nhtiles = 0, npieces = 0, nslices = 0, warm_end = 0;
    uint64_t heavy_entries = 0;  // entries of the warm pieces
    // Main pieces: a heavy row's entries outside the warm slices, swept with
    // the light rows; a piece's sum goes to partial mpart + its index.
    DevBuf mtiles, mtbase, moff, mpacked, mpeid;
    uint32_t nmtiles = 0, mbits = 0, mpart = 0;
    // Workgroup b < heavy_wgs takes segments seg[wseg[b] .. wseg[b + 1]).
    DevBuf seg, wseg;
    uint32_t heavy_wgs = 0;
};


// A pull orientation in a locality-friendly labelling: vertices renumbered by
// descending gather frequency (the row length of the opposite orientation:
// how often a vertex appears as an entry), rows in the new order and each
// row's entries sorted by their new label, edge ids kept.  Sweeps run in the
// new labels; vectors cross the boundary through perm / iperm.
struct Relabeled {
    bool ready = false;
    DevBuf perm;       // uint32[n]: old label -> new
    DevBuf iperm;      // uint32[n]: new label -> old
    Orient pull;       // relabelled pull orientation (eid = original edge ids)
    DevBuf push_off;   // uint32[n+1]: scanned push-row lengths in new order
    Sliced sl[2];      // sliced sweeps: [0] 4-byte values, [1] 8-byte values
    Blocks blk;        // block tiles (FP32 PageRank sweeps)
};

struct Workspace {
    DevBuf x[2];        // f64 iterates
    DevBuf c[2];        // f64 contributions x/deg
    DevBuf x32[2];      // f32 iterates
    DevBuf c32[2];      // f32 contributions
    DevBuf y3;          // staged f64 map / third eigenvector buffer
    DevBuf y[3];        // eigenvector iterates
    DevBuf deg;         // weighted out-degree, f64[n]
    DevBuf deg2;        // weighted out-degree in sweep labels
    DevBuf pers2;       // personalisation in sweep labels
    DevBuf wsrc;        // staged weight map (edge-id order)
    DevBuf wcsc;        // weights in pull order, value type of the call
    DevBuf wcsc64;      // weights in pull order, f64 (refinement sweeps)
    DevBuf active;      // uint8[n]
    DevBuf partials;    // f64[2 * grid]
    DevBuf long_part;   // f64[nslots]
    DevBuf state;       // device SweepState
    DevBuf pers;        // staged personalisation
    DevBuf lam[2];      // adjoint iterates
    DevBuf ax, ac;      // adjoint: staged x* and dL/dx (f64)
    DevBuf wcsr;        // adjoint: weights in the push orientation's packed order
    DevBuf mu;          // adjoint: M^T lambda, f64[n]
    DevBuf arl[4];      // adjoint, relabelled order: dL/dx, degree, personalisation, mu
    DevBuf gw, gp;      // adjoint: staged gradients for host outputs
    DevBuf bx[2];       // batched: [n][64] f32 iterates
    DevBuf bc[2];       // batched: [n][64] f32 contributions
    DevBuf bpart;       // batched: per-workgroup, per-column partials
    DevBuf blong;       // batched: split-row [64] partials
    DevBuf bstate;      // batched: loop state + seeds
    DevBuf blk_slots;   // block sweeps: int64 sums of the long rows
    DevBuf blk_slots_lo;  // block sweeps: their low words (FP64)
    DevBuf blk_part;    // block sweeps: per-tile delta, dangling and mass partials
    DevBuf blk_state;   // block sweeps: BlkState
    DevBuf blk_w, blk_w64;  // weighted block sweeps: weights in block order (f32, f64)
    DevBuf masks;       // staged vertex + edge masks
    DevBuf ub[16];      // uneven batch: per-call graph, tile and state arrays
    DevBuf btw[10];     // betweenness: [n][64] rows, lane masks, accumulators, pieces
    DevBuf btwx[3][9];  // betweenness: the other streams' batch state and accumulators
    Orient fpull;       // filtered view: compacted pull orientation
    Orient fpush;       // filtered view: compacted push orientation
    Orient fund;        // undirected filtered view: merged halves
};

struct gt_graph_impl;

}  // namespace gt

// Device-side loop state, shared by every sweep launch of one solve.
struct SweepState {
    double p_sink;      // dangling mass entering the next sweep (PageRank)
    double delta;       // L1 change of the last sweep
    double norm;        // eigenvector: norm of the last unnormalised iterate
    double norm_prev;   // eigenvector: norm of the one before
    double aux;         // spare reduction slot
    uint64_t iter;      // sweeps completed
    uint32_t done;      // set once converged or capped
    uint32_t counter;   // arrival ticket for the last-workgroup fold
    uint32_t launches;  // launches that found work
    uint32_t bail;      // set by a block sweep that cannot represent its sums: the loop
                        // ends and the solve runs again on the wave engine
};

struct gt_graph {
    uint64_t id = 0;
    uint32_t n = 0;
    uint64_t m = 0;
    uint64_t edge_bound = 0;
    gt::Orient in;   // CSC: in-halves
    gt::Orient out;  // CSR: out-halves
    gt::Relabeled rel[2];  // [0]: pull = in-halves, [1]: pull = out-halves
    gt::Orient und;        // whole blocks (out-half then in-half), built on first use
    // Vertex offsets of the last uneven batch found to have no edge joining
    // two of its graphs.
    std::vector<uint32_t> batch_checked;
    bool und_ready = false;
    // Edge list by edge id (source, target), kNoEdge where an id has no edge;
    // built on first use from the out-halves.
    gt::DevBuf esrc, etgt;
    bool edges_ready = false;
    hipStream_t own_stream = nullptr;
    hipStream_t side_stream[3] = {};  // betweenness: the other streams, made on first use
    hipStream_t stream = nullptr;
    gt::Workspace ws;
    gt_tuning tune{};       // effective launch / schedule parameters
    bool profiling = false;
    // Summed event time and launch count of sweep launches, per kernel kind
    // (gt_sweep_kind); prof_kind is the kind being launched.
    double prof_ms[GT_SWEEP_KINDS] = {};
    uint64_t prof_launches[GT_SWEEP_KINDS] = {};
    int prof_kind = 0;
    uint64_t last_sweeps = 0;
    // Working sweeps the last loop of each kind took (first batch of the
    // next): [0, GT_SWEEP_KINDS) by gt_sweep_kind, then the FP64 refinement
    // phase of the FP32 form.
    uint64_t sweep_hint[GT_SWEEP_KINDS + 1] = {};
    std::vector<hipEvent_t> ev;  // profiling events
    gt_sweep_layout layout{};    // the last solve's swept orientation
    SweepState *h_state = nullptr;  // pinned mirror of the device state
};

namespace gt {

// Build (once) the pull-sweep schedule of an orientation: wave tiles from the
// host offsets, then the packed entry stream and its edge ids on the device.
gt_status ensure_schedule(gt_graph *g, Orient &o, uint32_t n,
                          const std::vector<uint32_t> *bounds = nullptr);
// Record `o` (scheduled) as the orientation the current solve sweeps.
void note_layout(gt_graph *g, const Orient &o);
// Build (once per tile size) the SpMM schedule of an orientation.
gt_status ensure_schedule_mm(Orient &o, uint32_t n, uint32_t tile_nnz, uint32_t tile_rows);
// Allocate a graph shell with its stream.
gt_status graph_new(uint64_t graph_id, uint32_t n, uint64_t m, uint64_t edge_bound,
                    gt_graph **out);
// Workgroups of `kernel` (256 threads) that are resident at once on the
// device, at most `per_cu` per compute unit and 2048 in all: a persistent
// tile loop launched with this many workgroups has no second wave of
// workgroups.
int occupancy_cap(const void *kernel, int per_cu = 8);
// Default resident workgroups per compute unit for the wave-engine sweeps.
// More wavefronts than 12 per CU only add contention for the L2 the gathers
// live in (R-MAT scale 22: 3 per CU is the fastest f32 and f64 sweep, 8 per
// CU is 40% slower).
constexpr int kSweepBlocksPerCU = 3;
// The sliced split of R's pull orientation for `width`-label slices (value
// size `vbytes`), built once per width, row threshold, tile rows and grid.
gt_status ensure_sliced(gt_graph *g, Relabeled &R, uint32_t width, uint32_t vbytes, int grid,
                        Sliced **out);
// Sort and pack the entries of `ntiles` wave tiles (their idx / eid at the
// tiles' entry positions) into `packed` / `peid`, each tile's sources
// relative to its smallest (tbase).  Returns the slot bits: kWaveSlotBits,
// or 0 (plain sources, bases 0) when some tile spans 2^22 labels or more,
// which is looked for only when `wide` is set.  Waits for the stream.
gt_status pack_stream(gt_graph *g, const Tile *tiles, uint32_t ntiles, const uint32_t *idx,
                      const uint32_t *eid, bool wide, uint32_t *tbase, uint32_t *packed,
                      uint32_t *peid, uint32_t *bits);
// Default short-row threshold: rows with fewer entries are summed by one lane
// (R-MAT 22: 64 beats 32 by 6 % on the FP32 sweep; 96..256 are within 1 %).
constexpr uint32_t kShortRowDefault = 64;
// The tuning in effect for `t` (zero fields replaced by defaults, others
// clamped to their ranges).
gt_tuning resolve_tuning(const gt_tuning &t);
// min(cap, units), at least 1.
int fit_grid(int cap, uint32_t units);
// Most workgroups of a sweep launch; the partials buffer holds two f64
// partials per sweep workgroup and two per fold workgroup.
constexpr int kMaxSweepGrid = 2048;
constexpr size_t kPartials = 4 * (size_t)kMaxSweepGrid;
// Workgroups of a fold kernel over `nlong` split rows (one thread each).
int fold_grid(uint32_t nlong);
// Workgroups that give each wavefront at least one of `ntiles` wave tiles,
// at most `cap`.
int wave_grid(int cap, uint32_t ntiles);
// The relabelled pull orientation (0: in-halves, 1: out-halves), built once.
gt_status ensure_relabeled(gt_graph *g, int which, Relabeled **out);
// The edge list (g->esrc, g->etgt over [0, edge_bound)), built once.
gt_status ensure_edge_list(gt_graph *g);

}  // namespace gt
