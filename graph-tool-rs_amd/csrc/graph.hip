// Device graph handle: upload of a host adjacency, download, the SpMV row
// schedule, streams and the error channel of the C ABI.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <mutex>
#include <unordered_map>

#include "device_common.hpp"
#include "gt_internal.hpp"

namespace gt {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }

gt_status fail(gt_status code, const std::string &msg) {
    set_error(msg);
    return code;
}

gt_status hip_fail(hipError_t e, const char *what, const char *file, int line) {
    char buf[512];
    std::snprintf(buf, sizeof buf, "HIP error %d (%s) in %s at %s:%d", (int)e,
                  hipGetErrorString(e), what, file, line);
    set_error(buf);
    return (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? GT_ERR_OOM
                                                                       : GT_ERR_HIP;
}

// Device bytes held by the library's buffers, and the cap on them (0: none).
static std::atomic<uint64_t> g_devmem_in_use{0};
static std::atomic<uint64_t> g_devmem_limit{0};

void devmem_released(size_t b) { g_devmem_in_use.fetch_sub(b); }

gt_status DevBuf::reserve(size_t b) {
    if (b <= bytes && ptr) return GT_OK;
    release();
    if (b == 0) b = 256;
    const uint64_t lim = g_devmem_limit.load(), used = g_devmem_in_use.load();
    if (lim && used + b > lim) {
        char msg[240];
        std::snprintf(msg, sizeof msg,
                      "device allocation of %zu bytes exceeds the library's memory limit of "
                      "%llu bytes (%llu in use)",
                      b, (unsigned long long)lim, (unsigned long long)used);
        return fail(GT_ERR_OOM, msg);
    }
    hipError_t e = hipMalloc(&ptr, b);
    if (e != hipSuccess) {
        ptr = nullptr;
        return hip_fail(e, "hipMalloc", __FILE__, __LINE__);
    }
    bytes = b;
    g_devmem_in_use.fetch_add(b);
    return GT_OK;
}

int occupancy_cap(const void *kernel, int per_cu) {
    // Compute-unit count and per-kernel occupancy are looked up once.
    static std::mutex mu;
    static std::unordered_map<const void *, int> occ;
    static int cus = -1;
    int per = 0;
    {
        std::lock_guard<std::mutex> lk(mu);
        if (cus < 0) {
            int dev = 0;
            if (hipGetDevice(&dev) != hipSuccess ||
                hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
                    hipSuccess)
                cus = 0;
        }
        auto it = occ.find(kernel);
        if (it == occ.end()) {
            int p = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&p, kernel, kBlock, 0) != hipSuccess)
                p = 0;
            it = occ.emplace(kernel, p).first;
        }
        per = it->second;
    }
    if (cus <= 0 || per <= 0) return 2048;
    if (per > per_cu) per = per_cu;
    const int g = per * cus;
    return g < 2048 ? g : 2048;
}

gt_tuning resolve_tuning(const gt_tuning &t) {
    gt_tuning r{};
    r.sweep_blocks_per_cu = t.sweep_blocks_per_cu > 0 ? std::min(t.sweep_blocks_per_cu, 8)
                                                      : kSweepBlocksPerCU;
    r.tile_rows = t.tile_rows > 0 ? std::max(16, std::min<int>(t.tile_rows, kWaveTileRows))
                                  : (int)kWaveTileRows;
    r.short_row = t.short_row > 0 ? std::min(t.short_row, 1024) : (int)kShortRowDefault;
    r.mm_tile_nnz = t.mm_tile_nnz > 0 ? std::max(256, std::min(t.mm_tile_nnz, 8192))
                                      : (int)kMmTileNnz;
    r.mm_blocks_per_cu = t.mm_blocks_per_cu > 0 ? std::min(t.mm_blocks_per_cu, 8) : 8;
    r.sweep_hot_kb = t.sweep_hot_kb < 0 ? -1 : t.sweep_hot_kb > 0 ? std::min(t.sweep_hot_kb, 128) : 128;
    r.sliced_min_len =
        t.sliced_min_len > 0 ? std::max(2, std::min(t.sliced_min_len, 1024)) : -1;
    r.sliced_share = t.sliced_share > 0 ? std::min(t.sliced_share, 99) : 0;
    r.block_tile_nnz = t.block_tile_nnz < 0    ? -1
                       : t.block_tile_nnz == 0 ? (int)kBlkTileNnzDefault
                                               : std::max(1024, std::min(t.block_tile_nnz, 1 << 20));
    r.block_long_min = t.block_long_min > 0 ? std::max(16, std::min(t.block_long_min, 65536))
                                            : (int)kBlkLongMinDefault;
    r.block_per_cu = t.block_per_cu > 0 ? std::min(t.block_per_cu, 3) : 2;
    r.block_f64 = t.block_f64 < 0 ? -1 : 1;
    return r;
}

int fit_grid(int cap, uint32_t units) {
    const int g = (uint32_t)cap < units ? cap : (int)units;
    return g > 0 ? g : 1;
}

int fold_grid(uint32_t nlong) {
    return fit_grid(kMaxSweepGrid, (nlong + kBlock - 1) / kBlock);
}

int wave_grid(int cap, uint32_t ntiles) {
    return fit_grid(cap, (ntiles + kWavesPerBlock - 1) / kWavesPerBlock);
}

gt_status graph_new(uint64_t graph_id, uint32_t n, uint64_t m, uint64_t edge_bound,
                    gt_graph **out) {
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0)
        return fail(GT_ERR_HIP, "no HIP device is available");
    gt_graph *g = new gt_graph();
    g->id = graph_id;
    g->n = n;
    g->m = m;
    g->edge_bound = edge_bound;
    e = hipStreamCreateWithFlags(&g->own_stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete g;
        return hip_fail(e, "hipStreamCreate", __FILE__, __LINE__);
    }
    g->stream = g->own_stream;
    g->tune = resolve_tuning(gt_tuning{});
    e = hipHostMalloc((void **)&g->h_state, sizeof(SweepState), hipHostMallocDefault);
    if (e != hipSuccess) {
        (void)hipStreamDestroy(g->own_stream);
        delete g;
        return hip_fail(e, "hipHostMalloc", __FILE__, __LINE__);
    }
    *out = g;
    return GT_OK;
}

// Smallest source of each wave tile (its packing base) and, through one
// atomic per tile, the widest span (largest - smallest source) of any tile.
// One workgroup per tile.
__global__ __launch_bounds__(kBlock) void tile_span(const Tile *__restrict__ tiles,
                                                    const uint32_t *__restrict__ idx,
                                                    uint32_t *__restrict__ tbase,
                                                    uint32_t *__restrict__ span_max) {
    __shared__ uint32_t lo[kWavesPerBlock], hi[kWavesPerBlock];
    const Tile t = tiles[blockIdx.x];
    uint32_t a = 0xffffffffu, b = 0u;
    for (uint32_t j = t.nnz_begin + threadIdx.x; j < t.nnz_end; j += kBlock) {
        const uint32_t v = idx[j];
        a = v < a ? v : a;
        b = v > b ? v : b;
    }
    for (int o = kWave / 2; o > 0; o >>= 1) {
        const uint32_t a2 = __shfl_xor(a, o, kWave), b2 = __shfl_xor(b, o, kWave);
        a = a2 < a ? a2 : a;
        b = b2 > b ? b2 : b;
    }
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    if (lane == 0) {
        lo[wid] = a;
        hi[wid] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kWavesPerBlock; ++w) {
            a = lo[w] < a ? lo[w] : a;
            b = hi[w] > b ? hi[w] : b;
        }
        if (t.nnz_end == t.nnz_begin) a = b = 0u;
        tbase[blockIdx.x] = a;
        atomicMax(span_max, b - a);
    }
}

// The packed entry stream of one wave tile (one workgroup per tile).  A tile
// of whole rows is sorted by packed word -- by source, ties by slot -- with a
// bitonic network over kWaveTileNnz LDS keys (padding keys are all ones and
// sort last); split chunks keep row order.  Packed sources are relative to
// tbase[tile]; with bits == 0 the stream is the plain sources in row order
// and every base is set to 0.
__global__ __launch_bounds__(kBlock) void pack_tiles(const Tile *__restrict__ tiles,
                                                     const uint32_t *__restrict__ idx,
                                                     const uint32_t *__restrict__ eid,
                                                     uint32_t bits, uint32_t *__restrict__ tbase,
                                                     uint32_t *__restrict__ packed,
                                                     uint32_t *__restrict__ peid) {
    __shared__ uint32_t key[kWaveTileNnz];
    const Tile t = tiles[blockIdx.x];
    const uint32_t nb = t.nnz_begin, cnt = t.nnz_end - t.nnz_begin;
    const int tid = threadIdx.x;
    const uint32_t base = bits ? tbase[blockIdx.x] : 0u;
    if (bits == 0 || t.row_end == kSplit) {
        for (uint32_t j = tid; j < cnt; j += kBlock) {
            packed[nb + j] = bits ? ((idx[nb + j] - base) << bits) | j : idx[nb + j];
            peid[nb + j] = eid[nb + j];
        }
        if (bits == 0 && tid == 0) tbase[blockIdx.x] = 0u;
        return;
    }
    for (uint32_t j = tid; j < kWaveTileNnz; j += kBlock)
        key[j] = j < cnt ? ((idx[nb + j] - base) << bits) | j : 0xffffffffu;
    __syncthreads();
    for (uint32_t k = 2; k <= kWaveTileNnz; k <<= 1) {
        for (uint32_t h = k >> 1; h > 0; h >>= 1) {
            for (uint32_t i = tid; i < kWaveTileNnz; i += kBlock) {
                const uint32_t o = i ^ h;
                if (o > i) {
                    const uint32_t x = key[i], y = key[o];
                    if ((x > y) == ((i & k) == 0)) {
                        key[i] = y;
                        key[o] = x;
                    }
                }
            }
            __syncthreads();
        }
    }
    const uint32_t mask = (1u << bits) - 1u;
    for (uint32_t j = tid; j < cnt; j += kBlock) {
        const uint32_t p = key[j];
        packed[nb + j] = p;
        peid[nb + j] = eid[nb + (p & mask)];
    }
}

// Wave tiles of at most kWaveTileNnz entries and kWaveTileRows rows; a row
// longer than kWaveTileNnz becomes ceil(len / kWaveTileNnz) split chunks with
// one partial slot each.  With `bounds` (ascending row indices) no tile spans
// a bound, and the host keeps the tile list.
gt_status ensure_schedule(gt_graph *g, Orient &o, uint32_t n,
                          const std::vector<uint32_t> *bounds) {
    const uint32_t max_rows = (uint32_t)g->tune.tile_rows;
    if (o.scheduled && o.sched_rows == max_rows && (!bounds || o.bounds == *bounds)) return GT_OK;
    o.batch_bounds.clear();
    const std::vector<uint32_t> &off = o.h_off;
    std::vector<Tile> tiles;
    std::vector<uint32_t> long_rows, long_slot;
    uint32_t nslots = 0;
    tiles.reserve(n / 64 + 16);
    uint32_t r = 0;
    size_t bi = 0;
    while (r < n) {
        uint32_t stop = n;
        if (bounds) {
            while (bi < bounds->size() && (*bounds)[bi] <= r) ++bi;
            if (bi < bounds->size()) stop = std::min(n, (*bounds)[bi]);
        }
        const uint32_t len = off[r + 1] - off[r];
        if (len > kWaveTileNnz) {
            long_rows.push_back(r);
            long_slot.push_back(nslots);
            for (uint32_t b = off[r]; b < off[r + 1]; b += kWaveTileNnz) {
                const uint32_t e = std::min<uint64_t>((uint64_t)b + kWaveTileNnz, off[r + 1]);
                tiles.push_back(Tile{nslots++, kSplit, b, e});
            }
            ++r;
            continue;
        }
        const uint32_t r0 = r;
        uint32_t nnz = 0;
        while (r < stop && r - r0 < max_rows) {
            const uint32_t l = off[r + 1] - off[r];
            if (l > kWaveTileNnz || nnz + l > kWaveTileNnz) break;
            nnz += l;
            ++r;
        }
        tiles.push_back(Tile{r0, r, off[r0], off[r]});
    }
    long_slot.push_back(nslots);
    if (bounds) {
        o.bounds = *bounds;
        o.h_tiles = tiles;
        o.h_long_rows = long_rows;
    } else {
        o.bounds.clear();
        o.h_tiles.clear();
        o.h_long_rows.clear();
    }
    o.ntiles = (uint32_t)tiles.size();
    o.nlong = (uint32_t)long_rows.size();
    o.nslots = nslots;
    o.sched_rows = max_rows;
    const uint64_t m = off[n];
    GT_TRY(o.tiles.reserve(sizeof(Tile) * std::max<size_t>(1, tiles.size())));
    GT_TRY(o.tbase.reserve(sizeof(uint32_t) * std::max<size_t>(2, tiles.size() + 1)));
    GT_TRY(o.long_rows.reserve(sizeof(uint32_t) * std::max<size_t>(1, long_rows.size())));
    GT_TRY(o.long_slot.reserve(sizeof(uint32_t) * long_slot.size()));
    GT_TRY(o.packed.reserve(sizeof(uint32_t) * std::max<uint64_t>(1, m)));
    GT_TRY(o.peid.reserve(sizeof(uint32_t) * std::max<uint64_t>(1, m)));
    if (!tiles.empty())
        GT_HIP(hipMemcpy(o.tiles.ptr, tiles.data(), sizeof(Tile) * tiles.size(),
                         hipMemcpyHostToDevice));
    if (!long_rows.empty())
        GT_HIP(hipMemcpy(o.long_rows.ptr, long_rows.data(),
                         sizeof(uint32_t) * long_rows.size(), hipMemcpyHostToDevice));
    GT_HIP(hipMemcpy(o.long_slot.ptr, long_slot.data(), sizeof(uint32_t) * long_slot.size(),
                     hipMemcpyHostToDevice));
    GT_TRY(pack_stream(g, o.tiles.as<Tile>(), o.ntiles, o.idx.as<uint32_t>(),
                       o.eid.as<uint32_t>(), (uint64_t)n > (1ull << (32 - kWaveSlotBits)),
                       o.tbase.as<uint32_t>(), o.packed.as<uint32_t>(), o.peid.as<uint32_t>(),
                       &o.bits));
    o.scheduled = true;
    return GT_OK;
}

void note_layout(gt_graph *g, const Orient &o) {
    g->layout = gt_sweep_layout{};
    g->layout.slot_bits = o.bits;
    g->layout.tiles = o.ntiles;
    g->layout.split_rows = o.nlong;
}

gt_status pack_stream(gt_graph *g, const Tile *tiles, uint32_t ntiles, const uint32_t *idx,
                      const uint32_t *eid, bool wide, uint32_t *tbase, uint32_t *packed,
                      uint32_t *peid, uint32_t *bits) {
    // Each tile's smallest source is its packing base.  A source and a
    // kWaveSlotBits-bit slot share one 32-bit word when every tile's sources
    // span fewer than 2^(32 - kWaveSlotBits) labels: always below that many
    // vertices; above, when no tile reaches that wide (a batch of graphs of
    // that size, a graph with local edges).  Tiles go out in launches of at
    // most 65535 workgroups; the widest span lands in tbase[ntiles].
    uint32_t *span = tbase + ntiles;
    GT_HIP(hipMemsetAsync(span, 0, sizeof(uint32_t), g->stream));
    for (uint32_t t0 = 0; t0 < ntiles; t0 += 65535) {
        const uint32_t nt = std::min<uint32_t>(65535, ntiles - t0);
        hipLaunchKernelGGL(tile_span, dim3(nt), dim3(kBlock), 0, g->stream, tiles + t0, idx,
                           tbase + t0, span);
        GT_HIP(hipGetLastError());
    }
    uint32_t widest = 0;
    if (wide && ntiles) {
        GT_HIP(hipMemcpyAsync(&widest, span, sizeof(uint32_t), hipMemcpyDeviceToHost, g->stream));
        GT_HIP(hipStreamSynchronize(g->stream));
    }
    *bits = (uint64_t)widest < (1ull << (32 - kWaveSlotBits)) ? kWaveSlotBits : 0;
    for (uint32_t t0 = 0; t0 < ntiles; t0 += 65535) {
        const uint32_t nt = std::min<uint32_t>(65535, ntiles - t0);
        hipLaunchKernelGGL(pack_tiles, dim3(nt), dim3(kBlock), 0, g->stream, tiles + t0, idx, eid,
                           *bits, tbase + t0, packed, peid);
        GT_HIP(hipGetLastError());
    }
    GT_HIP(hipStreamSynchronize(g->stream));
    return GT_OK;
}

gt_status ensure_schedule_mm(Orient &o, uint32_t n, uint32_t tile_nnz, uint32_t tile_rows) {
    Schedule &S = o.mm;
    if (S.ready && S.tile_nnz == tile_nnz && S.tile_rows == tile_rows) return GT_OK;
    const uint32_t tnnz = tile_nnz;
    const uint32_t trows = std::max<uint32_t>(1, tile_rows);
    const std::vector<uint32_t> &off = o.h_off;
    std::vector<Tile> tiles;
    std::vector<uint32_t> long_rows, long_slot;
    uint32_t nslots = 0, r = 0;
    while (r < n) {
        const uint32_t len = off[r + 1] - off[r];
        if (len > tnnz) {
            long_rows.push_back(r);
            long_slot.push_back(nslots);
            for (uint32_t b = off[r]; b < off[r + 1]; b += tnnz) {
                const uint32_t e = std::min<uint64_t>((uint64_t)b + tnnz, off[r + 1]);
                tiles.push_back(Tile{nslots++, kSplit, b, e});
            }
            ++r;
            continue;
        }
        const uint32_t r0 = r;
        uint32_t nnz = 0;
        while (r < n && r - r0 < trows) {
            const uint32_t l = off[r + 1] - off[r];
            if (l > tnnz || nnz + l > tnnz) break;
            nnz += l;
            ++r;
        }
        tiles.push_back(Tile{r0, r, off[r0], off[r]});
    }
    long_slot.push_back(nslots);
    S.ntiles = (uint32_t)tiles.size();
    S.nlong = (uint32_t)long_rows.size();
    S.nslots = nslots;
    GT_TRY(S.tiles.reserve(sizeof(Tile) * std::max<size_t>(1, tiles.size())));
    GT_TRY(S.long_rows.reserve(sizeof(uint32_t) * std::max<size_t>(1, long_rows.size())));
    GT_TRY(S.long_slot.reserve(sizeof(uint32_t) * long_slot.size()));
    if (!tiles.empty())
        GT_HIP(hipMemcpy(S.tiles.ptr, tiles.data(), sizeof(Tile) * tiles.size(),
                         hipMemcpyHostToDevice));
    if (!long_rows.empty())
        GT_HIP(hipMemcpy(S.long_rows.ptr, long_rows.data(), sizeof(uint32_t) * long_rows.size(),
                         hipMemcpyHostToDevice));
    GT_HIP(hipMemcpy(S.long_slot.ptr, long_slot.data(), sizeof(uint32_t) * long_slot.size(),
                     hipMemcpyHostToDevice));
    S.tile_nnz = tile_nnz;
    S.tile_rows = tile_rows;
    S.ready = true;
    return GT_OK;
}

// esrc[e] = kNoEdge for every id, then the out-halves' entries (row s,
// entry (t, e)) set esrc[e] = s, etgt[e] = t: one wavefront per row.
__global__ __launch_bounds__(kBlock) void fill_u32(uint32_t *p, uint64_t len, uint32_t v) {
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < len;
         i += (uint64_t)gridDim.x * kBlock)
        p[i] = v;
}

__global__ __launch_bounds__(kBlock) void scatter_edges(const uint32_t *__restrict__ off,
                                                        const uint32_t *__restrict__ idx,
                                                        const uint32_t *__restrict__ eid,
                                                        uint32_t n, uint32_t *__restrict__ es,
                                                        uint32_t *__restrict__ et) {
    const int lane = threadIdx.x & 63;
    const uint32_t waves = gridDim.x * kWavesPerBlock;
    for (uint32_t r = blockIdx.x * kWavesPerBlock + threadIdx.x / 64; r < n; r += waves)
        for (uint32_t j = off[r] + lane; j < off[r + 1]; j += 64) {
            const uint32_t e = eid[j];
            es[e] = r;
            et[e] = idx[j];
        }
}

gt_status ensure_edge_list(gt_graph *g) {
    if (g->edges_ready) return GT_OK;
    const uint64_t eb = g->edge_bound;
    GT_TRY(g->esrc.reserve(sizeof(uint32_t) * std::max<uint64_t>(1, eb)));
    GT_TRY(g->etgt.reserve(sizeof(uint32_t) * std::max<uint64_t>(1, eb)));
    if (eb) {
        const unsigned fb = (unsigned)std::min<uint64_t>(8192, (eb + kBlock - 1) / kBlock);
        hipLaunchKernelGGL(fill_u32, dim3(fb), dim3(kBlock), 0, g->stream,
                           g->esrc.as<uint32_t>(), eb, kNoEdge);
        hipLaunchKernelGGL(fill_u32, dim3(fb), dim3(kBlock), 0, g->stream,
                           g->etgt.as<uint32_t>(), eb, kNoEdge);
        if (g->n && g->m) {
            const unsigned rb = (unsigned)std::min<uint64_t>(16384, ((uint64_t)g->n + 3) / 4);
            hipLaunchKernelGGL(scatter_edges, dim3(rb), dim3(kBlock), 0, g->stream,
                               g->out.off.as<uint32_t>(), g->out.idx.as<uint32_t>(),
                               g->out.eid.as<uint32_t>(), g->n, g->esrc.as<uint32_t>(),
                               g->etgt.as<uint32_t>());
        }
        GT_HIP(hipGetLastError());
    }
    g->edges_ready = true;
    return GT_OK;
}

static gt_status upload_orient(Orient &o, uint32_t n, uint64_t m, const uint32_t *off,
                               const uint32_t *idx, const uint32_t *eid) {
    GT_TRY(o.off.reserve(sizeof(uint32_t) * ((size_t)n + 1)));
    GT_TRY(o.idx.reserve(sizeof(uint32_t) * std::max<uint64_t>(1, m)));
    GT_TRY(o.eid.reserve(sizeof(uint32_t) * std::max<uint64_t>(1, m)));
    GT_HIP(hipMemcpy(o.off.ptr, off, sizeof(uint32_t) * ((size_t)n + 1), hipMemcpyHostToDevice));
    if (m) {
        GT_HIP(hipMemcpy(o.idx.ptr, idx, sizeof(uint32_t) * m, hipMemcpyHostToDevice));
        GT_HIP(hipMemcpy(o.eid.ptr, eid, sizeof(uint32_t) * m, hipMemcpyHostToDevice));
    }
    o.h_off.assign(off, off + (size_t)n + 1);
    return GT_OK;
}

// Host-side validation of an uploaded orientation: monotone offsets that end
// at m, endpoints < n, edge ids < edge_bound.
static gt_status check_orient(const char *name, uint32_t n, uint64_t m, uint64_t edge_bound,
                              const uint32_t *off, const uint32_t *idx, const uint32_t *eid) {
    if (off[0] != 0 || off[n] != m)
        return fail(GT_ERR_INVALID_ARG, std::string(name) + ": offsets must start at 0 and end at m");
    for (uint32_t v = 0; v < n; ++v)
        if (off[v + 1] < off[v])
            return fail(GT_ERR_INVALID_ARG, std::string(name) + ": offsets must be non-decreasing");
    for (uint64_t j = 0; j < m; ++j) {
        if (idx[j] >= n)
            return fail(GT_ERR_NO_SUCH_VERTEX, std::string(name) + ": endpoint out of range");
        if (eid[j] >= edge_bound)
            return fail(GT_ERR_INVALID_ARG, std::string(name) + ": edge id >= edge_bound");
    }
    return GT_OK;
}

// Profiling: HIP events around every sweep launch, on the stream it runs on.
gt_status prof_begin(gt_graph *g, uint64_t k) {
    if (!g->profiling) return GT_OK;
    GT_HIP(hipEventRecord(g->ev[2 * (k % 256)], g->stream));
    return GT_OK;
}
gt_status prof_end(gt_graph *g, uint64_t k) {
    if (!g->profiling) return GT_OK;
    GT_HIP(hipEventRecord(g->ev[2 * (k % 256) + 1], g->stream));
    if ((k % 256) == 255) {
        GT_HIP(hipStreamSynchronize(g->stream));
        for (int q = 0; q < 256; ++q) {
            float ms = 0.f;
            GT_HIP(hipEventElapsedTime(&ms, g->ev[2 * q], g->ev[2 * q + 1]));
            g->prof_ms[g->prof_kind] += ms;
        }
        g->prof_launches[g->prof_kind] += 256;
    }
    return GT_OK;
}
gt_status prof_flush(gt_graph *g, uint64_t launched) {
    if (!g->profiling) return GT_OK;
    const uint64_t rem = launched % 256;
    GT_HIP(hipStreamSynchronize(g->stream));
    for (uint64_t q = 0; q < rem; ++q) {
        float ms = 0.f;
        GT_HIP(hipEventElapsedTime(&ms, g->ev[2 * q], g->ev[2 * q + 1]));
        g->prof_ms[g->prof_kind] += ms;
    }
    g->prof_launches[g->prof_kind] += rem;
    return GT_OK;
}
gt_status prof_prepare(gt_graph *g) {
    if (g->profiling && g->ev.empty()) {
        g->ev.resize(512);
        for (auto &e : g->ev) GT_HIP(hipEventCreate(&e));
    }
    return GT_OK;
}

}  // namespace gt

using namespace gt;

extern "C" {

const char *gt_last_error(void) { return g_last_error.c_str(); }

gt_status gt_set_device_memory_limit(uint64_t bytes) {
    clear_error();
    g_devmem_limit.store(bytes);
    return GT_OK;
}

gt_status gt_device_memory_in_use(uint64_t *bytes) {
    clear_error();
    if (!bytes) return fail(GT_ERR_INVALID_ARG, "bytes is null");
    *bytes = g_devmem_in_use.load();
    return GT_OK;
}

gt_status gt_device_count(int32_t *count) {
    clear_error();
    if (!count) return fail(GT_ERR_INVALID_ARG, "count is null");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    *count = (e == hipSuccess) ? n : 0;
    return GT_OK;
}

gt_status gt_graph_upload(uint64_t graph_id, uint32_t n, uint64_t edge_bound,
                          const uint32_t *in_off, const uint32_t *in_other,
                          const uint32_t *in_eid, const uint32_t *out_off,
                          const uint32_t *out_other, const uint32_t *out_eid,
                          gt_graph **out) {
    clear_error();
    if (!out || !in_off || !out_off) return fail(GT_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    if (n == 0xffffffffu) return fail(GT_ERR_INVALID_ARG, "vertex count exceeds the index width");
    uint64_t m = in_off[n];
    if (out_off[n] != m)
        return fail(GT_ERR_INVALID_ARG, "in-halves and out-halves hold different edge counts");
    if (m > 0xffffffffull) return fail(GT_ERR_INVALID_ARG, "edge count exceeds the index width");
    if (m && (!in_other || !in_eid || !out_other || !out_eid))
        return fail(GT_ERR_INVALID_ARG, "null adjacency array");
    GT_TRY(check_orient("in-halves", n, m, edge_bound, in_off, in_other, in_eid));
    GT_TRY(check_orient("out-halves", n, m, edge_bound, out_off, out_other, out_eid));
    gt_graph *g = nullptr;
    GT_TRY(graph_new(graph_id, n, m, edge_bound, &g));
    gt_status st = upload_orient(g->in, n, m, in_off, in_other, in_eid);
    if (st == GT_OK) st = upload_orient(g->out, n, m, out_off, out_other, out_eid);
    if (st != GT_OK) {
        gt_graph_free(g);
        return st;
    }
    *out = g;
    return GT_OK;
}

gt_status gt_graph_free(gt_graph *g) {
    if (!g) return GT_OK;
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    for (hipEvent_t e : g->ev) (void)hipEventDestroy(e);
    if (g->h_state) (void)hipHostFree(g->h_state);
    for (hipStream_t st : g->side_stream)
        if (st) {
            (void)hipStreamSynchronize(st);
            (void)hipStreamDestroy(st);
        }
    if (g->own_stream) (void)hipStreamDestroy(g->own_stream);
    delete g;
    return GT_OK;
}

gt_status gt_graph_info(const gt_graph *g, uint32_t *n, uint64_t *m, uint64_t *edge_bound,
                        uint64_t *graph_id) {
    clear_error();
    if (!g) return fail(GT_ERR_INVALID_ARG, "null graph");
    if (n) *n = g->n;
    if (m) *m = g->m;
    if (edge_bound) *edge_bound = g->edge_bound;
    if (graph_id) *graph_id = g->id;
    return GT_OK;
}

gt_status gt_graph_download(const gt_graph *g, uint32_t *in_off, uint32_t *in_other,
                            uint32_t *in_eid, uint32_t *out_off, uint32_t *out_other,
                            uint32_t *out_eid) {
    clear_error();
    if (!g) return fail(GT_ERR_INVALID_ARG, "null graph");
    GT_HIP(hipStreamSynchronize(g->stream));
    size_t nb = sizeof(uint32_t) * ((size_t)g->n + 1), mb = sizeof(uint32_t) * g->m;
    if (in_off) GT_HIP(hipMemcpy(in_off, g->in.off.ptr, nb, hipMemcpyDeviceToHost));
    if (out_off) GT_HIP(hipMemcpy(out_off, g->out.off.ptr, nb, hipMemcpyDeviceToHost));
    if (g->m) {
        if (in_other) GT_HIP(hipMemcpy(in_other, g->in.idx.ptr, mb, hipMemcpyDeviceToHost));
        if (in_eid) GT_HIP(hipMemcpy(in_eid, g->in.eid.ptr, mb, hipMemcpyDeviceToHost));
        if (out_other) GT_HIP(hipMemcpy(out_other, g->out.idx.ptr, mb, hipMemcpyDeviceToHost));
        if (out_eid) GT_HIP(hipMemcpy(out_eid, g->out.eid.ptr, mb, hipMemcpyDeviceToHost));
    }
    return GT_OK;
}

gt_status gt_graph_set_stream(gt_graph *g, void *hip_stream) {
    clear_error();
    if (!g) return fail(GT_ERR_INVALID_ARG, "null graph");
    g->stream = hip_stream ? (hipStream_t)hip_stream : g->own_stream;
    return GT_OK;
}

gt_status gt_graph_set_tuning(gt_graph *g, const gt_tuning *t) {
    clear_error();
    if (!g || !t) return fail(GT_ERR_INVALID_ARG, "null argument");
    g->tune = resolve_tuning(*t);
    return GT_OK;
}

gt_status gt_last_layout(const gt_graph *g, gt_sweep_layout *out) {
    clear_error();
    if (!g || !out) return fail(GT_ERR_INVALID_ARG, "null argument");
    *out = g->layout;
    return GT_OK;
}

gt_status gt_graph_get_tuning(const gt_graph *g, gt_tuning *t) {
    clear_error();
    if (!g || !t) return fail(GT_ERR_INVALID_ARG, "null argument");
    *t = g->tune;
    return GT_OK;
}

}  // extern "C"
