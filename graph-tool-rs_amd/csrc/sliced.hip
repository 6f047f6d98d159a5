// The sliced split of a relabelled pull orientation (struct Sliced).
//
// A pull sweep spends most of its time in the gathers of the source vector:
// a random 4- or 8-byte load per entry, one cache-line request per distinct
// line in each 64-lane load.  On R-MAT graphs the long rows hold most of the
// entries (scale 22: rows of 64 entries or more, 2.8 % of the rows, hold 74 %
// of them).  The sliced sweep serves those rows from LDS: the source labels
// are cut into slices that fit beside a workgroup's wave tiles, every long
// row into one piece per slice it has entries in, and a workgroup that owns a
// run of a slice's pieces copies that slice into LDS once and gathers from it.
// Each piece's sum goes to its own slot; a row's slots are contiguous in slice
// order, and the fold kernel adds them in that order, as it adds a split row's
// chunks.  The other rows are swept as before by the launch's other
// workgroups.  Sums change order (piece by piece), not value beyond rounding.
//
// Built once per graph, slice width, row threshold, tile rows and grid:
// integer work on the host from the row offsets and the entries (copied back
// once), then the pieces' entries are sorted and packed on the device.
#include <algorithm>
#include <cmath>
#include <vector>

#include "gt_internal.hpp"

namespace gt {

namespace {

// A piece: up to two runs of a row's entries (orientation positions) and
// its first slot (a piece longer than a wave tile has one per chunk).
struct Piece {
    uint32_t b0, l0, b1, l1;
    uint32_t slot;
};

template <typename T>
gt_status upload(DevBuf &d, const std::vector<T> &h, size_t min_count = 1) {
    GT_TRY(d.reserve(sizeof(T) * std::max(h.size(), min_count)));
    if (!h.empty()) GT_HIP(hipMemcpy(d.ptr, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
    return GT_OK;
}

}  // namespace

gt_status ensure_sliced(gt_graph *g, Relabeled &R, uint32_t width, uint32_t vbytes, int grid,
                        Sliced **out) {
    Sliced &S = R.sl[vbytes == 8 ? 1 : 0];
    *out = &S;
    const uint32_t D = (uint32_t)std::max(2, g->tune.sliced_min_len);
    const uint32_t max_rows = (uint32_t)g->tune.tile_rows;
    const uint32_t share = (uint32_t)g->tune.sliced_share;
    if (S.ready && S.width == width && S.min_len == D && S.sched_rows == max_rows &&
        S.share == share && S.grid == (uint32_t)grid)
        return GT_OK;
    S.ready = false;
    S.active = false;
    S.width = width;
    S.min_len = D;
    S.sched_rows = max_rows;
    S.share = share;
    S.grid = (uint32_t)grid;
    Orient &o = R.pull;
    const uint32_t n = g->n;
    const std::vector<uint32_t> &off = o.h_off;
    const uint64_t m = off[n];
    std::vector<uint32_t> heavy;
    uint64_t eh = 0;
    for (uint32_t r = 0; r < n; ++r) {
        const uint32_t l = off[r + 1] - off[r];
        if (l >= D) {
            heavy.push_back(r);
            eh += l;
        }
    }
    if (heavy.empty() || grid < 2 || width < 4) {
        S.ready = true;
        return GT_OK;
    }

    // Light rows: runs of consecutive light rows, at most kWaveTileNnz
    // entries and max_rows rows per tile (a light row is shorter than D <=
    // kWaveTileNnz, so it always fits an empty tile).
    std::vector<Tile> lt;
    uint64_t el = 0, nlight = 0;
    for (uint32_t r = 0; r < n;) {
        if (off[r + 1] - off[r] >= D) {
            ++r;
            continue;
        }
        const uint32_t r0 = r;
        uint32_t nnz = 0;
        while (r < n && r - r0 < max_rows) {
            const uint32_t l = off[r + 1] - off[r];
            if (l >= D || nnz + l > kWaveTileNnz) break;
            nnz += l;
            ++r;
        }
        lt.push_back(Tile{r0, r, off[r0], off[r]});
        el += nnz;
        nlight += r - r0;
    }

    // The entries, to cut the heavy rows by slice.
    std::vector<uint32_t> idx(m), eid(m);
    GT_HIP(hipMemcpy(idx.data(), o.idx.ptr, sizeof(uint32_t) * m, hipMemcpyDeviceToHost));
    GT_HIP(hipMemcpy(eid.data(), o.eid.ptr, sizeof(uint32_t) * m, hipMemcpyDeviceToHost));

    // Warm slices: 1 .. warm_end - 1, as long as the heavy rows' runs in a
    // slice average at least kMinPiece entries (slice 0 is the hot slice the
    // light walk holds anyway; further out, runs of one or two entries cost
    // more as pieces than their global gathers).
    const uint32_t nsl = (uint32_t)(((uint64_t)n + width - 1) / width);
    std::vector<uint64_t> sl_entries(nsl, 0), sl_runs(nsl, 0);
    for (uint32_t r : heavy)
        for (uint32_t j = off[r]; j < off[r + 1];) {
            const uint32_t sl = idx[j] / width;
            uint32_t je = j + 1;
            while (je < off[r + 1] && idx[je] / width == sl) ++je;
            sl_entries[sl] += je - j;
            sl_runs[sl] += 1;
            j = je;
        }
    constexpr uint64_t kMinPiece = 8;
    uint32_t wend = 1;
    while (wend < nsl && sl_runs[wend] && sl_entries[wend] >= kMinPiece * sl_runs[wend]) ++wend;
    if (wend == 1) {
        S.ready = true;
        return GT_OK;
    }

    // Pieces of heavy row k, in its slot order: the main piece (the entries
    // below the first warm label and from the last warm label on: a prefix
    // and a suffix of the row), then one piece per warm slice it has entries
    // in.  A piece longer than a wave tile takes one slot per chunk.
    std::vector<std::vector<Piece>> by(wend);  // by[0]: main pieces
    std::vector<uint32_t> long_slot(heavy.size() + 1);
    uint64_t slot = 0;
    const uint32_t wlo = width, whi = (uint32_t)std::min<uint64_t>((uint64_t)wend * width, n);
    auto chunks = [](uint32_t len) { return (len + kWaveTileNnz - 1) / kWaveTileNnz; };
    for (size_t k = 0; k < heavy.size(); ++k) {
        const uint32_t r = heavy[k], b = off[r], e = off[r + 1];
        long_slot[k] = (uint32_t)slot;
        uint32_t p = b;
        while (p < e && idx[p] < wlo) ++p;
        uint32_t s = p;
        while (s < e && idx[s] < whi) ++s;
        if (p - b + e - s > 0) {
            by[0].push_back(Piece{b, p - b, s, e - s, (uint32_t)slot});
            slot += chunks(p - b + e - s);
        }
        for (uint32_t j = p; j < s;) {
            const uint32_t sl = idx[j] / width;
            uint32_t je = j + 1;
            while (je < s && idx[je] / width == sl) ++je;
            by[sl].push_back(Piece{j, je - j, 0, 0, (uint32_t)slot});
            slot += chunks(je - j);
            j = je;
        }
    }
    if (slot > 0x7fffffffull) return fail(GT_ERR_UNSUPPORTED, "sliced sweep: too many pieces");
    long_slot[heavy.size()] = (uint32_t)slot;

    // Tiles of pieces: runs of pieces (at most kWaveTileNnz entries,
    // max_rows pieces), a piece longer than a tile as chunk tiles.  Each run
    // is followed by one sentinel offset (its end), so a tile's last piece
    // ends where the tile does.  A piece's sum goes to partial `part0 + its
    // offset index`; chunks' partials follow all pieces of the stream.
    // long_idx[slot] is the partial a slot reads.
    std::vector<uint32_t> long_idx(slot);
    struct Stream {
        std::vector<Tile> tiles;
        std::vector<uint32_t> offs, src, eid, first;
        std::vector<std::pair<uint32_t, uint32_t>> chunk;  // (tile, slot)
        uint32_t pos = 0;
    };
    std::vector<uint32_t> buf;
    auto build = [&](Stream &T, const std::vector<Piece> &P, uint32_t base, bool by_slice) {
        for (size_t i = 0; i < P.size();) {
            const uint32_t len = P[i].l0 + P[i].l1;
            buf.clear();
            if (len > kWaveTileNnz) {
                for (uint32_t q = 0; q < P[i].l0; ++q) buf.push_back(P[i].b0 + q);
                for (uint32_t q = 0; q < P[i].l1; ++q) buf.push_back(P[i].b1 + q);
                for (uint32_t c = 0; c * kWaveTileNnz < len; ++c) {
                    const uint32_t cl = std::min(kWaveTileNnz, len - c * kWaveTileNnz);
                    T.chunk.emplace_back((uint32_t)T.tiles.size(), P[i].slot + c);
                    T.tiles.push_back(Tile{0, kSplit, T.pos, T.pos + cl});
                    for (uint32_t q = 0; q < cl; ++q) {
                        T.src.push_back(idx[buf[c * kWaveTileNnz + q]] - base);
                        T.eid.push_back(eid[buf[c * kWaveTileNnz + q]]);
                    }
                    T.pos += cl;
                }
                ++i;
                continue;
            }
            const uint32_t pr0 = (uint32_t)T.offs.size(), nb = T.pos;
            uint32_t nnz = 0;
            while (i < P.size() && T.offs.size() - pr0 < max_rows &&
                   P[i].l0 + P[i].l1 <= kWaveTileNnz && nnz + P[i].l0 + P[i].l1 <= kWaveTileNnz) {
                long_idx[P[i].slot] = (uint32_t)T.offs.size();  // offset index, rebased below
                T.offs.push_back(T.pos);
                for (uint32_t q = 0; q < P[i].l0; ++q) {
                    T.src.push_back(idx[P[i].b0 + q] - base);
                    T.eid.push_back(eid[P[i].b0 + q]);
                }
                for (uint32_t q = 0; q < P[i].l1; ++q) {
                    T.src.push_back(idx[P[i].b1 + q] - base);
                    T.eid.push_back(eid[P[i].b1 + q]);
                }
                T.pos += P[i].l0 + P[i].l1;
                nnz += P[i].l0 + P[i].l1;
                ++i;
            }
            const uint32_t pr1 = (uint32_t)T.offs.size();
            T.offs.push_back(T.pos);  // sentinel: the end of the tile's last piece
            T.tiles.push_back(Tile{pr0, pr1, nb, T.pos});
        }
        (void)by_slice;
    };
    // Warm pieces slice by slice (sources relative to the slice), then the
    // main pieces (sources as they are).
    Stream W, M;
    W.first.resize(wend + 1, 0);
    for (uint32_t sl = 1; sl < wend; ++sl) {
        W.first[sl] = (uint32_t)W.tiles.size();
        build(W, by[sl], sl * width, true);
        std::vector<Piece>().swap(by[sl]);
    }
    W.first[wend] = (uint32_t)W.tiles.size();
    // Offset indices of the warm stream are its partials; the main stream's
    // start after the warm chunks.
    std::vector<uint8_t> is_main(slot, 0);
    for (const Piece &pc : by[0])
        for (uint32_t c = 0; c < chunks(pc.l0 + pc.l1); ++c) is_main[pc.slot + c] = 1;
    build(M, by[0], 0, false);
    std::vector<Piece>().swap(by[0]);
    const uint32_t wpo = (uint32_t)W.offs.size();
    const uint32_t mpart = wpo + (uint32_t)W.chunk.size();
    const uint32_t mpo = (uint32_t)M.offs.size();
    for (uint64_t q = 0; q < slot; ++q)
        if (is_main[q]) long_idx[q] += mpart;
    for (size_t c = 0; c < W.chunk.size(); ++c) {
        W.tiles[W.chunk[c].first].row_begin = wpo + (uint32_t)c;
        long_idx[W.chunk[c].second] = wpo + (uint32_t)c;
    }
    for (size_t c = 0; c < M.chunk.size(); ++c) {
        M.tiles[M.chunk[c].first].row_begin = mpart + mpo + (uint32_t)c;
        long_idx[M.chunk[c].second] = mpart + mpo + (uint32_t)c;
    }
    const uint32_t nout = mpart + mpo + (uint32_t)M.chunk.size();
    uint64_t ew = 0;
    for (uint32_t sl = 1; sl < wend; ++sl) ew += sl_entries[sl];

    // Workgroups: ga of the grid take the warm tiles, in contiguous runs of
    // about equal cost, each run cut into one segment per slice it covers;
    // the others walk the light rows and the main pieces.  Costs in units of
    // an entry gathered from LDS: a global gather 3, a row's update 12, a
    // piece's 8, a tile 64, a slice fill a 32nd of its bytes.
    const double fill_cost = (double)width * vbytes / 32.0, tile_cost = 64.0;
    auto wcost = [&](const Tile &x) {
        return tile_cost + (x.nnz_end - x.nnz_begin) +
               (x.row_end == kSplit ? 0.0 : 8.0 * (x.row_end - x.row_begin));
    };
    double wtotal = 0.0;
    for (uint32_t sl = 1; sl < wend; ++sl) {
        if (W.first[sl] == W.first[sl + 1]) continue;
        wtotal += fill_cost;
        for (uint32_t q = W.first[sl]; q < W.first[sl + 1]; ++q) wtotal += wcost(W.tiles[q]);
    }
    uint64_t lglob = 0, mglob = 0;
    for (uint32_t r = 0; r < n; ++r) {
        const bool h = off[r + 1] - off[r] >= D;
        for (uint32_t j = off[r]; j < off[r + 1]; ++j) {
            if (idx[j] < width) continue;
            if (!h) ++lglob;
            else if (idx[j] >= whi) ++mglob;
        }
    }
    const double ltotal = (double)el + 2.0 * lglob + 12.0 * nlight + 64.0 * lt.size() +
                          (double)(eh - ew) + 2.0 * mglob + 8.0 * mpo + 64.0 * M.tiles.size();
    uint32_t ga;
    if (share > 0)
        ga = (uint32_t)std::lround(grid * share / 100.0);
    else
        ga = (uint32_t)std::lround(grid * wtotal / (wtotal + ltotal));
    ga = std::max<uint32_t>(1, std::min<uint32_t>(ga, (uint32_t)grid - 1));
    std::vector<SliceSeg> seg;
    std::vector<uint32_t> wseg(ga + 1, 0);
    uint32_t b = 0;
    double acc = 0.0;
    for (uint32_t sl = 1; sl < wend; ++sl) {
        if (W.first[sl] == W.first[sl + 1]) continue;
        const uint32_t base = sl * width;
        const uint32_t len = (std::min(width, n - base) + 3u) & ~3u;
        bool open = false;
        acc += fill_cost;
        for (uint32_t q = W.first[sl]; q < W.first[sl + 1]; ++q) {
            if (!open) {
                seg.push_back(SliceSeg{base, len, q, q});
                open = true;
            }
            acc += wcost(W.tiles[q]);
            seg.back().t1 = q + 1;
            if (b + 1 < ga && acc >= wtotal * (b + 1) / ga) {
                wseg[++b] = (uint32_t)seg.size();
                open = false;
                acc += fill_cost;
            }
        }
    }
    for (uint32_t q = b + 1; q <= ga; ++q) wseg[q] = (uint32_t)seg.size();

    // Upload, then sort and pack the three streams on the device.
    GT_TRY(upload(S.ltiles, lt));
    GT_TRY(S.ltbase.reserve(sizeof(uint32_t) * (lt.size() + 2)));
    GT_TRY(S.lpacked.reserve(sizeof(uint32_t) * std::max<uint64_t>(1, m)));
    GT_TRY(S.lpeid.reserve(sizeof(uint32_t) * std::max<uint64_t>(1, m)));
    GT_TRY(upload(S.long_rows, heavy));
    GT_TRY(upload(S.long_slot, long_slot));
    GT_TRY(upload(S.long_idx, long_idx));
    GT_TRY(upload(S.seg, seg));
    GT_TRY(upload(S.wseg, wseg));
    auto put = [&](Stream &T, DevBuf &tiles, DevBuf &tbase, DevBuf &offs, DevBuf &packed,
                   DevBuf &peid, bool wide, uint32_t *bits) -> gt_status {
        GT_TRY(upload(tiles, T.tiles));
        GT_TRY(tbase.reserve(sizeof(uint32_t) * (T.tiles.size() + 2)));
        GT_TRY(upload(offs, T.offs));
        GT_TRY(packed.reserve(sizeof(uint32_t) * std::max<size_t>(1, T.src.size())));
        GT_TRY(peid.reserve(sizeof(uint32_t) * std::max<size_t>(1, T.src.size())));
        DevBuf dsrc, deid;
        GT_TRY(upload(dsrc, T.src));
        GT_TRY(upload(deid, T.eid));
        return pack_stream(g, tiles.as<Tile>(), (uint32_t)T.tiles.size(), dsrc.as<uint32_t>(),
                           deid.as<uint32_t>(), wide, tbase.as<uint32_t>(), packed.as<uint32_t>(),
                           peid.as<uint32_t>(), bits);
    };
    uint32_t wbits = 0;
    GT_TRY(put(W, S.htiles, S.htbase, S.hoff, S.hpacked, S.hpeid, false, &wbits));
    if (wbits != kWaveSlotBits) return fail(GT_ERR_HIP, "sliced sweep: piece stream not packed");
    const bool wide = (uint64_t)n > (1ull << (32 - kWaveSlotBits));
    GT_TRY(put(M, S.mtiles, S.mtbase, S.moff, S.mpacked, S.mpeid, wide, &S.mbits));
    GT_TRY(pack_stream(g, S.ltiles.as<Tile>(), (uint32_t)lt.size(), o.idx.as<uint32_t>(),
                       o.eid.as<uint32_t>(), wide, S.ltbase.as<uint32_t>(),
                       S.lpacked.as<uint32_t>(), S.lpeid.as<uint32_t>(), &S.lbits));
    S.nltiles = (uint32_t)lt.size();
    S.nmtiles = (uint32_t)M.tiles.size();
    S.mpart = mpart;
    S.light_entries = el;
    S.nlong = (uint32_t)heavy.size();
    S.nslots = (uint32_t)slot;
    S.nout = nout;
    S.nhtiles = (uint32_t)W.tiles.size();
    S.npieces = wpo + mpo;
    S.nslices = wend - 1;
    S.warm_end = wend;
    S.heavy_entries = ew;
    S.heavy_wgs = ga;
    S.active = true;
    S.ready = true;
    return GT_OK;
}

}  // namespace gt
