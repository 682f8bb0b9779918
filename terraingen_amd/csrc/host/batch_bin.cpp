// SPDX-License-Identifier: BSD-3-Clause
//
// Binning a crater catalogue into the bins of many tiles at once, for the batched kernel. For each
// tile this is what ApplyCraters does for one grid (src/layers/craters/shaping/apply.cpp:225-240,
// 119-140, prepared.cpp:60-87): craters too small for the tile's pixel size are dropped
// (ResolutionWeight 0), each remaining crater's footprint box becomes the tile's pixel window, and
// the crater joins every 16 x 16 bin that window meets, in catalogue (apply) order.
//
// Candidates come from a spatial index: craters grouped by diameter level (powers of two, as the
// generator's bands) and, within a level, by square cells a few diameters across, so a tile only
// visits the cells its box (grown by the level's largest footprint) meets, and skips the levels
// whose craters its pixel size cannot resolve.

#include "batch.h"
#include "common.h"

#include <atomic>
#include <mutex>
#include <thread>
#include <unordered_map>

namespace tgh {

namespace {

std::uint64_t CellKey(std::int64_t i, std::int64_t j) {
    return (std::uint64_t(std::uint32_t(std::int32_t(i))) << 32) |
           std::uint32_t(std::int32_t(j));
}

// ResolutionWeight(D, spacing) != 0 (src/layers/fade.cpp:48-52 with the default fade 2-4).
bool Resolved(double diameter, double pixel) {
    const double x = diameter / pixel;
    if (x <= 2.0) {
        return false;
    }
    if (x >= 4.0) {
        return true;
    }
    double t = std::clamp((x - 2.0) / (4.0 - 2.0), 0.0, 1.0);
    return 1.0 - (1.0 - t * t * (3.0 - 2.0 * t)) != 0.0;
}


}  // namespace

IndexPtr MakeIndex(const double* cx, const double* cy, const double* reach,
                                 const double* diameter, std::size_t n) {
    IndexPtr index(new tgx_index());
    index->cx.assign(cx, cx + n);
    index->cy.assign(cy, cy + n);
    index->reach.assign(reach, reach + n);
    index->diameter.assign(diameter, diameter + n);
    int base = 0;
    if (n > 0) {
        double smallest = INFINITY;
        for (std::size_t k = 0; k < n; ++k) {
            smallest = std::fmin(smallest, diameter[k]);
        }
        base = std::ilogb(smallest);
    }
    std::vector<IndexLevel>& levels = index->levels;
    for (std::size_t k = 0; k < n; ++k) {
        if (!(diameter[k] > 0.0) || !std::isfinite(cx[k]) || !std::isfinite(cy[k]) ||
            !(reach[k] >= 0.0)) {
            throw std::invalid_argument("catalogue crater " + std::to_string(k) + " is malformed");
        }
        const std::size_t l = std::size_t(std::max(0, std::ilogb(diameter[k]) - base));
        if (l >= levels.size()) {
            levels.resize(l + 1);
        }
        IndexLevel& level = levels[l];
        if (level.cell == 0.0) {
            level.cell = 8.0 * std::ldexp(1.0, base + int(l) + 1);
        }
        level.maxReach = std::fmax(level.maxReach, reach[k]);
        level.maxDiameter = std::fmax(level.maxDiameter, diameter[k]);
        const auto i = std::int64_t(std::floor(cx[k] / level.cell));
        const auto j = std::int64_t(std::floor(cy[k] / level.cell));
        level.cells[CellKey(i, j)].push_back(std::uint32_t(k));
    }
    return index;
}

void ListWindows(const std::vector<tgx_window>& windows, const tgx_tile& tile, std::uint32_t side,
                 std::size_t row0, std::size_t row1, TileBins& out) {
    const std::size_t binsX = (tile.width + side - 1) / side;
    std::vector<std::vector<std::uint32_t>> lists(binsX * (row1 - row0));
    for (const tgx_window& w : windows) {
        const std::size_t by0 = std::max<std::size_t>(w.row0 / side, row0);
        const std::size_t by1 = std::min<std::size_t>(w.row1 / side + 1, row1);
        for (std::size_t by = by0; by < by1; ++by) {
            for (std::size_t bx = w.col0 / side; bx <= w.col1 / side; ++bx) {
                lists[(by - row0) * binsX + bx].push_back(w.crater);
            }
        }
    }
    out.counts.resize(lists.size());
    out.entries.clear();
    for (std::size_t b = 0; b < lists.size(); ++b) {
        out.counts[b] = std::uint32_t(lists[b].size());
        out.entries.insert(out.entries.end(), lists[b].begin(), lists[b].end());
    }
}

namespace {

// The windows of one tile's candidates against the index: for each crater resolved at the tile's
// pixel size whose footprint box meets the tile, in catalogue order, the cells its pixel window
// covers (PixelWindowIn over the whole tile). Candidates are gathered for the tile's rows
// [rowFirst, rowEnd) only, so a strip of a large tile finds the craters that can reach it.
void PlaneWindows(const tgx_index& index, const tgx_tile& tile, std::size_t rowFirst,
                  std::size_t rowEnd, std::vector<std::uint32_t>& candidates,
                  std::vector<tgx_window>& windows) {
    const Grid g{tile.width, tile.height, tile.origin_x, tile.origin_y, tile.pixel};
    const double minX = tile.origin_x, maxX = tile.origin_x + tile.width * tile.pixel;
    const double maxY = tile.origin_y - double(rowFirst) * tile.pixel;
    const double minY = tile.origin_y - double(std::min<std::size_t>(rowEnd, tile.height)) * tile.pixel;
    // The candidates are marked in a bitmap over the catalogue, which then lists them in
    // catalogue order in one pass over the marked span, rather than sorting them.
    thread_local std::vector<std::uint64_t> marks;
    if (marks.size() < (index.cx.size() + 63) / 64) {
        marks.assign((index.cx.size() + 63) / 64, 0ull);
    }
    std::size_t lowWord = marks.size(), highWord = 0;
    const auto mark = [&](const std::vector<std::uint32_t>& list) {
        for (const std::uint32_t k : list) {
            marks[k >> 6u] |= 1ull << (k & 63u);
            lowWord = std::min<std::size_t>(lowWord, k >> 6u);
            highWord = std::max<std::size_t>(highWord, k >> 6u);
        }
    };
    for (const IndexLevel& level : index.levels) {
        if (level.cell == 0.0 || !Resolved(level.maxDiameter, tile.pixel)) {
            continue;
        }
        const double grow = level.maxReach;
        const auto i0 = std::int64_t(std::floor((minX - grow) / level.cell));
        const auto i1 = std::int64_t(std::floor((maxX + grow) / level.cell));
        const auto j0 = std::int64_t(std::floor((minY - grow) / level.cell));
        const auto j1 = std::int64_t(std::floor((maxY + grow) / level.cell));
        if ((i1 - i0 + 1) * (j1 - j0 + 1) > std::int64_t(level.cells.size())) {
            for (const auto& [key, list] : level.cells) {
                mark(list);
            }
            continue;
        }
        for (std::int64_t j = j0; j <= j1; ++j) {
            for (std::int64_t i = i0; i <= i1; ++i) {
                const auto it = level.cells.find(CellKey(i, j));
                if (it != level.cells.end()) {
                    mark(it->second);
                }
            }
        }
    }
    candidates.clear();
    for (std::size_t w = lowWord; w <= highWord && w < marks.size(); ++w) {
        std::uint64_t bits = marks[w];
        marks[w] = 0ull;  // left clear for the next tile
        while (bits != 0ull) {
            candidates.push_back(std::uint32_t(w * 64u + std::uint32_t(__builtin_ctzll(bits))));
            bits &= bits - 1ull;
        }
    }
    windows.clear();
    for (const std::uint32_t k : candidates) {
        if (!Resolved(index.diameter[k], tile.pixel)) {
            continue;
        }
        const double r = index.reach[k];
        const PixelWindow w = PixelWindowIn(g, index.cx[k] - r, index.cy[k] - r,
                                            index.cx[k] + r, index.cy[k] + r);
        if (w.Empty() || w.rowEnd <= rowFirst || w.rowBegin >= rowEnd) {
            continue;
        }
        windows.push_back({k, std::uint32_t(w.colBegin), std::uint32_t(w.colEnd - 1),
                           std::uint32_t(w.rowBegin), std::uint32_t(w.rowEnd - 1)});
    }
}

// One tile's bins against the index: per-bin counts and the entries in bin order.
// The bins of bin rows [row0, row1) of `tile`: each crater's pixel window is the whole tile's, so
// a strip's bins are those of the whole tile; only its candidates are gathered from the strip.
void BinTile(const tgx_index& index, const tgx_tile& tile, std::size_t row0, std::size_t row1,
             std::vector<std::uint32_t>& candidates, TileBins& out) {
    constexpr std::uint32_t kSide = 16;
    std::vector<tgx_window> windows;
    PlaneWindows(index, tile, row0 * kSide, row1 * kSide, candidates, windows);
    ListWindows(windows, tile, kSide, row0, row1, out);
}

}  // namespace

void CheckTile(const tgx_tile& tile, std::size_t t) {
    if (tile.width == 0 || tile.height == 0 || !(tile.pixel > 0.0) ||
        !std::isfinite(tile.origin_x) || !std::isfinite(tile.origin_y) ||
        !std::isfinite(tile.pixel)) {
        throw std::invalid_argument("tile " + std::to_string(t) + " is malformed");
    }
}

void BinBatch(const tgx_index& index, const tgx_tile* tiles, std::size_t tileCount, unsigned threads,
              std::vector<tg_tile_desc>& descs, std::vector<tg_bin>& bins,
              std::vector<std::uint32_t>& entries) {
    if (index.sphere) {
        BinSphereBatch(*index.sphere, tiles, tileCount, threads, descs, bins, entries);
        return;
    }
    constexpr std::uint32_t kSide = 16;
    for (std::size_t t = 0; t < tileCount; ++t) {
        CheckTile(tiles[t], t);
    }
    // Each tile's bins in strips of bin rows, in parallel, so one large tile spreads over threads.
    const std::vector<Strip> strips = MakeStrips(tiles, tileCount);
    std::vector<TileBins> per(strips.size());
    std::atomic<std::size_t> next{0};
    auto work = [&] {
        std::vector<std::uint32_t> candidates;
        for (;;) {
            const std::size_t n = next.fetch_add(1);
            if (n >= strips.size()) {
                return;
            }
            const Strip& s = strips[n];
            BinTile(index, tiles[s.tile], s.row0, s.row1, candidates, per[n]);
        }
    };
    if (threads == 0) {
        threads = std::max(1u, std::thread::hardware_concurrency());
    }
    threads = unsigned(std::min<std::size_t>(threads, std::max<std::size_t>(strips.size(), 1)));
    std::vector<std::thread> pool;
    for (unsigned i = 1; i < threads; ++i) {
        pool.emplace_back(work);
    }
    work();
    for (std::thread& th : pool) {
        th.join();
    }
    // Concatenated in tile order.
    TileDescs(tiles, tileCount, kSide, descs);
    bins.clear();
    entries.clear();
    AppendStrips(strips, per, bins, entries);
}

void TileDescs(const tgx_tile* tiles, std::size_t tileCount, std::uint32_t side,
               std::vector<tg_tile_desc>& descs) {
    descs.clear();
    std::uint64_t out = 0;
    std::uint64_t binCursor = 0;
    for (std::size_t t = 0; t < tileCount; ++t) {
        const tgx_tile& tile = tiles[t];
        tg_tile_desc d{};
        d.origin_x = tile.origin_x;
        d.origin_y = tile.origin_y;
        d.pixel = tile.pixel;
        d.width = tile.width;
        d.height = tile.height;
        d.out_offset = out;
        d.bin_first = std::uint32_t(binCursor);
        d.bins_across = (tile.width + side - 1) / side;
        descs.push_back(d);
        out += std::uint64_t(tile.width) * tile.height;
        binCursor += std::uint64_t(d.bins_across) * ((tile.height + side - 1) / side);
        if (binCursor > 0xFFFFFFFFull) {
            throw std::invalid_argument("batch too large for 32-bit bin and entry indices");
        }
    }
}

void WindowBatch(const tgx_index& index, const tgx_tile* tiles, std::size_t tileCount,
                 std::uint32_t side, unsigned threads, std::vector<tg_tile_desc>& descs,
                 std::vector<std::uint32_t>& tileWindows, std::vector<tgx_window>& windows) {
    if (side != 8 && side != 16 && side != 32) {
        throw std::invalid_argument("the bin side must be 8, 16 or 32 cells");
    }
    for (std::size_t t = 0; t < tileCount; ++t) {
        CheckTile(tiles[t], t);
        if (index.sphere && !(tiles[t].origin_y <= 90.0 &&
                              tiles[t].origin_y - tiles[t].height * tiles[t].pixel >= -90.0)) {
            throw std::invalid_argument("tile " + std::to_string(t) +
                                        " reaches past a pole (its latitudes leave [-90, 90])");
        }
    }
    TileDescs(tiles, tileCount, side, descs);
    std::vector<std::vector<tgx_window>> per(tileCount);
    ForEachIndex(tileCount, threads, [&](std::size_t t) {
        std::vector<std::uint32_t> candidates;
        if (index.sphere) {
            SphereWindows(*index.sphere, tiles[t], 0, tiles[t].height, candidates, per[t]);
        } else {
            PlaneWindows(index, tiles[t], 0, tiles[t].height, candidates, per[t]);
        }
    });
    tileWindows.assign(2 * tileCount, 0u);
    windows.clear();
    for (std::size_t t = 0; t < tileCount; ++t) {
        tileWindows[2 * t] = std::uint32_t(windows.size());
        tileWindows[2 * t + 1] = std::uint32_t(per[t].size());
        windows.insert(windows.end(), per[t].begin(), per[t].end());
        if (windows.size() > 0xFFFFFFFFull) {
            throw std::invalid_argument("batch too large for 32-bit window indices");
        }
    }
}

std::vector<Strip> MakeStrips(const tgx_tile* tiles, std::size_t tileCount) {
    constexpr std::size_t kSide = 16, kRows = 8;
    std::vector<Strip> strips;
    for (std::size_t t = 0; t < tileCount; ++t) {
        const std::size_t rows = (tiles[t].height + kSide - 1) / kSide;
        for (std::size_t r = 0; r < rows; r += kRows) {
            strips.push_back({t, r, std::min(rows, r + kRows)});
        }
    }
    return strips;
}

void AppendStrips(const std::vector<Strip>& strips, const std::vector<TileBins>& per,
                  std::vector<tg_bin>& bins, std::vector<std::uint32_t>& entries) {
    for (std::size_t n = 0; n < strips.size(); ++n) {
        std::uint64_t first = entries.size();
        for (const std::uint32_t count : per[n].counts) {
            bins.push_back({std::uint32_t(strips[n].tile), std::uint32_t(first), count, 0u});
            first += count;
        }
        entries.insert(entries.end(), per[n].entries.begin(), per[n].entries.end());
        if (entries.size() > 0xFFFFFFFFull || bins.size() > 0xFFFFFFFFull) {
            throw std::invalid_argument("batch too large for 32-bit bin and entry indices");
        }
    }
}

// The (cell, crater) pairs of a binned batch whose FP64 cell centre lies strictly inside the
// crater's footprint radius (reach radii x outline bound x R), the pairs a crater kernel must
// evaluate in full, with the summed per-crater cost of those pairs. Independent of any kernel: the
// cell centres are the FP64 PixelToWorld centres.
void BatchPairOps(const tg_tile_desc* descs, std::size_t tileCount, const tg_bin* bins,
                  std::size_t binCount, const std::uint32_t* entries, std::size_t entryCount,
                  const double* cx, const double* cy, const double* reach, const double* cost,
                  std::size_t n, unsigned threads, std::uint64_t& listed, std::uint64_t& inside,
                  double& ops) {
    constexpr std::uint32_t kSide = 16;
    for (std::size_t b = 0; b < binCount; ++b) {
        const tg_bin& bin = bins[b];
        if (bin.tile >= tileCount || std::uint64_t(bin.first) + bin.count > entryCount) {
            throw std::invalid_argument("bin " + std::to_string(b) + " is out of range");
        }
    }
    for (std::size_t e = 0; e < entryCount; ++e) {
        if (entries[e] >= n) {
            throw std::invalid_argument("entry " + std::to_string(e) + " is out of range");
        }
    }
    if (threads == 0) {
        threads = std::max(1u, std::thread::hardware_concurrency());
    }
    std::atomic<std::size_t> next{0};
    std::mutex sumMutex;
    listed = inside = 0;
    ops = 0.0;
    auto work = [&] {
        std::uint64_t myListed = 0, myInside = 0;
        double myOps = 0.0;
        constexpr std::size_t kChunk = 256;
        for (;;) {
            const std::size_t begin = next.fetch_add(kChunk);
            if (begin >= binCount) {
                break;
            }
            const std::size_t end = std::min(binCount, begin + kChunk);
            for (std::size_t b = begin; b < end; ++b) {
                const tg_bin& bin = bins[b];
                const tg_tile_desc& t = descs[bin.tile];
                const std::uint32_t local = std::uint32_t(b) - t.bin_first;
                const std::uint32_t c0 = (local % t.bins_across) * kSide;
                const std::uint32_t r0 = (local / t.bins_across) * kSide;
                const std::uint32_t c1 = std::min(t.width, c0 + kSide);
                const std::uint32_t r1 = std::min(t.height, r0 + kSide);
                for (std::uint32_t e = bin.first; e < bin.first + bin.count; ++e) {
                    const std::uint32_t k = entries[e];
                    const double r2 = reach[k] * reach[k];
                    std::uint64_t count = 0;
                    for (std::uint32_t row = r0; row < r1; ++row) {
                        const double dy = t.origin_y - (double(row) + 0.5) * t.pixel - cy[k];
                        for (std::uint32_t col = c0; col < c1; ++col) {
                            const double dx = t.origin_x + (double(col) + 0.5) * t.pixel - cx[k];
                            count += (dx * dx + dy * dy < r2) ? 1u : 0u;
                        }
                    }
                    myListed += std::uint64_t(r1 - r0) * (c1 - c0);
                    myInside += count;
                    myOps += double(count) * cost[k];
                }
            }
        }
        std::lock_guard<std::mutex> lock(sumMutex);
        listed += myListed;
        inside += myInside;
        ops += myOps;
    };
    std::vector<std::thread> pool;
    for (unsigned i = 1; i < threads; ++i) {
        pool.emplace_back(work);
    }
    work();
    for (std::thread& th : pool) {
        th.join();
    }
}

}  // namespace tgh

extern "C" {

int tgx_batch_pair_ops(const tg_tile_desc* descs, size_t tileCount, const tg_bin* bins,
                       size_t binCount, const uint32_t* entries, size_t entryCount,
                       const double* cx, const double* cy, const double* reach,
                       const double* cost, size_t n, unsigned threads, uint64_t* listed,
                       uint64_t* inside, double* ops, char* err, size_t errLen) {
    return tgh::Guard(err, errLen, [&] {
        if (listed == nullptr || inside == nullptr || ops == nullptr ||
            (tileCount > 0 && descs == nullptr) || (binCount > 0 && bins == nullptr) ||
            (entryCount > 0 && entries == nullptr) ||
            (n > 0 && (cx == nullptr || cy == nullptr || reach == nullptr || cost == nullptr))) {
            throw std::invalid_argument("tgx_batch_pair_ops: null argument");
        }
        tgh::BatchPairOps(descs, tileCount, bins, binCount, entries, entryCount, cx, cy, reach,
                          cost, n, threads, *listed, *inside, *ops);
    });
}

int tgx_batch_bin(const double* cx, const double* cy, const double* reach,
                  const double* diameter, size_t n, const tgx_tile* tiles, size_t tileCount,
                  unsigned threads, tg_tile_desc** descs, tg_bin** bins, size_t* binCount,
                  uint32_t** entries, size_t* entryCount, char* err, size_t errLen) {
    return tgh::Guard(err, errLen, [&] {
        if (descs == nullptr || bins == nullptr || entries == nullptr || binCount == nullptr ||
            entryCount == nullptr || (tileCount > 0 && tiles == nullptr) ||
            (n > 0 && (cx == nullptr || cy == nullptr || reach == nullptr || diameter == nullptr))) {
            throw std::invalid_argument("tgx_batch_bin: null argument");
        }
        const tgh::IndexPtr index = tgh::MakeIndex(cx, cy, reach, diameter, n);
        std::vector<tg_tile_desc> d;
        std::vector<tg_bin> b;
        std::vector<std::uint32_t> e;
        tgh::BinBatch(*index, tiles, tileCount, threads, d, b, e);
        *descs = tgh::Export(d);
        *bins = tgh::Export(b);
        *binCount = b.size();
        *entries = tgh::Export(e);
        *entryCount = e.size();
    });
}

int tgx_index_create(const double* cx, const double* cy, const double* reach,
                     const double* diameter, size_t n, tgx_index** out, char* err,
                     size_t errLen) {
    return tgh::Guard(err, errLen, [&] {
        if (out == nullptr ||
            (n > 0 && (cx == nullptr || cy == nullptr || reach == nullptr || diameter == nullptr))) {
            throw std::invalid_argument("tgx_index_create: null argument");
        }
        *out = nullptr;
        *out = tgh::MakeIndex(cx, cy, reach, diameter, n).release();
    });
}

void tgx_index_free(tgx_index* index) {
    delete index;
}

size_t tgx_index_count(const tgx_index* index) {
    return index != nullptr ? index->cx.size() : 0;
}

int tgx_index_windows(const tgx_index* index, const tgx_tile* tiles, size_t tileCount,
                      uint32_t binSide, unsigned threads, tg_tile_desc** descs,
                      uint32_t** tileWindows, tgx_window** windows, size_t* windowCount,
                      char* err, size_t errLen) {
    return tgh::Guard(err, errLen, [&] {
        if (index == nullptr || descs == nullptr || tileWindows == nullptr ||
            windows == nullptr || windowCount == nullptr || (tileCount > 0 && tiles == nullptr)) {
            throw std::invalid_argument("tgx_index_windows: null argument");
        }
        std::vector<tg_tile_desc> d;
        std::vector<std::uint32_t> tw;
        std::vector<tgx_window> w;
        tgh::WindowBatch(*index, tiles, tileCount, binSide, threads, d, tw, w);
        *descs = tgh::Export(d);
        *tileWindows = tgh::Export(tw);
        *windows = tgh::Export(w);
        *windowCount = w.size();
    });
}

int tgx_index_bin(const tgx_index* index, const tgx_tile* tiles, size_t tileCount,
                  unsigned threads, tg_tile_desc** descs, tg_bin** bins, size_t* binCount,
                  uint32_t** entries, size_t* entryCount, char* err, size_t errLen) {
    return tgh::Guard(err, errLen, [&] {
        if (index == nullptr || descs == nullptr || bins == nullptr || entries == nullptr ||
            binCount == nullptr || entryCount == nullptr || (tileCount > 0 && tiles == nullptr)) {
            throw std::invalid_argument("tgx_index_bin: null argument");
        }
        std::vector<tg_tile_desc> d;
        std::vector<tg_bin> b;
        std::vector<std::uint32_t> e;
        tgh::BinBatch(*index, tiles, tileCount, threads, d, b, e);
        *descs = tgh::Export(d);
        *bins = tgh::Export(b);
        *binCount = b.size();
        *entries = tgh::Export(e);
        *entryCount = e.size();
    });
}

}  // extern "C"
