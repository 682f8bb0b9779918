// SPDX-License-Identifier: BSD-3-Clause
//
// Host binning of tile batches against a catalogue index, shared by the C API and the tile server.
#pragma once

#include "terraingen_amd.h"

#include <cstdint>
#include <memory>
#include <unordered_map>
#include <vector>

namespace tgh {

struct IndexLevel {
    double cell = 0.0;       // index cell side (m)
    double maxReach = 0.0;   // largest footprint half-width in the level (m)
    double maxDiameter = 0.0;
    std::unordered_map<std::uint64_t, std::vector<std::uint32_t>> cells;
};

// A catalogue on a sphere indexed by 1-degree longitude-latitude cells (sphere.cpp).
struct SphereCatalogue;

}  // namespace tgh

// The catalogue's spatial index: on a plane, craters grouped by diameter level and, within a
// level, by cell; on a sphere, the sphere catalogue (its centres in degrees in cx, cy).
struct tgx_index {
    std::vector<double> cx, cy, reach, diameter;
    std::vector<tgh::IndexLevel> levels;
    std::shared_ptr<const tgh::SphereCatalogue> sphere;
};

namespace tgh {

struct IndexDeleter {
    void operator()(tgx_index* index) const {
        tgx_index_free(index);
    }
};
using IndexPtr = std::unique_ptr<tgx_index, IndexDeleter>;

IndexPtr MakeIndex(const double* cx, const double* cy, const double* reach,
                                     const double* diameter, std::size_t n);

// Throws std::invalid_argument for a tile with no cells, a non-positive or non-finite pixel size
// or a non-finite origin.
void CheckTile(const tgx_tile& tile, std::size_t t);

// A tile's bins of bin rows [row0, row1), one unit of parallel binning work, and what binning
// it gives: per bin its count, and the entries in bin order.
struct Strip {
    std::size_t tile, row0, row1;
};
struct TileBins {
    std::vector<std::uint32_t> counts;
    std::vector<std::uint32_t> entries;
};
// Tiles cut into strips of at most 8 bin rows, in tile then row order.
std::vector<Strip> MakeStrips(const tgx_tile* tiles, std::size_t tileCount);
// Appends the strips' bins (tile `strips[n].tile`, counts and entries `per[n]`) in order.
void AppendStrips(const std::vector<Strip>& strips, const std::vector<TileBins>& per,
                  std::vector<tg_bin>& bins, std::vector<std::uint32_t>& entries);

// The sphere catalogue of craters centred at (lon, lat) degrees with footprint half-widths
// `reach` in meters of arc, on a sphere of `radius` meters.
std::shared_ptr<const SphereCatalogue> MakeSphereCatalogue(const double* lon, const double* lat,
                                                           const double* reach,
                                                           const double* diameter,
                                                           std::size_t n, double radius);
// The sphere's form of BinBatch (tgx_sphere_bin).
void BinSphereBatch(const SphereCatalogue& catalogue, const tgx_tile* tiles,
                    std::size_t tileCount, unsigned threads, std::vector<tg_tile_desc>& descs,
                    std::vector<tg_bin>& bins, std::vector<std::uint32_t>& entries);
// The radius of the sphere an index lies on; 0 on a plane.
double IndexRadius(const tgx_index& index);

// The tile descriptors of `tiles` for bins of `side` cells, heights packed from offset 0.
void TileDescs(const tgx_tile* tiles, std::size_t tileCount, std::uint32_t side,
               std::vector<tg_tile_desc>& descs);
// Appends to `out` the bins of bin rows [row0, row1) of `tile` for bins of `side` cells: each
// window's crater in every bin its cells meet, in window order.
void ListWindows(const std::vector<tgx_window>& windows, const tgx_tile& tile, std::uint32_t side,
                 std::size_t row0, std::size_t row1, TileBins& out);
// The sphere's form of the windows of one tile (sphere.cpp): cells within reach of each resolved
// candidate, for the tile's rows [rowFirst, rowEnd).
void SphereWindows(const SphereCatalogue& index, const tgx_tile& tile, std::size_t rowFirst,
                   std::size_t rowEnd, std::vector<std::uint32_t>& candidates,
                   std::vector<tgx_window>& windows);
// tgx_index_windows: the descriptors for bins of `side` cells, each tile's (first, count) of
// windows, and the windows, every tile's in catalogue order.
void WindowBatch(const tgx_index& index, const tgx_tile* tiles, std::size_t tileCount,
                 std::uint32_t side, unsigned threads, std::vector<tg_tile_desc>& descs,
                 std::vector<std::uint32_t>& tileWindows, std::vector<tgx_window>& windows);

// Bins `tiles` into 16 x 16 bins against the index (resolved craters only, in catalogue order),
// packing the tiles' heights one after another from offset 0, over `threads` threads (0: all).
void BinBatch(const tgx_index& index, const tgx_tile* tiles, std::size_t tileCount,
              unsigned threads, std::vector<tg_tile_desc>& descs, std::vector<tg_bin>& bins,
              std::vector<std::uint32_t>& entries);

}  // namespace tgh
