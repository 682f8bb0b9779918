// SPDX-License-Identifier: BSD-3-Clause
//
// Host binning of tile batches against a catalogue index, shared by the C API and the tile server.
#pragma once

#include "terraingen_amd.h"

#include <cstdint>
#include <memory>
#include <vector>

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

// Bins `tiles` into 16 x 16 bins against the index (resolved craters only, in catalogue order),
// packing the tiles' heights one after another from offset 0, over `threads` threads (0: all).
void BinBatch(const tgx_index& index, const tgx_tile* tiles, std::size_t tileCount,
              unsigned threads, std::vector<tg_tile_desc>& descs, std::vector<tg_bin>& bins,
              std::vector<std::uint32_t>& entries);

}  // namespace tgh
