// SPDX-License-Identifier: BSD-3-Clause
//
// Crater footprints binned into tiles of grid cells, the lists the crater kernel walks per cell
// (src/layers/craters/shaping/apply.cpp:28-32,109-140 of the reference).

#include "common.h"

#include <cstdlib>

namespace tgh {

// Tiles of `side` cells each crater's footprint box covers, crater order kept within each tile.
void BinPlane(const Grid& g, const double* cx, const double* cy, const double* reach,
              std::size_t n, std::uint32_t side, std::vector<std::uint32_t>& heads,
              std::vector<std::uint32_t>& entries) {
    if (side == 0) {
        throw std::invalid_argument("tile side must be positive");
    }
    const std::size_t tilesX = (g.width + side - 1) / side;
    const std::size_t tilesY = (g.height + side - 1) / side;
    // Count, then fill: one pass over the craters' tile ranges each time.
    std::vector<std::uint32_t> counts(tilesX * tilesY, 0);
    std::vector<PixelWindow> windows(n);
    for (std::size_t k = 0; k < n; ++k) {
        const PixelWindow w = PixelWindowIn(g, cx[k] - reach[k], cy[k] - reach[k],
                                            cx[k] + reach[k], cy[k] + reach[k]);
        windows[k] = w;
        if (w.Empty()) {
            continue;
        }
        for (std::size_t ty = w.rowBegin / side; ty <= (w.rowEnd - 1) / side; ++ty) {
            for (std::size_t tx = w.colBegin / side; tx <= (w.colEnd - 1) / side; ++tx) {
                ++counts[ty * tilesX + tx];
            }
        }
    }
    heads.assign(3 + 2 * tilesX * tilesY, 0);
    heads[0] = static_cast<std::uint32_t>(tilesX);
    heads[1] = side;
    heads[2] = 0;
    std::size_t total = 0;
    for (std::size_t t = 0; t < counts.size(); ++t) {
        heads[3 + 2 * t] = static_cast<std::uint32_t>(total);
        total += counts[t];
    }
    if (total > 0xFFFFFFFFull) {
        throw std::invalid_argument("too many tile entries for 32-bit offsets");
    }
    entries.assign(std::max<std::size_t>(total, 1), 0);
    for (std::size_t k = 0; k < n; ++k) {
        const PixelWindow& w = windows[k];
        if (w.Empty()) {
            continue;
        }
        for (std::size_t ty = w.rowBegin / side; ty <= (w.rowEnd - 1) / side; ++ty) {
            for (std::size_t tx = w.colBegin / side; tx <= (w.colEnd - 1) / side; ++tx) {
                const std::size_t t = ty * tilesX + tx;
                entries[heads[3 + 2 * t] + heads[4 + 2 * t]++] = static_cast<std::uint32_t>(k);
            }
        }
    }
}

}  // namespace tgh

extern "C" {

void tgx_free(void* p) {
    std::free(p);
}

double tgx_outline_max_scale(const tg_crater_record* r, const float* data) {
    double sum = 1.0;
    for (std::uint32_t i = 0; i < r->table_count; ++i) {
        sum = std::fmax(sum, 1.0 + double(data[r->table_offset + i]));
    }
    for (std::uint32_t i = 0; i < r->harmonic_count; ++i) {
        sum += std::fabs(double(data[r->harmonic_offset + 2 * i]));
    }
    return std::sqrt(double(r->ellipticity)) * sum;
}

int tgx_bin_plane(const tgx_grid* grid, const double* cx, const double* cy, const double* reach,
                  size_t n, uint32_t side, uint32_t** tiles, size_t* tileWords,
                  uint32_t** entries, size_t* entryCount, char* err, size_t errLen) {
    return tgh::Guard(err, errLen, [&] {
        if (grid == nullptr || tiles == nullptr || entries == nullptr ||
            (n > 0 && (cx == nullptr || cy == nullptr || reach == nullptr))) {
            throw std::invalid_argument("tgx_bin_plane: null argument");
        }
        std::vector<std::uint32_t> heads, list;
        tgh::BinPlane(tgh::Grid::From(*grid), cx, cy, reach, n, side, heads, list);
        *tiles = tgh::Export(heads);
        *tileWords = heads.size();
        *entries = tgh::Export(list);
        *entryCount = list.size();
    });
}

}  // extern "C"
