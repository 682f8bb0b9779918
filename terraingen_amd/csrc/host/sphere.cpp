// SPDX-License-Identifier: BSD-3-Clause
//
// Tiles on a sphere: cell positions and split points of longitude-latitude rasters, crater frames,
// and the binning of a catalogue into the 16 x 16 bins of many such tiles.
//   positions  Surface::Position of each cell centre (src/geometry/surface.cpp:45-50,261-267),
//              with the sine and cosine of sphere_math.h, the same bits the batched kernel uses
//   split      MakeSplitPoint (src/layers/craters/shaping/crater_kernel.cpp:303-310)
//   frames     MakeCraterFrame (crater_kernel.cpp:285-301) with LocalAxes (surface.cpp:18-24), in
//              FP64 with the C library's sine and cosine, as the reference computes them
//   bins       ApplyCraters' rule (apply.cpp:119-140) on a sphere: a crater joins every bin its
//              footprint's longitude-latitude box meets, in catalogue order, when its resolution
//              weight at the tile's spacing (cell side times metres per degree) is above 0.

#include "batch.h"
#include "common.h"
#include "sphere_math.h"

#include <atomic>
#include <cmath>
#include <map>
#include <thread>
#include <unordered_map>

namespace tgh {

// The craters of one diameter level (a power of two), grouped by longitude-latitude cell of
// `cellDeg` degrees, twice the level's largest footprint: a tile visits the cells its box grown
// by that footprint meets.
struct SphereLevel {
    double cellDeg = 0.0;
    double maxReachDeg = 0.0;
    double maxDiameter = 0.0;
    std::int64_t lonCells = 0, latCells = 0;
    std::unordered_map<std::int64_t, std::vector<std::uint32_t>> cells;

    std::int64_t LonCell(double lon) const {
        double l = std::fmod(lon + 180.0, 360.0);
        if (l < 0.0) {
            l += 360.0;
        }
        return std::min(lonCells - 1, std::int64_t(std::floor(l / cellDeg)));
    }
    std::int64_t LatCell(double lat) const {
        return std::clamp<std::int64_t>(std::int64_t(std::floor((lat + 90.0) / cellDeg)), 0,
                                        latCells - 1);
    }
    std::int64_t Key(std::int64_t i, std::int64_t j) const {
        return i * latCells + j;
    }
};

// Crater centres in degrees with footprints in degrees of arc, indexed by diameter level.
struct SphereCatalogue {
    std::vector<double> lon, lat, reachDeg, diameter;
    std::vector<double> lonHalfWidth;  // LonHalfWidth(lat, reachDeg) per crater
    double radius = 0.0;
    std::vector<SphereLevel> levels;
};

namespace {

constexpr double kPi = 3.141592653589793238462643383279502884;
constexpr double kDegree = kPi / 180.0;
constexpr std::uint32_t kSide = 16;

void Split(double v, float& hi, float& lo) {
    hi = static_cast<float>(v);
    lo = static_cast<float>(v - static_cast<double>(hi));
}

// ResolutionWeight(D, spacing) != 0 with the default fade of 2 to 4 samples (fade.cpp:48-52).
bool Resolved(double diameter, double spacing) {
    const double x = diameter / spacing;
    if (x <= 2.0) {
        return false;
    }
    if (x >= 4.0) {
        return true;
    }
    const double t = std::clamp((x - 2.0) / 2.0, 0.0, 1.0);
    return 1.0 - (1.0 - t * t * (3.0 - 2.0 * t)) != 0.0;
}

// The signed longitude difference a - b folded into [-180, 180).
double LonDelta(double a, double b) {
    double d = std::fmod(a - b + 180.0, 360.0);
    if (d < 0.0) {
        d += 360.0;
    }
    return d - 180.0;
}

// The half-width in degrees of longitude of every point within `arc` degrees of arc of a point at
// latitude `lat` or nearer the equator than |lat| + arc: |sin dLon| <= sin arc / cos lat' for a
// point at latitude lat', so asin(sin arc / cos(|lat| + arc)) bounds it; 360 when the cap holds a
// pole.
double LonHalfWidth(double lat, double arc) {
    const double polar = std::fabs(lat) + arc;
    if (polar >= 89.999999) {
        return 360.0;
    }
    const double s = std::sin(arc * kDegree) / std::cos(polar * kDegree);
    if (s >= 1.0) {
        return 360.0;
    }
    return std::asin(s) / kDegree * (1.0 + 1e-12) + 1e-12;
}

}  // namespace

// The windows of one tile: for each candidate resolved at the tile's spacing, in catalogue order,
// the cells within reach of its footprint (a cell to spare each side), for the tile's rows
// [rowFirst, rowEnd). Every bound is the whole tile's; only the candidates are gathered from the
// rows' latitudes, so a strip of a large tile finds the craters that can reach it.
void SphereWindows(const SphereCatalogue& index, const tgx_tile& tile, std::size_t rowFirst,
                   std::size_t rowEnd, std::vector<std::uint32_t>& candidates,
                   std::vector<tgx_window>& windows) {
    const double west = tile.origin_x, step = tile.pixel;
    const double north = tile.origin_y;
    const double span = tile.width * step;
    // The rows' latitudes, two cells wider each way: a window reaches a cell and a half past the
    // rows in reach of its crater, so every window meeting the rows is gathered, whichever rows
    // the tile is cut into.
    const double stripNorth = north - (double(rowFirst) - 2.0) * step;
    const double stripSouth =
        north - (double(std::min<std::size_t>(rowEnd, tile.height)) + 2.0) * step;
    const double spacing = step * index.radius * kDegree;
    // Candidates: per diameter level whose largest crater the tile can resolve, the cells of the
    // tile's box grown by the level's largest footprint.
    candidates.clear();
    for (const SphereLevel& level : index.levels) {
        if (!Resolved(level.maxDiameter, spacing)) {
            continue;
        }
        const double grow = level.maxReachDeg;
        const double latLo = std::max(-90.0, stripSouth - grow);
        const double latHi = std::min(90.0, stripNorth + grow);
        const double lonGrow =
            LonHalfWidth(std::max(std::fabs(stripSouth), std::fabs(stripNorth)), grow);
        const std::int64_t j0 = level.LatCell(latLo), j1 = level.LatCell(latHi);
        const double lonSpan = span + 2.0 * lonGrow;
        const std::int64_t lonSteps =
            lonSpan >= 359.0 ? level.lonCells
                             : std::min(level.lonCells,
                                        std::int64_t(std::ceil(lonSpan / level.cellDeg)) + 2);
        if (lonSteps * (j1 - j0 + 1) > std::int64_t(level.cells.size())) {
            // Fewer occupied cells than cells to visit: filter the occupied ones by row.
            for (const auto& [key, list] : level.cells) {
                const std::int64_t j = key % level.latCells;
                if (j < j0 || j > j1) {
                    continue;
                }
                candidates.insert(candidates.end(), list.begin(), list.end());
            }
            continue;
        }
        const std::int64_t i0 = lonSpan >= 359.0 ? 0 : level.LonCell(west - lonGrow);
        for (std::int64_t j = j0; j <= j1; ++j) {
            for (std::int64_t k = 0; k < lonSteps; ++k) {
                const std::int64_t i = (i0 + k) % level.lonCells;
                const auto it = level.cells.find(level.Key(i, j));
                if (it != level.cells.end()) {
                    candidates.insert(candidates.end(), it->second.begin(), it->second.end());
                }
            }
        }
    }
    std::sort(candidates.begin(), candidates.end());
    candidates.erase(std::unique(candidates.begin(), candidates.end()), candidates.end());
    windows.clear();
    for (const std::uint32_t k : candidates) {
        if (!Resolved(index.diameter[k], spacing)) {
            continue;
        }
        const double r = index.reachDeg[k];
        const double lat = index.lat[k];
        // Rows whose cell centres lie within r of the crater's latitude (row r's centre is at
        // north - (r + 0.5) step): a meridian is a great circle, so no farther cell is in reach.
        const double rowA = (north - (lat + r)) / step - 0.5;
        const double rowB = (north - (lat - r)) / step - 0.5;
        if (rowB < -1.0 || rowA > double(tile.height)) {
            continue;
        }
        const double dLon = index.lonHalfWidth[k];
        double colA = -1.0, colB = double(tile.width);
        if (span + 2.0 * dLon < 359.0) {
            // The crater's longitude east of the tile's middle, folded into [-180, 180).
            const double offset = LonDelta(index.lon[k], west + 0.5 * span) + 0.5 * span;
            colA = (offset - dLon) / step - 0.5;
            colB = (offset + dLon) / step - 0.5;
            if (colB < -1.0 || colA > double(tile.width)) {
                continue;
            }
        }
        // Cells from floor(a) - 1 to ceil(b) + 1 cover the band with a cell to spare each side.
        const auto first = [](double a, std::uint32_t n) {
            return std::uint32_t(std::clamp(std::floor(a) - 1.0, 0.0, double(n - 1)));
        };
        const auto last = [](double b, std::uint32_t n) {
            return std::uint32_t(std::clamp(std::ceil(b) + 1.0, 0.0, double(n - 1)));
        };
        const tgx_window w{k, first(colA, tile.width), last(colB, tile.width),
                           first(rowA, tile.height), last(rowB, tile.height)};
        if (w.row1 < rowFirst || w.row0 >= rowEnd) {
            continue;
        }
        windows.push_back(w);
    }
}

namespace {

// One tile's bins: per-bin counts and the entries in bin order, for bin rows [row0, row1).
void BinSphereTile(const SphereCatalogue& index, const tgx_tile& tile, std::size_t row0,
                   std::size_t row1, std::vector<std::uint32_t>& candidates, TileBins& out) {
    std::vector<tgx_window> windows;
    SphereWindows(index, tile, row0 * kSide, row1 * kSide, candidates, windows);
    ListWindows(windows, tile, kSide, row0, row1, out);
}

// The (cell, listed crater) pairs of tiles [first, last) with the cell centre strictly within
// the crater's reach along the sphere: arc < reach exactly when the chord is below
// 2 R sin(reach / 2R).
void SpherePairs(const tg_tile_desc* descs, std::size_t first, std::size_t last,
                 const tg_bin* bins, const std::uint32_t* entries, const double* qx,
                 const double* qy, const double* qz, const double* chord2, const double* cost,
                 double radius, std::uint64_t& listed, std::uint64_t& inside, double& ops) {
    std::vector<double> cosLon, sinLon, cosLat, sinLat;
    for (std::size_t t = first; t < last; ++t) {
        const tg_tile_desc& d = descs[t];
        cosLon.resize(d.width);
        sinLon.resize(d.width);
        cosLat.resize(d.height);
        sinLat.resize(d.height);
        for (std::uint32_t c = 0; c < d.width; ++c) {
            const tg::SinCos a = tg::SinCosDegrees(d.origin_x + (double(c) + 0.5) * d.pixel + 0.0);
            cosLon[c] = a.cos;
            sinLon[c] = a.sin;
        }
        for (std::uint32_t r = 0; r < d.height; ++r) {
            const tg::SinCos a = tg::SinCosDegrees(d.origin_y + (double(r) + 0.5) * -d.pixel);
            cosLat[r] = a.cos;
            sinLat[r] = a.sin;
        }
        const std::uint32_t binsDown = (d.height + kSide - 1) / kSide;
        for (std::uint32_t b = 0; b < d.bins_across * binsDown; ++b) {
            const tg_bin& bin = bins[d.bin_first + b];
            const std::uint32_t c0 = (b % d.bins_across) * kSide, r0 = (b / d.bins_across) * kSide;
            const std::uint32_t c1 = std::min(d.width, c0 + kSide), r1 = std::min(d.height, r0 + kSide);
            for (std::uint32_t e = bin.first; e < bin.first + bin.count; ++e) {
                const std::uint32_t k = entries[e];
                std::uint64_t in = 0;
                for (std::uint32_t r = r0; r < r1; ++r) {
                    for (std::uint32_t c = c0; c < c1; ++c) {
                        const double px = radius * (cosLat[r] * cosLon[c]) - qx[k];
                        const double py = radius * (cosLat[r] * sinLon[c]) - qy[k];
                        const double pz = radius * sinLat[r] - qz[k];
                        in += px * px + py * py + pz * pz < chord2[k] ? 1u : 0u;
                    }
                }
                listed += std::uint64_t(r1 - r0) * (c1 - c0);
                inside += in;
                ops += double(in) * cost[k];
            }
        }
    }
}

}  // namespace

std::shared_ptr<const SphereCatalogue> MakeSphereCatalogue(const double* lon, const double* lat,
                                                           const double* reach,
                                                           const double* diameter,
                                                           std::size_t n, double radius) {
    if ((n > 0 && (lon == nullptr || lat == nullptr || reach == nullptr || diameter == nullptr)) ||
        !(radius > 0.0 && std::isfinite(radius))) {
        throw std::invalid_argument("a sphere catalogue needs its columns and a positive radius");
    }
    auto c = std::make_shared<SphereCatalogue>();
    c->radius = radius;
    c->lon.assign(lon, lon + n);
    c->lat.assign(lat, lat + n);
    c->diameter.assign(diameter, diameter + n);
    c->reachDeg.resize(n);
    c->lonHalfWidth.resize(n);
    std::map<int, std::vector<std::uint32_t>> byLevel;
    for (std::size_t k = 0; k < n; ++k) {
        if (!std::isfinite(lon[k]) || !(std::fabs(lat[k]) <= 90.0) || !(reach[k] >= 0.0) ||
            !(diameter[k] > 0.0 && std::isfinite(diameter[k]))) {
            throw std::invalid_argument("crater " + std::to_string(k) + " is malformed");
        }
        // Footprints in degrees of arc: a crater's offsets are arc lengths (KernelFrameOffset).
        c->reachDeg[k] = reach[k] / radius / kDegree;
        c->lonHalfWidth[k] = LonHalfWidth(lat[k], c->reachDeg[k]);
        byLevel[std::ilogb(diameter[k])].push_back(std::uint32_t(k));
    }
    for (const auto& [exponent, members] : byLevel) {
        (void)exponent;
        SphereLevel level;
        for (const std::uint32_t k : members) {
            level.maxReachDeg = std::max(level.maxReachDeg, c->reachDeg[k]);
            level.maxDiameter = std::max(level.maxDiameter, c->diameter[k]);
        }
        // Cells at least twice the largest footprint, doubled until they hold a few craters
        // each on average, so that a tile visits few cells and reads few craters per cell.
        level.cellDeg = std::clamp(2.0 * level.maxReachDeg, 1e-7, 1.0);
        std::vector<std::int64_t> keys(members.size());
        for (;;) {
            level.lonCells = std::int64_t(std::ceil(360.0 / level.cellDeg));
            level.latCells = std::int64_t(std::ceil(180.0 / level.cellDeg)) + 1;
            for (std::size_t m = 0; m < members.size(); ++m) {
                const std::uint32_t k = members[m];
                keys[m] = level.Key(level.LonCell(c->lon[k]), level.LatCell(c->lat[k]));
            }
            if (level.cellDeg >= 1.0) {
                break;
            }
            std::vector<std::int64_t> distinct = keys;
            std::sort(distinct.begin(), distinct.end());
            const std::size_t occupied = std::size_t(
                std::unique(distinct.begin(), distinct.end()) - distinct.begin());
            if (members.size() >= 8 * occupied) {
                break;
            }
            level.cellDeg = std::min(1.0, 2.0 * level.cellDeg);
        }
        for (std::size_t m = 0; m < members.size(); ++m) {
            level.cells[keys[m]].push_back(members[m]);
        }
        c->levels.push_back(std::move(level));
    }
    return c;
}

void BinSphereBatch(const SphereCatalogue& catalogue, const tgx_tile* tiles,
                    std::size_t tileCount, unsigned threads, std::vector<tg_tile_desc>& d,
                    std::vector<tg_bin>& b, std::vector<std::uint32_t>& e) {
    for (std::size_t t = 0; t < tileCount; ++t) {
        CheckTile(tiles[t], t);
        // A longitude-latitude tile lies between the poles.
        const tgx_tile& tile = tiles[t];
        if (!(tile.origin_y <= 90.0 && tile.origin_y - tile.height * tile.pixel >= -90.0)) {
            throw std::invalid_argument("tile " + std::to_string(t) +
                                        " reaches past a pole (its latitudes leave [-90, 90])");
        }
    }
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
            BinSphereTile(catalogue, tiles[s.tile], s.row0, s.row1, candidates, per[n]);
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
    TileDescs(tiles, tileCount, kSide, d);
    b.clear();
    e.clear();
    AppendStrips(strips, per, b, e);
}

double IndexRadius(const tgx_index& index) {
    return index.sphere ? index.sphere->radius : 0.0;
}

}  // namespace tgh

extern "C" {

int tgx_sphere_positions(double west, double north, double step, uint32_t width, uint32_t height,
                         double radius, double* xyz, char* err, size_t errLen) {
    return tgh::Guard(err, errLen, [&] {
        if (xyz == nullptr || !(step > 0.0) || !(radius > 0.0)) {
            throw std::invalid_argument("tgx_sphere_positions: need a positive step and radius");
        }
        for (uint32_t r = 0; r < height; ++r) {
            for (uint32_t c = 0; c < width; ++c) {
                const tg::Position p = tg::CellPosition(west, north, step, c, r, radius);
                double* o = xyz + 3 * (std::size_t(r) * width + c);
                o[0] = p.x;
                o[1] = p.y;
                o[2] = p.z;
            }
        }
    });
}

int tgx_sphere_split_points(double west, double north, double step, uint32_t width,
                            uint32_t height, double radius, tg_split_point* out, char* err,
                            size_t errLen) {
    return tgh::Guard(err, errLen, [&] {
        if (out == nullptr || !(step > 0.0) || !(radius > 0.0)) {
            throw std::invalid_argument("tgx_sphere_split_points: need a positive step and radius");
        }
        for (uint32_t r = 0; r < height; ++r) {
            for (uint32_t c = 0; c < width; ++c) {
                const tg::Position p = tg::CellPosition(west, north, step, c, r, radius);
                tg_split_point& s = out[std::size_t(r) * width + c];
                tgh::Split(p.x, s.x, s.x_low);
                tgh::Split(p.y, s.y, s.y_low);
                tgh::Split(p.z, s.z, s.z_low);
            }
        }
    });
}

int tgx_sphere_frames(const double* lon, const double* lat, size_t n, double radius,
                      tg_crater_frame* out, char* err, size_t errLen) {
    return tgh::Guard(err, errLen, [&] {
        if ((n > 0 && (lon == nullptr || lat == nullptr || out == nullptr)) || !(radius > 0.0)) {
            throw std::invalid_argument("tgx_sphere_frames: null argument or radius not positive");
        }
        for (size_t i = 0; i < n; ++i) {
            const double l = lon[i] * tgh::kDegree, p = lat[i] * tgh::kDegree;
            const double cosLat = std::cos(p);
            const double ux = cosLat * std::cos(l), uy = cosLat * std::sin(l), uz = std::sin(p);
            tg_crater_frame& f = out[i];
            tgh::Split(ux * radius, f.x, f.x_low);
            tgh::Split(uy * radius, f.y, f.y_low);
            tgh::Split(uz * radius, f.z, f.z_low);
            const double sinLon = std::sin(l), cosLon = std::cos(l), sinLat = std::sin(p);
            f.east_x = static_cast<float>(-sinLon);
            f.east_y = static_cast<float>(cosLon);
            f.east_z = 0.0f;
            f.north_x = static_cast<float>(-sinLat * cosLon);
            f.north_y = static_cast<float>(-sinLat * sinLon);
            f.north_z = static_cast<float>(std::cos(p));
            f.sphere_radius = static_cast<float>(radius);
        }
    });
}

int tgx_sphere_bin(const double* lon, const double* lat, const double* reach,
                   const double* diameter, size_t n, double radius, const tgx_tile* tiles,
                   size_t tileCount, unsigned threads, tg_tile_desc** descs, tg_bin** bins,
                   size_t* binCount, uint32_t** entries, size_t* entryCount, char* err,
                   size_t errLen) {
    return tgh::Guard(err, errLen, [&] {
        if (descs == nullptr || bins == nullptr || entries == nullptr || binCount == nullptr ||
            entryCount == nullptr || (tileCount > 0 && tiles == nullptr)) {
            throw std::invalid_argument("tgx_sphere_bin: null argument");
        }
        const auto catalogue = tgh::MakeSphereCatalogue(lon, lat, reach, diameter, n, radius);
        std::vector<tg_tile_desc> d;
        std::vector<tg_bin> b;
        std::vector<uint32_t> e;
        tgh::BinSphereBatch(*catalogue, tiles, tileCount, threads, d, b, e);
        *descs = tgh::Export(d);
        *bins = tgh::Export(b);
        *binCount = b.size();
        *entries = tgh::Export(e);
        *entryCount = e.size();
    });
}

int tgx_sphere_index_create(const double* lon, const double* lat, const double* reach,
                            const double* diameter, size_t n, double radius, tgx_index** out,
                            char* err, size_t errLen) {
    return tgh::Guard(err, errLen, [&] {
        if (out == nullptr) {
            throw std::invalid_argument("tgx_sphere_index_create: null argument");
        }
        *out = nullptr;
        auto index = std::make_unique<tgx_index>();
        index->sphere = tgh::MakeSphereCatalogue(lon, lat, reach, diameter, n, radius);
        index->cx = index->sphere->lon;
        index->cy = index->sphere->lat;
        index->reach.assign(reach, reach + n);
        index->diameter = index->sphere->diameter;
        *out = index.release();
    });
}

int tgx_sphere_pair_ops(const tg_tile_desc* descs, size_t tileCount, const tg_bin* bins,
                        size_t binCount, const uint32_t* entries, size_t entryCount,
                        const double* lon, const double* lat, const double* reach,
                        const double* cost, size_t n, double radius, unsigned threads,
                        uint64_t* listed, uint64_t* inside, double* ops, char* err,
                        size_t errLen) {
    return tgh::Guard(err, errLen, [&] {
        if (listed == nullptr || inside == nullptr || ops == nullptr ||
            (tileCount > 0 && descs == nullptr) || (binCount > 0 && bins == nullptr) ||
            (entryCount > 0 && entries == nullptr) ||
            (n > 0 && (lon == nullptr || lat == nullptr || reach == nullptr || cost == nullptr)) ||
            !(radius > 0.0)) {
            throw std::invalid_argument("tgx_sphere_pair_ops: null argument or radius not positive");
        }
        for (size_t i = 0; i < entryCount; ++i) {
            if (entries[i] >= n) {
                throw std::invalid_argument("tgx_sphere_pair_ops: entry past the catalogue");
            }
        }
        for (size_t t = 0; t < tileCount; ++t) {
            const tg_tile_desc& d = descs[t];
            const uint64_t binsDown = (d.height + tgh::kSide - 1) / tgh::kSide;
            if (uint64_t(d.bin_first) + uint64_t(d.bins_across) * binsDown > binCount) {
                throw std::invalid_argument("tgx_sphere_pair_ops: a tile's bins run past the list");
            }
        }
        std::vector<double> qx(n), qy(n), qz(n), chord2(n);
        for (size_t k = 0; k < n; ++k) {
            const tg::Position q = tg::SpherePosition(lon[k], lat[k], radius);
            qx[k] = q.x;
            qy[k] = q.y;
            qz[k] = q.z;
            const double half = std::min(reach[k] / (2.0 * radius), 0.5 * tgh::kPi);
            const double chord = 2.0 * radius * std::sin(half);
            chord2[k] = chord * chord;
        }
        if (threads == 0) {
            threads = std::max(1u, std::thread::hardware_concurrency());
        }
        threads = unsigned(std::min<size_t>(threads, std::max<size_t>(tileCount, 1)));
        std::vector<uint64_t> l(threads, 0), in(threads, 0);
        std::vector<double> o(threads, 0.0);
        std::atomic<size_t> next{0};
        auto work = [&](unsigned w) {
            for (;;) {
                const size_t t = next.fetch_add(1);
                if (t >= tileCount) {
                    return;
                }
                tgh::SpherePairs(descs, t, t + 1, bins, entries, qx.data(), qy.data(), qz.data(),
                                 chord2.data(), cost, radius, l[w], in[w], o[w]);
            }
        };
        std::vector<std::thread> pool;
        for (unsigned i = 1; i < threads; ++i) {
            pool.emplace_back(work, i);
        }
        work(0);
        for (std::thread& th : pool) {
            th.join();
        }
        *listed = 0;
        *inside = 0;
        *ops = 0.0;
        for (unsigned i = 0; i < threads; ++i) {
            *listed += l[i];
            *inside += in[i];
            *ops += o[i];
        }
    });
}

}  // extern "C"
