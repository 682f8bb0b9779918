// SPDX-License-Identifier: BSD-3-Clause
//
// Cube-sphere placement cells and sphere regions (cube_cells.h). Every quantity a drawn crater
// depends on (cell index, edges, solid angle, sample point, region tests) is evaluated with the
// reference's floating-point steps; the covering search is the reference's too (a direct ring
// around the region's home cell when the region is small, else seed points every quarter cell
// with their rings, kept where the ring cell's centre lies within its corner angle of the region),
// so the same cells are drawn and the expected counts agree.

#include "cube_cells.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <stdexcept>

namespace tgh {

namespace {

constexpr double kPi = 3.141592653589793238462643383279502884;
constexpr double kDegree = kPi / 180.0;
constexpr double kRadian = 180.0 / kPi;
constexpr double kQuarter = kPi / 4.0;  // the equiangular half-width of a face

double Dot(const Vec3& a, const Vec3& b) {
    return a.x * b.x + a.y * b.y + a.z * b.z;
}

Vec3 Unit(const Vec3& v) {
    const double length = std::sqrt(Dot(v, v));
    const double inv = length > 0.0 ? 1.0 / length : 1.0;
    return length > 0.0 ? Vec3{v.x * inv, v.y * inv, v.z * inv} : v;
}

// The solid angle of the gnomonic rectangle [0, a] x [0, b] on a face, signed.
double Corner(double a, double b) {
    return std::atan(a * b / std::sqrt(1.0 + a * a + b * b));
}

// The gnomonic coordinate of edge k of n along a face.
double Edge(std::int32_t k, std::int32_t n) {
    return std::tan(-kQuarter + 2.0 * kQuarter * static_cast<double>(k) / static_cast<double>(n));
}

void SortUnique(std::vector<CellIndex>& cells) {
    std::sort(cells.begin(), cells.end());
    cells.erase(std::unique(cells.begin(), cells.end()), cells.end());
}

}  // namespace

// ---- sphere regions --------------------------------------------------------------------------

SphereRegions::SphereRegions(double radius) : m_radius(radius), m_unitLength(radius * kDegree) {
    if (!(radius > 0.0 && std::isfinite(radius))) {
        throw std::invalid_argument("A sphere's radius must be positive and finite");
    }
}

Vec3 UnitVector(double lonDeg, double latDeg) {
    const double lon = lonDeg * kDegree;
    const double lat = latDeg * kDegree;
    const double cosLat = std::cos(lat);
    return {cosLat * std::cos(lon), cosLat * std::sin(lon), std::sin(lat)};
}

void SphereRegions::Coordinates(const Vec3& d, double& lon, double& lat) const {
    lon = std::atan2(d.y, d.x) * kRadian;
    lat = std::atan2(d.z, std::hypot(d.x, d.y)) * kRadian;
    if (lon == -180.0) {
        lon = 180.0;
    }
}

Box SphereRegions::Expand(const Box& region, double meters) const {
    const double dLat = meters / m_unitLength;
    const double angle = meters / m_radius;
    Box out{region.minX, std::max(-90.0, region.minY - dLat), region.maxX,
            std::min(90.0, region.maxY + dLat)};
    // The widest longitude span a cap of `angle` takes, at the latitude nearest a pole.
    const double poleward = std::max(std::fabs(region.minY), std::fabs(region.maxY));
    const double reach = std::sin(std::min(angle, kPi / 2.0));
    const double cosLat = std::cos(std::min(poleward, 90.0) * kDegree);
    const double dLon =
        angle < kPi / 2.0 && reach < cosLat ? std::asin(reach / cosLat) * kRadian : 360.0;
    const bool everyLongitude = out.minY <= -90.0 || out.maxY >= 90.0 || dLon >= 180.0 ||
                                region.maxX - region.minX + 2.0 * dLon >= 360.0;
    out.minX = everyLongitude ? -180.0 : region.minX - dLon;
    out.maxX = everyLongitude ? 180.0 : region.maxX + dLon;
    return out;
}

bool SphereRegions::Contains(const Box& region, double lon, double lat) const {
    if (region.Contains(lon, lat)) {
        return true;
    }
    const double at = lat >= 90.0 && region.maxY >= 90.0 ? region.minY : lat;
    return region.Contains(lon, at) || region.Contains(lon - 360.0, at) ||
           region.Contains(lon + 360.0, at);
}

// ---- cube cells ------------------------------------------------------------------------------

const CubeCells::Face CubeCells::kFaces[6] = {
    {{1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}},
    {{-1.0, 0.0, 0.0}, {0.0, -1.0, 0.0}, {0.0, 0.0, 1.0}},
    {{0.0, 1.0, 0.0}, {-1.0, 0.0, 0.0}, {0.0, 0.0, 1.0}},
    {{0.0, -1.0, 0.0}, {1.0, 0.0, 0.0}, {0.0, 0.0, 1.0}},
    {{0.0, 0.0, 1.0}, {0.0, 1.0, 0.0}, {-1.0, 0.0, 0.0}},
    {{0.0, 0.0, -1.0}, {0.0, 1.0, 0.0}, {1.0, 0.0, 0.0}},
};

namespace {

// normal + a u + b v.
template <class F> Vec3 FacePoint(const F& f, double a, double b) {
    return {f.normal.x + a * f.u.x + b * f.v.x, f.normal.y + a * f.u.y + b * f.v.y,
            f.normal.z + a * f.u.z + b * f.v.z};
}

}  // namespace

CubeCells::CubeCells(double radius) : m_regions(radius) {}

std::int32_t CubeCells::PerEdge(double size) const {
    if (!(size > 0.0 && std::isfinite(size))) {
        throw std::invalid_argument("Cell size must be positive and finite");
    }
    const double n = std::ceil(m_regions.Radius() * 2.0 * kQuarter / size);
    if (!(n < static_cast<double>(std::numeric_limits<std::int32_t>::max() / 6))) {
        throw std::invalid_argument("Cells are too small for the sphere");
    }
    return std::max<std::int32_t>(1, static_cast<std::int32_t>(n));
}

CubeCells::Rect CubeCells::Bounds(const CellIndex& cell, std::int32_t n) const {
    const std::int32_t col = cell.i % n;
    return {Edge(col, n), Edge(col + 1, n), Edge(cell.j, n), Edge(cell.j + 1, n)};
}

double CubeCells::Area(const CellIndex& cell, double size) const {
    const Rect r = Bounds(cell, PerEdge(size));
    const double solid =
        Corner(r.a2, r.b2) - Corner(r.a1, r.b2) - Corner(r.a2, r.b1) + Corner(r.a1, r.b1);
    return m_regions.Radius() * m_regions.Radius() * solid;
}

void CubeCells::Sample(double u, double v, const CellIndex& cell, double size, double& lon,
                       double& lat) const {
    const std::int32_t n = PerEdge(size);
    const Rect r = Bounds(cell, n);
    // Along u: the solid angle of the strip from a1 is monotone in a; bisect it to u of the whole.
    // Its a1 corners are the same at every step (the same values, so the same bits).
    const double cornerA1B2 = Corner(r.a1, r.b2);
    const double cornerA1B1 = Corner(r.a1, r.b1);
    const auto strip = [&](double a) {
        return Corner(a, r.b2) - Corner(a, r.b1) - cornerA1B2 + cornerA1B1;
    };
    const double whole = strip(r.a2);
    double lo = r.a1, hi = r.a2;
    for (int step = 0; step < 64; ++step) {
        const double mid = 0.5 * (lo + hi);
        // Once the midpoint rounds to an end, this step either leaves lo and hi as they are or
        // sets them equal, and every later step computes the same midpoint and leaves them: the
        // remaining steps change nothing, and stopping here gives the 64 steps' bits.
        const bool settled = mid == lo || mid == hi;
        if (strip(mid) < u * whole) {
            lo = mid;
        } else {
            hi = mid;
        }
        if (settled) {
            break;
        }
    }
    const double a = 0.5 * (lo + hi);
    // Along v at that a the solid angle is linear in b / sqrt(1 + a^2 + b^2).
    const auto ratio = [a](double b) { return b / std::sqrt(1.0 + a * a + b * b); };
    const double k = ratio(r.b1) + v * (ratio(r.b2) - ratio(r.b1));
    const double b = k * std::sqrt(1.0 + a * a) / std::sqrt(1.0 - k * k);
    m_regions.Coordinates(FacePoint(kFaces[cell.i / n], a, b), lon, lat);
}

CellIndex CubeCells::Home(double lonDeg, double latDeg, double size) const {
    const std::int32_t n = PerEdge(size);
    const Vec3 d = UnitVector(lonDeg, latDeg);
    // The face whose normal the direction is nearest (the first on ties).
    std::size_t face = 0;
    double best = -2.0;
    for (std::size_t f = 0; f < 6; ++f) {
        const double along = Dot(d, kFaces[f].normal);
        if (along > best) {
            best = along;
            face = f;
        }
    }
    const Face& f = kFaces[face];
    const double along = Dot(d, f.normal);
    const double cells = static_cast<double>(n);
    const double s = (std::atan(Dot(d, f.u) / along) + kQuarter) / (2.0 * kQuarter) * cells;
    const double t = (std::atan(Dot(d, f.v) / along) + kQuarter) / (2.0 * kQuarter) * cells;
    const auto index = [n](double x) {
        return std::clamp<std::int32_t>(static_cast<std::int32_t>(std::floor(x)), 0, n - 1);
    };
    return {static_cast<std::int32_t>(face) * n + index(s), index(t)};
}

void CubeCells::AcrossEdge(const Face& face, std::int32_t col, std::int32_t row, std::int32_t n,
                           double size, std::vector<CellIndex>& out) const {
    // A cell off the face's edge: the cells of nine points of it, placed on the neighbouring face.
    const double step = 2.0 * kQuarter / static_cast<double>(n);
    for (const double su : {0.1, 0.5, 0.9}) {
        for (const double sv : {0.1, 0.5, 0.9}) {
            const double alpha = -kQuarter + (static_cast<double>(col) + su) * step;
            const double beta = -kQuarter + (static_cast<double>(row) + sv) * step;
            if (!(std::fabs(alpha) < 2.0 * kQuarter - step &&
                  std::fabs(beta) < 2.0 * kQuarter - step)) {
                continue;
            }
            double lon, lat;
            m_regions.Coordinates(FacePoint(face, std::tan(alpha), std::tan(beta)), lon, lat);
            out.push_back(Home(lon, lat, size));
        }
    }
}

void CubeCells::Ring(const CellIndex& cell, double size, std::vector<CellIndex>& out) const {
    const std::int32_t n = PerEdge(size);
    const std::int32_t face = cell.i / n;
    const std::int32_t col = cell.i % n;
    std::vector<CellIndex> cells;
    for (std::int32_t dj = -1; dj <= 1; ++dj) {
        for (std::int32_t di = -1; di <= 1; ++di) {
            const std::int32_t c = col + di;
            const std::int32_t r = cell.j + dj;
            if (c >= 0 && c < n && r >= 0 && r < n) {
                cells.push_back({face * n + c, r});
            } else {
                AcrossEdge(kFaces[face], c, r, n, size, cells);
            }
        }
    }
    SortUnique(cells);
    out.insert(out.end(), cells.begin(), cells.end());
}

std::vector<CellIndex> CubeCells::Near(const Box& box, double size) const {
    // Moving dtheta along the sphere moves a face's equiangular angle by at most dtheta /
    // sqrt(1 - v^2): under 1.53 dtheta within half a cell of a face of 16 cells or more per edge.
    // So when twice the region's angular half-size is under a cell, the 3 x 3 ring around the
    // region's home cell, each cell widened by that much, holds every cell the region reaches.
    constexpr std::int32_t kMinCells = 16;
    constexpr double kAngleRate = 2.0;
    constexpr double kMaxLatitude = 80.0;
    const std::int32_t n = PerEdge(size);
    const double halfLon = 0.5 * (box.maxX - box.minX) * kDegree;
    const double halfLat = 0.5 * (box.maxY - box.minY) * kDegree;
    const double nearestLat = box.minY > 0.0 ? box.minY : (box.maxY < 0.0 ? -box.maxY : 0.0);
    const double rho = halfLat + halfLon * std::cos(nearestLat * kDegree);
    const double step = 2.0 * kQuarter / static_cast<double>(n);
    if (n < kMinCells || std::fmax(-box.minY, box.maxY) > kMaxLatitude ||
        !(kAngleRate * rho < step)) {
        return {};
    }
    const Vec3 centre = UnitVector(0.5 * (box.minX + box.maxX), 0.5 * (box.minY + box.maxY));
    // The centre's face and equiangular coordinates (as Home).
    std::size_t homeFace = 0;
    double best = -2.0;
    for (std::size_t f = 0; f < 6; ++f) {
        const double along = Dot(centre, kFaces[f].normal);
        if (along > best) {
            best = along;
            homeFace = f;
        }
    }
    const double cellsN = static_cast<double>(n);
    const Face& hf = kFaces[homeFace];
    const double hAlong = Dot(centre, hf.normal);
    const double hs =
        (std::atan(Dot(centre, hf.u) / hAlong) + kQuarter) / (2.0 * kQuarter) * cellsN;
    const double ht =
        (std::atan(Dot(centre, hf.v) / hAlong) + kQuarter) / (2.0 * kQuarter) * cellsN;
    const auto index = [n](double x) {
        return std::clamp<std::int32_t>(static_cast<std::int32_t>(std::floor(x)), 0, n - 1);
    };
    const CellIndex home{static_cast<std::int32_t>(homeFace) * n + index(hs), index(ht)};
    const double widen = kAngleRate * rho;
    std::vector<CellIndex> ring;
    Ring(home, size, ring);
    std::vector<CellIndex> cells;
    for (const CellIndex& cell : ring) {
        const std::size_t faceIndex = static_cast<std::size_t>(cell.i / n);
        const Face& face = kFaces[faceIndex];
        const double along = Dot(centre, face.normal);
        const bool onHome = faceIndex == homeFace;
        const double alpha =
            onHome ? hs * step - kQuarter : std::atan2(Dot(centre, face.u), along);
        const double beta = onHome ? ht * step - kQuarter : std::atan2(Dot(centre, face.v), along);
        const double alpha0 = -kQuarter + static_cast<double>(cell.i % n) * step;
        const double beta0 = -kQuarter + static_cast<double>(cell.j) * step;
        if (along <= 0.0 || (alpha >= alpha0 - widen && alpha <= alpha0 + step + widen &&
                             beta >= beta0 - widen && beta <= beta0 + step + widen)) {
            cells.push_back(cell);
        }
    }
    return cells;
}

std::vector<CellIndex> CubeCells::Seeds(const Box& box, double size) const {
    // The home cells of points every quarter cell over the box.
    const double step = size / 4.0 / m_regions.UnitLength();
    const auto rows = static_cast<std::int64_t>(std::ceil((box.maxY - box.minY) / step));
    std::vector<CellIndex> seeds;
    for (std::int64_t r = 0; r <= rows; ++r) {
        const double lat = std::min(box.maxY, box.minY + static_cast<double>(r) * step);
        const double width = std::min(box.maxX - box.minX, 360.0);
        const double cosLat = std::cos(lat * kDegree);
        const auto columns = static_cast<std::int64_t>(std::ceil(width * cosLat / step));
        for (std::int64_t c = 0; c <= columns; ++c) {
            const double lon = columns == 0 ? box.minX
                                            : box.minX + width * static_cast<double>(c) /
                                                             static_cast<double>(columns);
            seeds.push_back(Home(lon, lat, size));
        }
    }
    SortUnique(seeds);
    return seeds;
}

std::vector<CellIndex> CubeCells::Reaching(const Box& box, const std::vector<CellIndex>& seeds,
                                           const std::vector<CellIndex>& cells,
                                           double size) const {
    // Cell edges are great circles, so a ring cell reaches the region only if its centre lies
    // within its farthest corner's angle of it: one expansion by the widest such angle tests all.
    const std::int32_t n = PerEdge(size);
    std::vector<CellIndex> ring;
    std::vector<std::pair<double, double>> centres;
    double widest = 0.0;
    for (const CellIndex& cell : cells) {
        if (std::binary_search(seeds.begin(), seeds.end(), cell)) {
            continue;
        }
        const Face& face = kFaces[cell.i / n];
        const Rect r = Bounds(cell, n);
        const Vec3 centre = Unit(FacePoint(face, 0.5 * (r.a1 + r.a2), 0.5 * (r.b1 + r.b2)));
        double nearest = 1.0;
        for (const double a : {r.a1, r.a2}) {
            for (const double b : {r.b1, r.b2}) {
                nearest = std::min(nearest, Dot(centre, Unit(FacePoint(face, a, b))));
            }
        }
        widest = std::max(widest, std::acos(std::clamp(nearest, -1.0, 1.0)));
        ring.push_back(cell);
        double lon, lat;
        m_regions.Coordinates(centre, lon, lat);
        centres.emplace_back(lon, lat);
    }
    if (ring.empty()) {
        return cells;
    }
    const Box reach =
        m_regions.Expand(box, m_regions.Radius() * (widest * (1.0 + 1e-9) + 1e-12));
    std::vector<CellIndex> kept = seeds;
    for (std::size_t k = 0; k < ring.size(); ++k) {
        if (m_regions.Contains(reach, centres[k].first, centres[k].second)) {
            kept.push_back(ring[k]);
        }
    }
    SortUnique(kept);
    return kept;
}

std::vector<CellIndex> CubeCells::Covering(const Box& region, double size) const {
    PerEdge(size);  // validates the size
    if (!(std::isfinite(region.minX) && std::isfinite(region.minY) &&
          std::isfinite(region.maxX) && std::isfinite(region.maxY) &&
          region.maxX >= region.minX && region.maxY >= region.minY)) {
        throw std::invalid_argument("Region must be finite and not inverted");
    }
    const Box box{region.minX, std::max(-90.0, region.minY), region.maxX,
                  std::min(90.0, region.maxY)};
    std::vector<CellIndex> near = Near(box, size);
    if (!near.empty()) {
        return near;
    }
    const std::vector<CellIndex> seeds = Seeds(box, size);
    std::vector<CellIndex> cells;
    for (const CellIndex& seed : seeds) {
        Ring(seed, size, cells);
    }
    SortUnique(cells);
    return Reaching(box, seeds, cells, size);
}

}  // namespace tgh
