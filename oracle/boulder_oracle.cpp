// SPDX-License-Identifier: BSD-3-Clause
//
// ORACLE — test infrastructure only (see noise_oracle.cpp). A CPU restatement of the reference's
// boulder relief in FP64, as BoulderLayer bakes it on the CPU:
//   DomeHeight                  src/layers/boulders/boulder_relief.cpp:44-57
//   Boulder::CutByFacets        include/terraingen/layers/boulders/boulder.h:99-104
//   BoulderFacet::HeightAt      boulder.h:37-41
//   BoulderLayer::AddRelief     boulder_relief.cpp:244-262 (no fillet)
//   BoulderLayer::PrepareRelief boulder_relief.cpp:373-423 (weight, band width, box)
//   DrawFeatures                include/terraingen/sources/sample_block.h:107-131 (base, order)
//   Surface::SphereOffset       src/geometry/surface.cpp:214-229 (offsets on a sphere), with
//   UnitVector and LocalAxes    surface.cpp:18-24,45-50
// `or_boulder_grid` keeps the reference's FP32 store after every feature (its heights grid is
// FP32); `or_boulder_grid_f64` keeps FP64 throughout, the oracle the device's FP32 is held to;
// `or_boulder_body_f64` is the same on a longitude-latitude grid of a sphere.

#include "terraingen_amd.h"

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <vector>

namespace {

double Weight(double diameter, double spacing) {
    const double x = diameter / spacing;
    if (x <= 2.0) {
        return 0.0;
    }
    if (x >= 4.0) {
        return 1.0;
    }
    const double t = std::clamp((x - 2.0) / (4.0 - 2.0), 0.0, 1.0);
    return 1.0 - (1.0 - t * t * (3.0 - 2.0 * t));
}

double Dome(const tgx_boulder& b, double distance, double width) {
    const double radius = 0.5 * b.diameter;
    if (!(width > 0.0)) {
        const double r2 = (distance * distance) / (radius * radius);
        return r2 < 1.0 ? b.exposed * std::sqrt(1.0 - r2) : 0.0;
    }
    const auto area = [](double u) { return 0.5 * (u * std::sqrt(1.0 - u * u) + std::asin(u)); };
    const double low = std::clamp((distance - 0.5 * width) / radius, -1.0, 1.0);
    const double high = std::clamp((distance + 0.5 * width) / radius, -1.0, 1.0);
    if (!(high > low)) {
        return 0.0;
    }
    return b.exposed * (area(high) - area(low)) * radius / width;
}

double Relief(const tgx_boulder& b, double dx, double dy, double width) {
    double top = Dome(b, std::hypot(dx, dy), width);
    for (std::uint32_t k = 0; k < b.facet_count; ++k) {
        const double azimuth = b.facets[3 * k], angle = b.facets[3 * k + 1];
        const double along = dx * std::cos(azimuth) + dy * std::sin(azimuth);
        top = std::fmin(top, (b.facets[3 * k + 2] - along) * std::tan(angle));
    }
    return std::fmax(top, 0.0);
}

// The grid cells of a relief box: floor/ceil of its corners in pixel coordinates, clamped.
void Window(const tgx_grid& g, double x, double y, double reach, std::size_t& c0, std::size_t& c1,
            std::size_t& r0, std::size_t& r1) {
    const auto clampTo = [](double v, std::size_t limit) {
        return static_cast<std::size_t>(std::clamp(v, 0.0, static_cast<double>(limit)));
    };
    c0 = clampTo(std::floor((x - reach - g.origin_x) / g.pixel), g.width);
    c1 = clampTo(std::ceil((x + reach - g.origin_x) / g.pixel), g.width);
    r0 = clampTo(std::floor((g.origin_y - (y + reach)) / g.pixel), g.height);
    r1 = clampTo(std::ceil((g.origin_y - (y - reach)) / g.pixel), g.height);
}

template <class H>
void Draw(const tgx_boulder* boulders, std::size_t n, const tgx_grid& g, int dome, H* heights) {
    const double width = dome != 0 ? g.pixel : 0.0;
    const std::vector<H> base(heights, heights + std::size_t(g.width) * g.height);
    for (std::size_t i = 0; i < n; ++i) {
        const tgx_boulder& b = boulders[i];
        const double weight = Weight(b.diameter, g.pixel);
        if (b.model != 0 || !(weight > 0.0)) {
            continue;
        }
        const double reach = 0.5 * (b.diameter + width);
        std::size_t c0, c1, r0, r1;
        Window(g, b.x, b.y, reach, c0, c1, r0, r1);
        for (std::size_t r = r0; r < r1; ++r) {
            const double cy = g.origin_y + (double(r) + 0.5) * -g.pixel;
            for (std::size_t c = c0; c < c1; ++c) {
                const double cx = g.origin_x + (double(c) + 0.5) * g.pixel + 0.0;
                const double relief = Relief(b, cx - b.x, cy - b.y, width);
                H& cell = heights[r * g.width + c];
                if (relief > 0.0) {
                    cell = static_cast<H>(std::fmax(static_cast<double>(cell),
                                                    static_cast<double>(base[r * g.width + c]) +
                                                        weight * relief));
                }
            }
        }
    }
}

constexpr double kPi = 3.141592653589793238462643383279502884;
constexpr double kDegree = kPi / 180.0;

struct V3 {
    double x, y, z;
};

double Dot(const V3& a, const V3& b) {
    return a.x * b.x + a.y * b.y + a.z * b.z;
}

V3 Unit(double lonDeg, double latDeg) {
    const double lon = lonDeg * kDegree, lat = latDeg * kDegree;
    const double cosLat = std::cos(lat);
    return {cosLat * std::cos(lon), cosLat * std::sin(lon), std::sin(lat)};
}

// The offset east and north (m) of `point` from `center` (longitude, latitude in degrees) on a
// sphere of `radius`: the direction along the tangent plane at the centre, the length the arc.
void SphereOffset(double radius, double clon, double clat, double plon, double plat, double& e,
                  double& n) {
    const V3 c = Unit(clon, clat), p = Unit(plon, plat);
    const double lon = clon * kDegree, lat = clat * kDegree;
    const V3 east{-std::sin(lon), std::cos(lon), 0.0};
    const V3 north{-std::sin(lat) * std::cos(lon), -std::sin(lat) * std::sin(lon), std::cos(lat)};
    const V3 cross{c.y * p.z - c.z * p.y, c.z * p.x - c.x * p.z, c.x * p.y - c.y * p.x};
    const double angle = std::atan2(std::sqrt(Dot(cross, cross)), Dot(c, p));
    const double pe = Dot(p, east), pn = Dot(p, north);
    const double across = std::hypot(pe, pn);
    if (across == 0.0) {
        e = 0.0;
        n = angle == 0.0 ? 0.0 : radius * angle;
        return;
    }
    const double scale = radius * angle / across;
    e = pe * scale;
    n = pn * scale;
}

// The dome relief on a longitude-latitude grid (degrees) of a sphere: each boulder drawn at the
// grid's spacing (pixel x radius x pi / 180) over the cells within a window of its reach in
// degrees, a cell's offset the SphereOffset from the boulder's centre.
void DrawBody(const tgx_boulder* boulders, std::size_t n, const tgx_grid& g, double radius,
              int dome, double* heights) {
    const double spacing = g.pixel * radius * kDegree;
    const double width = dome != 0 ? spacing : 0.0;
    const std::vector<double> base(heights, heights + std::size_t(g.width) * g.height);
    for (std::size_t i = 0; i < n; ++i) {
        const tgx_boulder& b = boulders[i];
        const double weight = Weight(b.diameter, spacing);
        if (b.model != 0 || !(weight > 0.0)) {
            continue;
        }
        const double reach = 0.5 * (b.diameter + width);
        // Its reach in degrees of latitude, and of longitude at the highest latitude it spans
        // (a pole within reach takes every column), with a cell to spare either side.
        const double dLat = reach / (radius * kDegree);
        const double top = std::fabs(b.y) + dLat;
        const double dLon = top >= 89.0 ? 360.0 : dLat / std::cos(top * kDegree);
        std::size_t c0, c1, r0, r1;
        const double pad = g.pixel;
        const auto clampTo = [](double v, std::size_t limit) {
            return static_cast<std::size_t>(std::clamp(v, 0.0, static_cast<double>(limit)));
        };
        c0 = clampTo(std::floor((b.x - dLon - pad - g.origin_x) / g.pixel), g.width);
        c1 = clampTo(std::ceil((b.x + dLon + pad - g.origin_x) / g.pixel), g.width);
        r0 = clampTo(std::floor((g.origin_y - (b.y + dLat + pad)) / g.pixel), g.height);
        r1 = clampTo(std::ceil((g.origin_y - (b.y - dLat - pad)) / g.pixel), g.height);
        if (dLon >= 180.0) {
            c0 = 0;
            c1 = g.width;
        }
        for (std::size_t r = r0; r < r1; ++r) {
            const double lat = g.origin_y + (double(r) + 0.5) * -g.pixel;
            for (std::size_t c = c0; c < c1; ++c) {
                const double lon = g.origin_x + (double(c) + 0.5) * g.pixel + 0.0;
                double dx, dy;
                SphereOffset(radius, b.x, b.y, lon, lat, dx, dy);
                const double relief = Relief(b, dx, dy, width);
                double& cell = heights[r * g.width + c];
                if (relief > 0.0) {
                    cell = std::fmax(cell, base[r * g.width + c] + weight * relief);
                }
            }
        }
    }
}

}  // namespace

extern "C" {

void or_boulder_body_f64(const tgx_boulder* boulders, std::size_t n, const tgx_grid* grid,
                         double radius, int dome, double* heights) {
    DrawBody(boulders, n, *grid, radius, dome, heights);
}

// SphereOffset of a point from a centre (longitude, latitude in degrees), FP64.
void or_sphere_offset(double radius, double clon, double clat, double plon, double plat,
                      double* out) {
    SphereOffset(radius, clon, clat, plon, plat, out[0], out[1]);
}

void or_boulder_grid(const tgx_boulder* boulders, std::size_t n, const tgx_grid* grid, int dome,
                     float* heights) {
    Draw(boulders, n, *grid, dome, heights);
}

void or_boulder_grid_f64(const tgx_boulder* boulders, std::size_t n, const tgx_grid* grid,
                         int dome, double* heights) {
    Draw(boulders, n, *grid, dome, heights);
}

// One boulder's relief at (dx, dy) from its centre, FP64.
double or_boulder_relief_f64(const tgx_boulder* boulder, double dx, double dy, double width) {
    return Relief(*boulder, dx, dy, width);
}

}  // extern "C"
