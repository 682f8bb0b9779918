// SPDX-License-Identifier: BSD-3-Clause
//
// Cell positions on a sphere, the same bits on the host and on gfx950. The reference places a
// longitude-latitude raster's cell centre at Surface::Position (src/geometry/surface.cpp:45-50,
// 261-267): R (cos lat cos lon, cos lat sin lon, sin lat) with the C library's FP64 sine and
// cosine. Here the sine and cosine are this file's own: the angle is reduced in degrees, which is
// exact (x - 90 k is representable whenever it is no larger than x), then converted to radians and
// summed as Taylor series on [-pi/4, pi/4] by Horner's rule in a fixed order, so the host and the
// device (both built with FP contraction off) compute identical positions. They differ from the C
// library's by about an ulp, 2e-10 m at a lunar radius.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define TG_HOST_DEVICE __host__ __device__
#else
#define TG_HOST_DEVICE
#endif

namespace tg {

struct SinCos {
    double sin, cos;
};

// sin and cos of `degrees`.
TG_HOST_DEVICE inline SinCos SinCosDegrees(double degrees) {
    const double quarter = degrees / 90.0;
    // Nearest quarter turn, rounded half away from zero.
    const double k = quarter >= 0.0 ? static_cast<double>(static_cast<int64_t>(quarter + 0.5))
                                    : -static_cast<double>(static_cast<int64_t>(0.5 - quarter));
    const double r = (degrees - 90.0 * k) * 0.017453292519943295;  // pi / 180
    const double r2 = r * r;
    // r - r^3/3! + ... - r^15/15!, and 1 - r^2/2! + ... + r^16/16!.
    const double s =
        r + r * r2 *
                (-1.0 / 6.0 +
                 r2 * (1.0 / 120.0 +
                       r2 * (-1.0 / 5040.0 +
                             r2 * (1.0 / 362880.0 +
                                   r2 * (-1.0 / 39916800.0 +
                                         r2 * (1.0 / 6227020800.0 +
                                               r2 * (-1.0 / 1307674368000.0)))))));
    const double c =
        1.0 + r2 * (-0.5 +
                    r2 * (1.0 / 24.0 +
                          r2 * (-1.0 / 720.0 +
                                r2 * (1.0 / 40320.0 +
                                      r2 * (-1.0 / 3628800.0 +
                                            r2 * (1.0 / 479001600.0 +
                                                  r2 * (-1.0 / 87178291200.0 +
                                                        r2 * (1.0 / 20922789888000.0))))))));
    const int64_t q = static_cast<int64_t>(k) & 3;
    switch (q) {
    case 0:
        return {s, c};
    case 1:
        return {c, -s};
    case 2:
        return {-s, -c};
    default:
        return {-c, s};
    }
}

struct Position {
    double x, y, z;
};

// Surface::Position of (`lonDeg`, `latDeg`) on a sphere of `radius` meters.
TG_HOST_DEVICE inline Position SpherePosition(double lonDeg, double latDeg, double radius) {
    const SinCos lon = SinCosDegrees(lonDeg);
    const SinCos lat = SinCosDegrees(latDeg);
    return {radius * (lat.cos * lon.cos), radius * (lat.cos * lon.sin), radius * lat.sin};
}

// GeoTransform::PixelToWorld of a north-up raster's cell centre, in degrees
// (src/raster/grid.cpp; src/gpu/kernel.cpp:55-70 for the plane form).
TG_HOST_DEVICE inline Position CellPosition(double west, double north, double step, uint32_t col,
                                            uint32_t row, double radius) {
    const double lon = west + (double(col) + 0.5) * step + 0.0;
    const double lat = north + (double(row) + 0.5) * -step;
    return SpherePosition(lon, lat, radius);
}

}  // namespace tg
