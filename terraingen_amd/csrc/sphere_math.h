// SPDX-License-Identifier: BSD-3-Clause
//
// Cell positions on a sphere, the same bits on the host and on gfx950. The reference places a
// longitude-latitude raster's cell centre at Surface::Position (src/geometry/surface.cpp:45-50,
// 261-267): R (cos lat cos lon, cos lat sin lon, sin lat) with the C library's FP64 sine and
// cosine. Here the sine and cosine are this file's own: the angle is reduced in degrees, which is
// exact (x - 90 k is representable whenever it is no larger than x), then converted to radians and
// summed as Taylor series on [-pi/4, pi/4] by Horner's rule in a fixed order, so the host and the
// device (both built with FP contraction off, both rounding to nearest) compute identical
// positions. They differ from the C library's by about an ulp, 2e-10 m at a lunar radius.
#pragma once

#include <cmath>
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
    // The nearest quarter turn k (ties to even; any k within a turn of the nearest keeps the
    // series accurate), and the remainder r in degrees, which is exact.
    const double k = rint(degrees * (1.0 / 90.0));
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
    // k mod 4, exactly (every step is exact for an integral k), picks the quadrant by selects.
    const double quadrant = k - 4.0 * floor(k * 0.25);
    const bool odd = quadrant == 1.0 || quadrant == 3.0;
    const double a = odd ? c : s;
    const double b = odd ? s : c;
    const bool negateSin = quadrant >= 2.0;
    const bool negateCos = quadrant == 1.0 || quadrant == 2.0;
    return {negateSin ? -a : a, negateCos ? -b : b};
}

struct Position {
    double x, y, z;
};

// Surface::Position from the sine and cosine of the longitude and of the latitude.
TG_HOST_DEVICE inline Position SpherePositionOf(SinCos lon, SinCos lat, double radius) {
    return {radius * (lat.cos * lon.cos), radius * (lat.cos * lon.sin), radius * lat.sin};
}

// Surface::Position of (`lonDeg`, `latDeg`) on a sphere of `radius` meters.
TG_HOST_DEVICE inline Position SpherePosition(double lonDeg, double latDeg, double radius) {
    return SpherePositionOf(SinCosDegrees(lonDeg), SinCosDegrees(latDeg), radius);
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
