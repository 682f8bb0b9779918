// SPDX-License-Identifier: BSD-3-Clause
//
// ORACLE — test infrastructure only (see noise_oracle.cpp). A CPU restatement of the reference's
// kernel upsampling of a post lattice onto a north-up grid, in the reference's FP64 arithmetic:
//   CubicWeight, BSplineWeight, Sinc, Lanczos3Weight  src/raster/resample.cpp:19-61
//   KernelRadius, KernelWeights                      src/raster/resample.cpp:87-135
//   InterpolateLatticeWeights, InterpolateLattice    src/raster/kernels.h:31-78
//   RasterPosts::Read / detail::PostIndex            src/sources/upsampling.cpp:267-283,
//                                                    src/sources/lattice.h:126-134
//   CenteredPost, AtPostSpacing, Upsampler::Fill     src/sources/upsampling.cpp:31-40,289-312
//   RasterHeights::Covers                            src/sources/raster_source.cpp:78-82
//   GeoTransform::PixelToWorld / WorldToPixel        include/terraingen/raster/geo_transform.h:
//                                                    63-79
// One cell at a time, each step as written there.

#include "tg_layouts.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <limits>
#include <optional>

namespace {

constexpr double kPi = 3.141592653589793238462643383279502884;

double Cubic(double x) {
    constexpr double a = -0.5;
    x = std::fabs(x);
    if (x <= 1.0) {
        return ((a + 2.0) * x - (a + 3.0)) * x * x + 1.0;
    }
    if (x < 2.0) {
        return ((a * x - 5.0 * a) * x + 8.0 * a) * x - 4.0 * a;
    }
    return 0.0;
}

double Sinc(double x) {
    if (x == 0.0) {
        return 1.0;
    }
    const double angle = kPi * x;
    return std::sin(angle) / angle;
}

double Lanczos3(double x) {
    x = std::fabs(x);
    if (x >= 3.0 || (x > 0.0 && x == std::floor(x))) {
        return 0.0;
    }
    return Sinc(x) * Sinc(x / 3.0);
}

double BSpline(double x) {
    x = std::fabs(x);
    if (x < 1.0) {
        return 2.0 / 3.0 - x * x + 0.5 * x * x * x;
    }
    if (x < 2.0) {
        const double rest = 2.0 - x;
        return rest * rest * rest / 6.0;
    }
    return 0.0;
}

int Radius(std::uint32_t method) {
    switch (method) {
    case TG_RESAMPLE_NEAREST:
    case TG_RESAMPLE_BILINEAR:
        return 1;
    case TG_RESAMPLE_LANCZOS3:
        return 3;
    default:
        return 2;
    }
}

std::array<double, 6> Weights(std::uint32_t method, double fraction) {
    std::array<double, 6> w{};
    if (method == TG_RESAMPLE_NEAREST) {
        return w;
    }
    if (method == TG_RESAMPLE_BILINEAR) {
        w[0] = 1.0 - fraction;
        w[1] = fraction;
        return w;
    }
    const int radius = Radius(method);
    double total = 0.0;
    for (int i = 1 - radius; i <= radius; ++i) {
        const double x = fraction - i;
        const double v = method == TG_RESAMPLE_LANCZOS3       ? Lanczos3(x)
                         : method == TG_RESAMPLE_CUBIC_BSPLINE ? BSpline(x)
                                                               : Cubic(x);
        w[static_cast<std::size_t>(i + radius - 1)] = v;
        total += v;
    }
    if (method == TG_RESAMPLE_LANCZOS3) {
        for (double& v : w) {
            v /= total;
        }
    }
    return w;
}

// PostIndex: wrapped into [0, count) when wrapping, then clamped to [begin, end).
std::int64_t Index(std::int64_t at, bool wrap, std::int64_t count, std::int64_t begin,
                   std::int64_t end) {
    const std::int64_t wrapped = wrap ? ((at % count) + count) % count : at;
    return std::clamp<std::int64_t>(wrapped, begin, end - 1);
}

float Post(const tg_upsample& u, const float* posts, std::int64_t col, std::int64_t row) {
    const std::int64_t c = Index(col, u.wrap_columns != 0, u.lattice_columns, u.col0,
                                 u.col0 + static_cast<std::int64_t>(u.window_width));
    const std::int64_t r = Index(row, false, 1, u.row0,
                                 u.row0 + static_cast<std::int64_t>(u.window_height));
    return posts[static_cast<std::size_t>(r - u.row0) * u.window_width +
                 static_cast<std::size_t>(c - u.col0)];
}

std::optional<std::int64_t> Whole(double v) {
    const double rounded = std::round(v);
    if (!(std::fabs(v - rounded) <= 1e-6)) {
        return std::nullopt;
    }
    return static_cast<std::int64_t>(rounded);
}

float Cell(const tg_upsample& u, const float* posts, std::size_t col, std::size_t row) {
    // Raster::CellCenter of the north-up grid.
    const double pw = u.pixel, ph = -u.pixel, rot = 0.0;
    const double x = u.origin_x + (static_cast<double>(col) + 0.5) * pw +
                     (static_cast<double>(row) + 0.5) * rot;
    const double y = u.origin_y + (static_cast<double>(col) + 0.5) * rot +
                     (static_cast<double>(row) + 0.5) * ph;
    // WorldToPixel of the lattice's north-up transform.
    const double lw = u.post, lh = -(u.post_height > 0.0 ? u.post_height : u.post);
    const double det = lw * lh - rot * rot;
    const double dx = x - u.lattice_x;
    const double dy = y - u.lattice_y;
    const double px = (dx * lh - dy * rot) / det;
    const double py = (dy * lw - dx * rot) / det;
    const double lu = px - 0.5, lv = py - 0.5;
    constexpr float kNaN = std::numeric_limits<float>::quiet_NaN();
    if (u.covered != 0 &&
        !(px >= static_cast<double>(u.cover_col0) && py >= static_cast<double>(u.cover_row0) &&
          px < static_cast<double>(u.cover_col1) && py < static_cast<double>(u.cover_row1))) {
        return kNaN;  // RasterHeights::Covers
    }
    if (u.at_post_spacing != 0) {
        const auto c = Whole(lu);
        const auto r = Whole(lv);
        if (c && r) {
            return Post(u, posts, *c, *r);
        }
    }
    if (u.method == TG_RESAMPLE_NEAREST) {
        const float v = Post(u, posts, static_cast<std::int64_t>(std::round(lu)),
                             static_cast<std::int64_t>(std::round(lv)));
        return std::isnan(v) ? kNaN : v;
    }
    const double u0 = std::floor(lu), v0 = std::floor(lv);
    const std::array<double, 6> wxs = Weights(u.method, lu - u0);
    const std::array<double, 6> wys = Weights(u.method, lv - v0);
    const int radius = Radius(u.method);
    double sum = 0.0;
    for (int j = 1 - radius; j <= radius; ++j) {
        const double wy = wys[static_cast<std::size_t>(j + radius - 1)];
        for (int i = 1 - radius; i <= radius; ++i) {
            const double wx = wxs[static_cast<std::size_t>(i + radius - 1)];
            if (wx * wy == 0.0) {
                continue;
            }
            const float v = Post(u, posts, static_cast<std::int64_t>(u0 + i),
                                 static_cast<std::int64_t>(v0 + j));
            if (std::isnan(v)) {
                return kNaN;
            }
            sum += wx * wy * static_cast<double>(v);
        }
    }
    return static_cast<float>(sum);
}

}  // namespace

extern "C" {

// The grid `spec` describes, filled from `posts` (the window it names) into `out`, row by row.
void or_upsample_grid(const tg_upsample* spec, const float* posts, float* out) {
    const std::size_t pitch = spec->out_pitch != 0 ? spec->out_pitch : spec->width;
    for (std::size_t row = 0; row < spec->height; ++row) {
        for (std::size_t col = 0; col < spec->width; ++col) {
            out[row * pitch + col] = Cell(*spec, posts, col, row);
        }
    }
}

// KernelWeights for one fraction, for closed-form checks.
void or_kernel_weights(std::uint32_t method, double fraction, double* out) {
    const std::array<double, 6> w = Weights(method, fraction);
    std::copy(w.begin(), w.end(), out);
}

}  // extern "C"
