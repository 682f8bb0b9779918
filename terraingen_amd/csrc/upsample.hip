// SPDX-License-Identifier: BSD-3-Clause
//
// Kernel upsampling of a post lattice onto a grid on the device: the reference's KernelUpsampler
// (src/sources/upsampling.cpp:318-335) and the post-spacing rule of Upsampler::Fill (:297-312),
// which the reference runs on the CPU for every source with posts (a SyntheticSource with a pixel
// size, GeoTiffSource, RasterSource). One lane per grid cell: the cell centre in FP64
// (GeoTransform::PixelToWorld), its lattice position (GeoTransform::WorldToPixel less half a
// post), the kernel's weights along each axis in FP64 (KernelWeights, src/raster/resample.cpp:
// 112-135) and their FP64 sum over the taps (InterpolateLatticeWeights, src/raster/kernels.h:
// 31-51), narrowed to FP32 once; NaN where a tap of non-zero weight has no data. The taps are read
// from the posts held in device memory, neighbouring lanes sharing them through the caches.
// Built with -ffp-contract=off, as the reference is, so every product and sum rounds as its does.

#include "internal.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

using namespace tgi;

namespace {

// Keys' cubic convolution kernel with a = -0.5 (Catmull-Rom) at an offset of x cells.
__device__ __forceinline__ double CubicAt(double x) {
    constexpr double a = -0.5;
    x = fabs(x);
    if (x <= 1.0) {
        return ((a + 2.0) * x - (a + 3.0)) * x * x + 1.0;
    }
    if (x < 2.0) {
        return ((a * x - 5.0 * a) * x + 8.0 * a) * x - 4.0 * a;
    }
    return 0.0;
}

// The uniform cubic B-spline at an offset of x cells.
__device__ __forceinline__ double BSplineAt(double x) {
    x = fabs(x);
    if (x < 1.0) {
        return 2.0 / 3.0 - x * x + 0.5 * x * x * x;
    }
    if (x < 2.0) {
        const double rest = 2.0 - x;
        return rest * rest * rest / 6.0;
    }
    return 0.0;
}

// sin(pi x) / (pi x), 1 at 0.
__device__ __forceinline__ double SincAt(double x) {
    if (x == 0.0) {
        return 1.0;
    }
    const double angle = 3.141592653589793238462643383279502884 * x;
    return sin(angle) / angle;
}

// Three-lobed Lanczos at an offset of x cells, exactly 0 at the whole offsets but 0.
__device__ __forceinline__ double LanczosAt(double x) {
    x = fabs(x);
    if (x >= 3.0 || (x > 0.0 && x == floor(x))) {
        return 0.0;
    }
    return SincAt(x) * SincAt(x / 3.0);
}

__host__ __device__ constexpr int RadiusOf(uint32_t method) {
    return method == TG_RESAMPLE_LANCZOS3 ? 3
           : (method == TG_RESAMPLE_CUBIC || method == TG_RESAMPLE_CUBIC_BSPLINE) ? 2
                                                                                  : 1;
}

// The weights of the taps -radius + 1 .. radius around a point `fraction` of the way from tap 0
// to tap 1, in that order (Lanczos normalized to sum to 1).
__device__ __forceinline__ void AxisWeights(uint32_t method, double fraction, double w[6]) {
    for (int k = 0; k < 6; ++k) {
        w[k] = 0.0;
    }
    if (method == TG_RESAMPLE_BILINEAR) {
        w[0] = 1.0 - fraction;
        w[1] = fraction;
        return;
    }
    const int radius = RadiusOf(method);
    double total = 0.0;
    for (int i = 1 - radius; i <= radius; ++i) {
        const double at = fraction - i;
        const double v = method == TG_RESAMPLE_LANCZOS3       ? LanczosAt(at)
                         : method == TG_RESAMPLE_CUBIC_BSPLINE ? BSplineAt(at)
                                                               : CubicAt(at);
        w[i + radius - 1] = v;
        total += v;
    }
    if (method == TG_RESAMPLE_LANCZOS3) {
        for (int k = 0; k < 6; ++k) {
            w[k] /= total;
        }
    }
}

// The column in the window of lattice column `col`: wrapped first when the lattice wraps, then
// clamped to the window, so reads beyond it repeat its edge.
__device__ __forceinline__ uint32_t WindowColumn(const tg_upsample& u, int64_t col) {
    if (u.wrap_columns != 0u) {
        const int64_t n = int64_t(u.lattice_columns);
        col = ((col % n) + n) % n;
    }
    return uint32_t(col < u.col0 ? 0 : (col >= u.col0 + int64_t(u.window_width)
                                            ? int64_t(u.window_width) - 1
                                            : col - u.col0));
}

// The offset in the window of the first post of lattice row `row`, clamped to the window.
__device__ __forceinline__ size_t WindowRow(const tg_upsample& u, int64_t row) {
    const int64_t r = row < u.row0 ? 0 : (row >= u.row0 + int64_t(u.window_height)
                                              ? int64_t(u.window_height) - 1
                                              : row - u.row0);
    return size_t(r) * u.window_width;
}

// The post at lattice (col, row) from the window held.
__device__ __forceinline__ float PostAt(const tg_upsample& u, const float* __restrict__ posts,
                                        int64_t col, int64_t row) {
    return posts[WindowRow(u, row) + WindowColumn(u, col)];
}

// `v` as a whole number when it lies within 1e-6 of one (detail::NearestWhole, lattice.h:33-39).
__device__ __forceinline__ bool NearWhole(double v, int64_t& whole) {
    const double rounded = round(v);
    if (!(fabs(v - rounded) <= 1e-6)) {
        return false;
    }
    whole = int64_t(rounded);
    return true;
}

constexpr uint32_t kSide = 16;

// One variant per kernel (kMethod), so the taps' loops unroll and their indices and weights stay
// in registers. Each tap's window index is found once per axis rather than once per tap.
template <uint32_t kMethod>
__global__ void __launch_bounds__(kSide * kSide)
    UpsampleGrid(tg_upsample u, const float* __restrict__ posts, float* __restrict__ out) {
    const uint32_t col = blockIdx.x * kSide + (threadIdx.x % kSide);
    const uint32_t row = blockIdx.y * kSide + (threadIdx.x / kSide);
    if (col >= u.width || row >= u.height) {
        return;
    }
    // The cell centre (Raster::CellCenter of a north-up grid) and its position on the lattice
    // (GeoTransform::WorldToPixel of the lattice's north-up transform, less half a post).
    const double zero = 0.0;
    const double x = u.origin_x + (double(col) + 0.5) * u.pixel + (double(row) + 0.5) * zero;
    const double y = u.origin_y + (double(col) + 0.5) * zero + (double(row) + 0.5) * -u.pixel;
    const double pw = u.post, ph = -(u.post_height > 0.0 ? u.post_height : u.post);
    const double det = pw * ph - zero * zero;
    const double dx = x - u.lattice_x;
    const double dy = y - u.lattice_y;
    const double px = (dx * ph - dy * zero) / det;
    const double py = (dy * pw - dx * zero) / det;
    const double lu = px - 0.5;
    const double lv = py - 0.5;
    float* cell = out + size_t(row) * (u.out_pitch != 0u ? u.out_pitch : u.width) + col;
    if (u.covered != 0u) {
        // WorldToPixel of the centre against the covered cells.
        if (!(px >= double(u.cover_col0) && py >= double(u.cover_row0) &&
              px < double(u.cover_col1) && py < double(u.cover_row1))) {
            *cell = __builtin_nanf("");
            return;
        }
    }
    if (u.at_post_spacing != 0u) {
        int64_t pc, pr;
        if (NearWhole(lu, pc) && NearWhole(lv, pr)) {
            *cell = PostAt(u, posts, pc, pr);  // a cell centred on a post keeps the post
            return;
        }
    }
    if constexpr (kMethod == TG_RESAMPLE_NEAREST) {
        *cell = PostAt(u, posts, int64_t(round(lu)), int64_t(round(lv)));
    } else {
        constexpr int kRadius = RadiusOf(kMethod);
        constexpr int kTaps = 2 * kRadius;
        const double u0 = floor(lu);
        const double v0 = floor(lv);
        double wx[6], wy[6];
        AxisWeights(kMethod, lu - u0, wx);
        AxisWeights(kMethod, lv - v0, wy);
        // Tap k of an axis lies k + 1 - radius posts from (u0, v0).
        uint32_t cols[kTaps];
        size_t rows[kTaps];
#pragma unroll
        for (int k = 0; k < kTaps; ++k) {
            cols[k] = WindowColumn(u, int64_t(u0) + (k + 1 - kRadius));
            rows[k] = WindowRow(u, int64_t(v0) + (k + 1 - kRadius));
        }
        double sum = 0.0;
#pragma unroll
        for (int j = 0; j < kTaps; ++j) {
            const double ww = wy[j];
#pragma unroll
            for (int i = 0; i < kTaps; ++i) {
                const double w = wx[i];
                if (w * ww == 0.0) {
                    continue;  // a tap of no weight is not read
                }
                const float p = posts[rows[j] + cols[i]];
                if (isnan(p)) {
                    *cell = __builtin_nanf("");
                    return;
                }
                sum += w * ww * double(p);
            }
        }
        *cell = float(sum);
    }
}

}  // namespace

namespace tgi {

void ValidateUpsample(const tg_upsample& u) {
    if (!(u.post > 0.0) || !std::isfinite(u.post) || !(u.post_height >= 0.0) ||
        !std::isfinite(u.post_height) || !(u.pixel > 0.0) || !std::isfinite(u.pixel) ||
        !std::isfinite(u.lattice_x) || !std::isfinite(u.lattice_y) ||
        !std::isfinite(u.origin_x) || !std::isfinite(u.origin_y)) {
        throw HipError{"upsampling: the lattice and grid need finite origins and positive sizes"};
    }
    if (u.window_width == 0 || u.window_height == 0) {
        throw HipError{"upsampling: the posts window is empty"};
    }
    if (u.out_pitch != 0u && u.out_pitch < u.width) {
        throw HipError{"upsampling: the output's row pitch is below the grid's width"};
    }
    if (u.method > TG_RESAMPLE_CUBIC_BSPLINE) {
        throw HipError{"upsampling: unknown resample method"};
    }
    if (u.wrap_columns != 0u &&
        (u.lattice_columns == 0 || u.col0 != 0 || u.window_width != u.lattice_columns)) {
        throw HipError{"upsampling: wrapping posts must span every column of the lattice"};
    }
}

void LaunchUpsample(const tg_upsample& u, const float* posts, float* out, hipStream_t stream) {
    if (u.width == 0 || u.height == 0) {
        return;
    }
    const dim3 grid((u.width + kSide - 1) / kSide, (u.height + kSide - 1) / kSide, 1);
    using Kernel = decltype(&UpsampleGrid<TG_RESAMPLE_CUBIC>);
    const Kernel kernels[] = {UpsampleGrid<TG_RESAMPLE_NEAREST>, UpsampleGrid<TG_RESAMPLE_BILINEAR>,
                              UpsampleGrid<TG_RESAMPLE_CUBIC>, UpsampleGrid<TG_RESAMPLE_LANCZOS3>,
                              UpsampleGrid<TG_RESAMPLE_CUBIC_BSPLINE>};
    hipLaunchKernelGGL(kernels[u.method], grid, dim3(kSide * kSide), 0, stream, u, posts, out);
    Check(hipGetLastError(), "launching the upsampling kernel");
}

}  // namespace tgi

extern "C" {

int tg_hip_upsample(tg_dev* dev, const tg_upsample* spec, const float* posts, float* heights,
                    void* stream, char* err, size_t errLen) {
    if (dev == nullptr || spec == nullptr || posts == nullptr ||
        (size_t(spec->width) * spec->height > 0 && heights == nullptr)) {
        SetErr(err, errLen, "tg_hip_upsample: null argument");
        return TG_ERR_ARGUMENT;
    }
    return Guard(err, errLen, [&] {
        ValidateUpsample(*spec);
        Check(hipSetDevice(dev->ordinal), "hipSetDevice");
        LaunchUpsample(*spec, posts, heights, static_cast<hipStream_t>(stream));
    });
}

}  // extern "C"
