// SPDX-License-Identifier: BSD-3-Clause
//
// ORACLE — test infrastructure only (see noise_oracle.cpp). The reference's own CPU path over one
// tile, built into oracle/_ref/libtgref.so from its sources, for the CPU baseline of the
// benchmark: the noise of a fractal stack (FractalStackRows, src/noise/surface_noise.cpp:22-37)
// row by row on ParallelFor as NoiseLayer's CPU path runs it (src/layers/noise_layer.cpp:28-43),
// then crater shaping tile by tile (16 x 16 cells on ForEachTile, each cell through
// ShapeCratersOnBody or ShapeCraters at its MakeSplitPoint) as ShapeTilesOnCpu runs it
// (src/layers/craters/shaping/apply.cpp:199-222). Inputs are in this repository's layouts, which
// are the reference's (tg_layouts.h), with the crater batch and tile lists the caller prepared.
// Nothing of the reference is copied here: this file calls its functions.

#include "layers/craters/shaping/crater_kernel.h"
#include "terraingen/geometry/surface.h"
#include "terraingen/layers/kernels/craters.h"
#include "terraingen/noise/fractals.h"
#include "terraingen/noise/surface_noise.h"
#include "terraingen/raster/cell_loops.h"
#include "terraingen/raster/raster.h"
#include "terraingen/util/parallel.h"

#include "tg_layouts.h"

#include <cstring>
#include <exception>
#include <string>
#include <vector>

using namespace terraingen;

namespace {

thread_local std::string gBenchError;

Fractal Widen(const tg_fractal& f) {
    Fractal out;
    out.enabled = f.enabled != 0;
    out.seed = f.seed;
    out.wavelength = f.wavelength;
    out.amplitude = f.amplitude;
    out.octaves = f.octaves;
    out.lacunarity = f.lacunarity;
    out.gain = f.gain;
    out.basis = static_cast<NoiseBasis>(f.basis);
    out.type = static_cast<FractalType>(f.type);
    out.offset = f.offset;
    out.blend = static_cast<FractalBlend>(f.blend);
    out.warpAmplitude = f.warp_amplitude;
    out.warpWavelength = f.warp_wavelength;
    return out;
}

// One tile through the reference's CPU path; `records` empty means noise only.
int RunTile(const tg_fractal* fractals, std::size_t fractalCount, double originX, double originY,
            double pixel, std::uint32_t width, std::uint32_t height, double radius,
            const void* kernel, const void* records, const float* data, const void* frames,
            const std::uint32_t* heads, const std::uint32_t* entries, unsigned threads,
            float* heights) {
    try {
        static_assert(sizeof(CraterShapeKernel) == sizeof(tg_crater_shape), "kernel layout");
        static_assert(sizeof(CraterRecord) == sizeof(tg_crater_record), "record layout");
        static_assert(sizeof(CraterFrame) == sizeof(tg_crater_frame), "frame layout");
        const Surface surface = radius > 0.0 ? Surface::Sphere(radius) : Surface::Plane();
        std::vector<Fractal> stack;
        for (std::size_t i = 0; i < fractalCount; ++i) {
            stack.push_back(Widen(fractals[i]));
        }
        const RasterGrid grid{width, height, GeoTransform::NorthUp(originX, originY, pixel)};
        Raster<float> raster(grid, 0.0f);
        const Vec2d origin{0.0, 0.0};  // SyntheticSource's origin; unused on a body
        ParallelFor(height, threads, [&](std::size_t row) {
            std::vector<Vec2d> points(width);
            std::vector<float> noise(width);
            for (std::size_t col = 0; col < width; ++col) {
                points[col] = raster.CellCenter(col, row);
            }
            FractalStackRows(stack, origin, surface, points, noise);
            for (std::size_t col = 0; col < width; ++col) {
                raster(col, row) = raster(col, row) + noise[col];
            }
        });
        if (records != nullptr) {
            CraterShapeKernel k;
            std::memcpy(&k, kernel, sizeof k);
            const auto* recs = static_cast<const CraterRecord*>(records);
            const auto* frs = static_cast<const CraterFrame*>(frames);
            ForEachTile(width, height, 16, threads, [&](const PixelWindow& cells) {
                const std::uint32_t head =
                    CraterTileHead(heads, static_cast<std::uint32_t>(cells.colBegin),
                                   static_cast<std::uint32_t>(cells.rowBegin));
                const std::uint32_t first = heads[head];
                const std::uint32_t count = heads[head + 1];
                if (count == 0) {
                    return;
                }
                for (std::size_t row = cells.rowBegin; row < cells.rowEnd; ++row) {
                    for (std::size_t col = cells.colBegin; col < cells.colEnd; ++col) {
                        float& cell = raster(col, row);
                        const SplitPoint at = MakeSplitPoint(surface, raster.CellCenter(col, row));
                        const CraterData d{data};
                        cell = surface.IsBody() ? ShapeCratersOnBody(k, d, recs, frs, entries,
                                                                     first, count, cell, at)
                                                : ShapeCraters(k, d, recs, entries, first, count,
                                                               cell, at);
                    }
                }
            });
        }
        std::memcpy(heights, raster.Values().data(), raster.Size() * sizeof(float));
        return 0;
    } catch (const std::exception& e) {
        gBenchError = e.what();
        return 1;
    }
}

}  // namespace

extern "C" {

const char* ref_bench_error() {
    return gBenchError.c_str();
}

// A longitude-latitude tile on a sphere of `radius` meters (north-west corner and cell side in
// degrees): noise at each cell's body-fixed position, then craters in their frames.
int ref_body_tile(const tg_fractal* fractals, std::size_t fractalCount, double lon, double lat,
                  double step, std::uint32_t width, std::uint32_t height, double radius,
                  const void* kernel, const void* records, const float* data, const void* frames,
                  const std::uint32_t* heads, const std::uint32_t* entries, unsigned threads,
                  float* heights) {
    return RunTile(fractals, fractalCount, lon, lat, step, width, height, radius, kernel, records,
                   data, frames, heads, entries, threads, heights);
}

// A plane tile (north-west corner and pixel in meters): the synthetic source's noise (origin
// {0, 0}), then craters.
int ref_plane_tile(const tg_fractal* fractals, std::size_t fractalCount, double x, double y,
                   double pixel, std::uint32_t width, std::uint32_t height, const void* kernel,
                   const void* records, const float* data, const std::uint32_t* heads,
                   const std::uint32_t* entries, unsigned threads, float* heights) {
    return RunTile(fractals, fractalCount, x, y, pixel, width, height, 0.0, kernel, records, data,
                   nullptr, heads, entries, threads, heights);
}

}  // extern "C"
