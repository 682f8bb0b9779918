// SPDX-License-Identifier: BSD-3-Clause
//
// ORACLE — test infrastructure only (see noise_oracle.cpp). A driver of the reference's own CPU
// pipeline for sources with posts, built into oracle/_ref/libtgref.so from its sources: a terrain
// whose only layer is its source (a SyntheticSource with a pixel size, or a GeoTiffSource of a
// file), upsampled with a KernelUpsampler, evaluated on a north-up grid as a caller of the
// reference's C++ API would. Nothing of the reference is copied here.

#include "terraingen/noise/fractals.h"
#include "terraingen/pipeline/terrain.h"
#include "terraingen/raster/geotiff/geotiff.h"
#include "terraingen/sources/geotiff_source.h"
#include "terraingen/sources/synthetic_source.h"
#include "terraingen/sources/upsampling.h"

#include "tg_layouts.h"

#include <cstring>
#include <memory>
#include <string>
#include <vector>

using namespace terraingen;

namespace {

thread_local std::string gSourceError;
unsigned gThreads = 8;

Fractal FromLayoutS(const tg_fractal& f) {
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

int Evaluate(std::shared_ptr<const Source> source, std::uint32_t width, std::uint32_t height,
             double originX, double originY, double pixel, double spacing, float* heights,
             double radius = 0.0) {
    const Terrain terrain(radius > 0.0 ? Surface::Sphere(radius) : Surface{}, std::move(source),
                          std::vector<LayerPtr>{});
    const RasterGrid grid{width, height, GeoTransform::NorthUp(originX, originY, pixel)};
    EvaluationOptions options;
    options.On(Backend::Cpu).Threads(gThreads);
    options.spacing = spacing;
    const Raster<float> raster = terrain.Evaluate(grid, options);
    std::memcpy(heights, raster.Values().data(), raster.Size() * sizeof(float));
    return 0;
}

}  // namespace

extern "C" {

const char* ref_sources_error() {
    return gSourceError.c_str();
}

// The CPU threads the grids below are evaluated on (8 until set; 0 leaves it).
void ref_sources_threads(std::uint32_t threads) {
    if (threads != 0) {
        gThreads = threads;
    }
}

// A SyntheticSource of `count` fractals at `elevation` whose posts lie `post` apart, upsampled
// by KernelUpsampler(`method`), evaluated on the grid of width x height cells of `pixel` from its
// north-west corner (originX, originY) at the sample spacing `spacing` (0: the grid's).
int ref_synthetic_grid(const tg_fractal* fractals, std::uint32_t count, float elevation,
                       double post, std::uint32_t method, std::uint32_t width,
                       std::uint32_t height, double originX, double originY, double pixel,
                       double spacing, float* heights) {
    try {
        std::vector<Fractal> stack;
        for (std::uint32_t i = 0; i < count; ++i) {
            stack.push_back(FromLayoutS(fractals[i]));
        }
        auto upsampler = std::make_shared<KernelUpsampler>(static_cast<ResampleMethod>(method));
        auto source = std::make_shared<SyntheticSource>(std::move(stack), elevation, post,
                                                        std::move(upsampler));
        return Evaluate(source, width, height, originX, originY, pixel, spacing, heights);
    } catch (const std::exception& e) {
        gSourceError = e.what();
        return 1;
    }
}

// A GeoTiffSource of the file at `path` (band 0), upsampled by KernelUpsampler(`method`), with an
// optional crop (crop[4] = minX, minY, maxX, maxY; null for the whole file), evaluated on the grid
// as ref_synthetic_grid evaluates: on a plane, or on a sphere of `radius` meters when it is
// positive (a file in longitude and latitude).
int ref_geotiff_grid(const char* path, std::uint32_t method, const double* crop,
                     std::uint32_t width, std::uint32_t height, double originX, double originY,
                     double pixel, double spacing, double radius, float* heights) {
    try {
        GeoTiffSourceOptions options;
        options.upsampler =
            std::make_shared<KernelUpsampler>(static_cast<ResampleMethod>(method));
        if (crop != nullptr) {
            options.crop = Bounds2d{crop[0], crop[1], crop[2], crop[3]};
        }
        auto source = std::make_shared<GeoTiffSource>(path, options);
        return Evaluate(source, width, height, originX, originY, pixel, spacing, heights, radius);
    } catch (const std::exception& e) {
        gSourceError = e.what();
        return 1;
    }
}

// The reference's ReadGeoTiff of band `band`: the raw samples as floats (band scaling not applied)
// into `values` (width x height, when not null), the transform (GDAL order) and the no-data
// value. Returns the cell count, or -1 on error.
long ref_geotiff_read(const char* path, std::uint32_t band, float* values, double* transform,
                      double* noData, int* hasNoData) {
    try {
        const GeoTiff tiff = ReadGeoTiff(path, band);
        const Raster<float>& r = tiff.raster;
        if (values != nullptr) {
            std::memcpy(values, r.Values().data(), r.Size() * sizeof(float));
        }
        if (transform != nullptr) {
            const GeoTransform& t = r.Transform();
            const double g[6] = {t.originX, t.pixelWidth, t.rowRotation,
                                 t.originY, t.colRotation, t.pixelHeight};
            std::memcpy(transform, g, sizeof g);
        }
        if (hasNoData != nullptr) {
            *hasNoData = r.NoData().has_value() ? 1 : 0;
            if (noData != nullptr && r.NoData()) {
                *noData = *r.NoData();
            }
        }
        return long(r.Size());
    } catch (const std::exception& e) {
        gSourceError = e.what();
        return -1;
    }
}

}  // extern "C"
