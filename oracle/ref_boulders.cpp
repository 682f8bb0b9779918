// SPDX-License-Identifier: BSD-3-Clause
//
// ORACLE — test infrastructure only (see noise_oracle.cpp). A driver of the reference's own CPU
// pipeline for boulders, built into oracle/_ref/libtgref.so from its sources: a flat or fractal
// source under a BoulderLayer, set up as a caller of the reference's C++ API would, returning the
// boulders the layer draws into a grid and the FP32 heights its CPU backend evaluates. Nothing of
// the reference is copied here.

#include "terraingen/distribution/power_law.h"
#include "terraingen/layers/boulders/boulder_layer.h"
#include "terraingen/layers/layer.h"
#include "terraingen/pipeline/terrain.h"
#include "terraingen/raster/grid.h"
#include "terraingen/sources/flat_source.h"
#include "terraingen/sources/synthetic_source.h"

#include "tg_layouts.h"

#include <cstring>
#include <memory>
#include <string>
#include <vector>

using namespace terraingen;

extern "C" {

typedef struct ref_boulder_case {
    std::uint32_t width, height;
    double pixel, center_x, center_y;
    float elevation;
    std::uint32_t source_kind;    // 0 FlatSource, 1 SyntheticSource (fractals)
    std::uint32_t fractal_count;
    tg_fractal fractals[8];
    std::uint32_t sizes_kind;     // 0 the layer's default sizes, 1 a power law
    double sfd_coefficient, sfd_exponent, min_d, max_d;
    double ventifact_share;       // BoulderVentifacts::share
    std::uint32_t seed;
    std::uint32_t threads;
} ref_boulder_case;

// A boulder layer on a sphere of `radius` meters over a longitude-latitude raster of `side` x
// `side` cells of `step` degrees from its north-west corner (lon, lat), flat at 0 beneath.
typedef struct ref_boulder_sphere_case {
    std::uint32_t side;
    double lon, lat, step, radius;
    double sfd_coefficient, sfd_exponent, min_d, max_d;
    double ventifact_share;
    std::uint32_t seed;
    std::uint32_t threads;
} ref_boulder_sphere_case;

// One boulder as the relief reads it (boulder_relief.cpp:244-275): centre, diameter, exposed
// height, model flag and up to two facets (azimuth, angle, reach).
typedef struct ref_boulder {
    double x, y, diameter, exposed;
    double facets[6];
    std::uint32_t facet_count;
    std::uint32_t model;
} ref_boulder;

}  // extern "C"

namespace {

Fractal FromLayoutB(const tg_fractal& f) {
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

struct BoulderCase {
    ref_boulder_case spec;
    RasterGrid grid;
    Terrain terrain;
    std::shared_ptr<const BoulderLayer> layer;
};

thread_local std::string gBoulderError;

}  // namespace

extern "C" {

const char* ref_boulders_error() {
    return gBoulderError.c_str();
}

void* ref_boulders_open(const ref_boulder_case* spec) {
    try {
        auto out = std::make_unique<BoulderCase>();
        out->spec = *spec;
        const double halfX = 0.5 * double(spec->width) * spec->pixel;
        const double halfY = 0.5 * double(spec->height) * spec->pixel;
        out->grid = RasterGrid{spec->width, spec->height,
                               GeoTransform::NorthUp(spec->center_x - halfX,
                                                     spec->center_y + halfY, spec->pixel)};
        std::shared_ptr<const Source> source;
        if (spec->source_kind == 1) {
            std::vector<Fractal> fractals;
            for (std::uint32_t i = 0; i < spec->fractal_count; ++i) {
                fractals.push_back(FromLayoutB(spec->fractals[i]));
            }
            source = std::make_shared<SyntheticSource>(fractals, spec->elevation);
        } else {
            source = std::make_shared<FlatSource>(spec->width, spec->height, spec->pixel,
                                                  spec->elevation,
                                                  Vec2d{spec->center_x, spec->center_y});
        }
        BoulderLayerOptions options;
        options.seed = spec->seed;
        if (spec->sizes_kind == 1) {
            options.sizes.distribution =
                std::make_shared<PowerLawSfd>(spec->sfd_coefficient, spec->sfd_exponent);
            options.sizes.minDiameter = spec->min_d;
            options.sizes.maxDiameter = spec->max_d;
        }
        options.ventifacts.share = spec->ventifact_share;
        auto layer = std::make_shared<BoulderLayer>(options);
        out->layer = layer;
        out->terrain = Terrain(Surface{}, source, std::vector<LayerPtr>{layer});
        return out.release();
    } catch (const std::exception& e) {
        gBoulderError = e.what();
        return nullptr;
    }
}

void ref_boulders_close(void* handle) {
    delete static_cast<BoulderCase*>(handle);
}

int ref_boulders_evaluate(void* handle, float* heights) {
    BoulderCase& c = *static_cast<BoulderCase*>(handle);
    try {
        EvaluationOptions options;
        options.On(Backend::Cpu).Threads(c.spec.threads);
        const Raster<float> raster = c.terrain.Evaluate(c.grid, options);
        std::memcpy(heights, raster.Values().data(), raster.Size() * sizeof(float));
        return 0;
    } catch (const std::exception& e) {
        gBoulderError = e.what();
        return 1;
    }
}

namespace {

void Export(const BoulderList& list, ref_boulder* out) {
    for (std::size_t n = 0; n < list.Size(); ++n) {
        const Boulder b = list.At(n);
        ref_boulder& r = out[n];
        r.x = b.x;
        r.y = b.y;
        r.diameter = b.diameter;
        r.exposed = b.ExposedHeight();
        for (int k = 0; k < 2; ++k) {
            r.facets[3 * k] = b.facets[k].azimuth;
            r.facets[3 * k + 1] = b.facets[k].angle;
            r.facets[3 * k + 2] = b.facets[k].reach;
        }
        r.facet_count = b.facetCount;
        r.model = b.model ? 1u : 0u;
    }
}

}  // namespace

// The layer of a sphere case applied on the CPU to its raster of zeros: the FP32 heights into
// `heights` (side x side, row by row) and, when `out` is not null, the boulders it draws there
// (those of the raster's box grown by half a sample, at the raster's spacing, in the layer's
// order; at most `capacity`). Returns the boulder count, or -1 on error.
long ref_boulders_sphere(const ref_boulder_sphere_case* spec, float* heights, ref_boulder* out,
                         long capacity) {
    try {
        const Surface surface = Surface::Sphere(spec->radius);
        BoulderLayerOptions options;
        options.seed = spec->seed;
        options.sizes.distribution =
            std::make_shared<PowerLawSfd>(spec->sfd_coefficient, spec->sfd_exponent);
        options.sizes.minDiameter = spec->min_d;
        options.sizes.maxDiameter = spec->max_d;
        options.ventifacts.share = spec->ventifact_share;
        const BoulderLayer layer(options, nullptr, TerrainBase{surface});
        const RasterGrid grid{spec->side, spec->side,
                              GeoTransform::NorthUp(spec->lon, spec->lat, spec->step)};
        Raster<float> raster(grid, 0.0f);
        LayerContext context;
        context.threads = spec->threads;
        layer.Apply(raster, context);
        std::memcpy(heights, raster.Values().data(), raster.Size() * sizeof(float));
        const double spacing = ResolvedSpacing(grid, 0.0, surface);
        const BoulderList list = layer.Boulders(surface.Expand(grid.Bounds(), 0.5 * spacing),
                                                spacing);
        if (out != nullptr) {
            if (long(list.Size()) > capacity) {
                gBoulderError = "the boulder list is longer than the capacity given";
                return -1;
            }
            Export(list, out);
        }
        return long(list.Size());
    } catch (const std::exception& e) {
        gBoulderError = e.what();
        return -1;
    }
}

// The boulders the layer draws into the case's grid: those of the grid's box grown by half a
// sample, as BoulderLayer::ApplySamples gathers them for dome relief (boulder_relief.cpp:287-291,
// 429-432), in the layer's order. Returns the count; fills `out` when it is not null.
long ref_boulders_list(void* handle, ref_boulder* out) {
    BoulderCase& c = *static_cast<BoulderCase*>(handle);
    try {
        const double spacing = c.spec.pixel;
        const BoulderList list =
            c.layer->Boulders(Surface{}.Expand(c.grid.Bounds(), 0.5 * spacing), spacing);
        if (out != nullptr) {
            Export(list, out);
        }
        return long(list.Size());
    } catch (const std::exception& e) {
        gBoulderError = e.what();
        return -1;
    }
}

}  // extern "C"
