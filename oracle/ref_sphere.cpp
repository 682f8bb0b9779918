// SPDX-License-Identifier: BSD-3-Clause
//
// ORACLE — test infrastructure only (see noise_oracle.cpp). A driver of the reference's own CPU
// pipeline for craters on a sphere, built into oracle/_ref/libtgref.so from its sources: a
// CraterLayer on Surface::Sphere applied to a longitude-latitude raster, as the reference's
// tests/unit/layers/craters/sphere_test.cpp sets one up. It returns what the layer's CPU backend
// shapes, and what the crater kernel reads on a body: the prepared batch (records, data), each
// crater's frame (MakeCraterFrame) and each cell's split position (MakeSplitPoint). Nothing of the
// reference is copied here.

#include "layers/craters/shaping/crater_kernel.h"
#include "terraingen/distribution/power_law.h"
#include "terraingen/layers/craters/ages/age_model.h"
#include "terraingen/layers/craters/crater_layer.h"
#include "terraingen/layers/craters/degradation/degradation_model.h"
#include "terraingen/layers/craters/degradation/freshness_degradation.h"
#include "terraingen/layers/craters/profiles/fassett_thomson.h"
#include "terraingen/layers/craters/profiles/martin_parkes_dunstan.h"
#include "terraingen/layers/craters/shaping/morphometry.h"
#include "terraingen/layers/craters/shaping/prepared.h"
#include "terraingen/layers/kernels/craters.h"

#include "tg_layouts.h"

#include <cstring>
#include <limits>
#include <memory>
#include <string>
#include <vector>

using namespace terraingen;

extern "C" {

typedef struct ref_sphere_case {
    std::uint32_t side;       // raster cells per side
    double lon, lat, step;    // north-west corner and cell side, degrees
    double radius;            // sphere radius, meters
    // 0: Fassett-Thomson, no degradation (sphere_test.cpp's models);
    // 1: Martin-Parkes-Dunstan with FreshnessDegradation over UniformAges(lifetime)
    std::uint32_t models;
    double sfd_coefficient, sfd_exponent, min_d, max_d, lifetime;
    std::uint32_t irregular;  // CraterIrregularity enabled
    double elongated_share;
    std::uint32_t seed, threads;
    std::uint32_t manual_count;  // extra craters: lon, lat, diameter, age
    double manual[4][4];
} ref_sphere_case;

}  // extern "C"

namespace {

CraterModels SphereModels(const ref_sphere_case& c) {
    auto density = std::make_shared<PowerLawSfd>(c.sfd_coefficient, c.sfd_exponent);
    std::shared_ptr<const CraterAgeModel> ages = std::make_shared<UniformAges>(c.lifetime);
    if (c.models == 1) {
        return {density, ages,
                std::make_shared<FreshnessDegradation>(ages, LunarFreshCraterMorphometry),
                std::make_shared<MartinParkesDunstanProfile>()};
    }
    return {density, ages, std::make_shared<NoDegradation>(LunarFreshCraterMorphometry),
            std::make_shared<FassettThomsonProfile>()};
}

struct SphereCase {
    ref_sphere_case spec;
    CraterShapeKernel kernel{};
    std::vector<CraterRecord> records;
    std::vector<float> data;
    std::vector<CraterFrame> frames;
    std::vector<SplitPoint> points;
    std::vector<float> heights;
};

thread_local std::string gSphereError;

}  // namespace

extern "C" {

const char* ref_sphere_error() {
    return gSphereError.c_str();
}

// Builds the case: applies the layer to a raster of zeros on the CPU, then prepares every crater
// the layer listed for that raster (CraterLayer::ApplySamples's list, in its order) into the
// kernel's batch, with frames and the raster's split positions.
void* ref_sphere_open(const ref_sphere_case* spec) {
    try {
        auto out = std::make_unique<SphereCase>();
        out->spec = *spec;
        const Surface surface = Surface::Sphere(spec->radius);
        CraterLayerOptions options;
        options.seed = spec->seed;
        options.generation.minDiameter = spec->min_d;
        options.generation.maxDiameter = spec->max_d;
        options.apply.irregularity.enabled = spec->irregular != 0;
        options.apply.irregularity.elongatedShare = spec->elongated_share;
        for (std::uint32_t k = 0; k < spec->manual_count && k < 4; ++k) {
            const double* m = spec->manual[k];
            options.generation.manual.Append({m[0], m[1], m[2], m[3], 100u + k});
        }
        const CraterLayer layer(SphereModels(*spec), options, surface);
        const RasterGrid grid{spec->side, spec->side,
                              GeoTransform::NorthUp(spec->lon, spec->lat, spec->step)};
        Raster<float> raster(grid, 0.0f);
        LayerContext context;
        context.threads = spec->threads;
        layer.Apply(raster, context);
        out->heights.assign(raster.Values().begin(), raster.Values().end());

        const double spacing = ResolvedSpacing(grid, 0.0, surface);
        const CraterList list =
            layer.Craters(surface.Expand(grid.Bounds(), layer.GroundReach(spacing)), spacing, 0.0,
                          std::numeric_limits<double>::infinity(), spec->threads);
        const CraterShapeModels models = layer.Models().Shape();
        const std::vector<PreparedCrater> prepared =
            PrepareCraters(list, layer.Options().apply, spacing, models, spec->threads, {}, {});
        out->kernel = MakeCraterShapeKernel(*models.profile, layer.Options().apply);
        std::vector<std::size_t> all(prepared.size());
        for (std::size_t n = 0; n < all.size(); ++n) {
            all[n] = n;
            out->frames.push_back(MakeCraterFrame(surface, prepared[n].crater));
        }
        CraterKernelBatch batch = MakeCraterKernelBatch(out->kernel, *models.profile, prepared,
                                                        all, layer.Options().apply);
        out->records = std::move(batch.records);
        out->data = std::move(batch.data);
        out->points.resize(raster.Size());
        for (std::size_t row = 0; row < grid.height; ++row) {
            for (std::size_t col = 0; col < grid.width; ++col) {
                out->points[row * grid.width + col] =
                    MakeSplitPoint(surface, raster.CellCenter(col, row));
            }
        }
        return out.release();
    } catch (const std::exception& e) {
        gSphereError = e.what();
        return nullptr;
    }
}

void ref_sphere_close(void* handle) {
    delete static_cast<SphereCase*>(handle);
}

// The record count; `data_count` receives the data length.
long ref_sphere_sizes(void* handle, long* dataCount) {
    const SphereCase& c = *static_cast<SphereCase*>(handle);
    *dataCount = long(c.data.size());
    return long(c.records.size());
}

// Copies the kernel struct (188 B), records (144 B each), data, frames (52 B each), the split
// positions (24 B per cell, row-major) and the layer's FP32 heights.
void ref_sphere_copy(void* handle, void* kernel, void* records, float* data, void* frames,
                     void* points, float* heights) {
    const SphereCase& c = *static_cast<SphereCase*>(handle);
    static_assert(sizeof(CraterFrame) == sizeof(tg_crater_frame), "frame layout");
    static_assert(sizeof(SplitPoint) == sizeof(tg_split_point), "split point layout");
    std::memcpy(kernel, &c.kernel, sizeof c.kernel);
    std::memcpy(records, c.records.data(), c.records.size() * sizeof(CraterRecord));
    std::memcpy(data, c.data.data(), c.data.size() * sizeof(float));
    std::memcpy(frames, c.frames.data(), c.frames.size() * sizeof(CraterFrame));
    std::memcpy(points, c.points.data(), c.points.size() * sizeof(SplitPoint));
    std::memcpy(heights, c.heights.data(), c.heights.size() * sizeof(float));
}

}  // extern "C"
