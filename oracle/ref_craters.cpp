// SPDX-License-Identifier: BSD-3-Clause
//
// ORACLE — test infrastructure only (see noise_oracle.cpp). A driver of the reference's own CPU
// pipeline, built into oracle/_ref/libtgref.so from its sources: it sets up a terrain the way a
// caller of the reference's C++ API would (sources, noise, crater layer), then returns what the
// reference computes — generated crater lists, the prepared kernel batch (records and data) and
// the evaluated FP32 heights on its CPU backend. Nothing of the reference is copied here.

#include "layers/craters/shaping/crater_kernel.h"
#include "terraingen/distribution/equilibrium.h"
#include "terraingen/distribution/neukum.h"
#include "terraingen/distribution/power_law.h"
#include "terraingen/distribution/spliced.h"
#include "terraingen/layers/craters/ages/age_model.h"
#include "terraingen/layers/craters/crater_layer.h"
#include "terraingen/layers/craters/degradation/diffusion.h"
#include "terraingen/layers/craters/degradation/freshness_degradation.h"
#include "terraingen/layers/craters/profiles/complex.h"
#include "terraingen/layers/craters/profiles/fassett_thomson.h"
#include "terraingen/layers/craters/profiles/martin_parkes_dunstan.h"
#include "terraingen/layers/craters/shaping/apply.h"
#include "terraingen/layers/craters/shaping/morphometry.h"
#include "terraingen/layers/craters/shaping/prepared.h"
#include "terraingen/layers/kernels/craters.h"
#include "terraingen/layers/noise_layer.h"
#include "terraingen/pipeline/terrain.h"
#include "terraingen/sources/flat_source.h"
#include "terraingen/sources/synthetic_source.h"

#include "tg_layouts.h"

#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

using namespace terraingen;

extern "C" {

// A test case in the reference's vocabulary; tests/ fills it through ctypes.
typedef struct ref_case {
    std::uint32_t width, height;
    double pixel, center_x, center_y;
    float elevation;
    std::uint32_t source_kind;    // 0 FlatSource, 1 SyntheticSource (fractals)
    std::uint32_t fractal_count;  // SyntheticSource fractals
    tg_fractal fractals[8];
    std::uint32_t craters;  // 1: a CraterLayer
    // 0 DefaultCraterModels(); 1 power law + uniform ages + MPD + freshness degradation;
    // 2 as 1 with complex craters (MPD simple) from transition_diameter; 3 as 1 with FT profile
    std::uint32_t models;
    double sfd_coefficient, sfd_exponent, min_d, max_d, lifetime, transition_diameter;
    std::uint32_t irregular, peaks, terraces;
    double terrace_onset, terrace_full;
    std::uint32_t seed;
    std::uint32_t threads;
    // Optional shaping features (RimSmoothing, CraterNoise, CraterTexture; 0 leaves them off).
    double rim_sigma, noise_amplitude, noise_rim_amplitude, wall_streaks, ejecta;
} ref_case;

}  // extern "C"

namespace {

Fractal FromLayout(const tg_fractal& f) {
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

CraterModels ModelsFor(const ref_case& c) {
    if (c.models == 0) {
        return DefaultCraterModels();
    }
    auto density = std::make_shared<PowerLawSfd>(c.sfd_coefficient, c.sfd_exponent);
    std::shared_ptr<const CraterAgeModel> ages = std::make_shared<UniformAges>(c.lifetime);
    auto degradation = std::make_shared<FreshnessDegradation>(ages, LunarFreshCraterMorphometry);
    std::shared_ptr<const CraterProfile> profile;
    if (c.models == 3) {
        profile = std::make_shared<FassettThomsonProfile>();
    } else {
        profile = std::make_shared<MartinParkesDunstanProfile>();
    }
    if (c.models == 2) {
        ComplexCraterShape shape = LunarComplexCraterShape();
        shape.transitionDiameter = c.transition_diameter;
        profile = std::make_shared<ComplexCraterProfile>(profile, shape);
    }
    return {density, ages, degradation, profile};
}

struct Case {
    ref_case spec;
    RasterGrid grid;
    Terrain terrain;
    std::shared_ptr<const CraterLayer> craters;
    // Filled on demand.
    CraterList list;
    CraterShapeKernel kernel{};
    std::vector<CraterRecord> records;
    std::vector<float> data;
    std::string error;
};

RasterGrid GridOf(const ref_case& c) {
    const double halfX = 0.5 * double(c.width) * c.pixel;
    const double halfY = 0.5 * double(c.height) * c.pixel;
    return {c.width, c.height,
            GeoTransform::NorthUp(c.center_x - halfX, c.center_y + halfY, c.pixel)};
}

thread_local std::string gError;

}  // namespace

extern "C" {

const char* ref_last_error() {
    return gError.c_str();
}

void* ref_case_open(const ref_case* spec) {
    try {
        auto out = std::make_unique<Case>();
        out->spec = *spec;
        out->grid = GridOf(*spec);
        std::shared_ptr<const Source> source;
        if (spec->source_kind == 1) {
            std::vector<Fractal> fractals;
            for (std::uint32_t i = 0; i < spec->fractal_count; ++i) {
                fractals.push_back(FromLayout(spec->fractals[i]));
            }
            source = std::make_shared<SyntheticSource>(fractals, spec->elevation);
        } else {
            source = std::make_shared<FlatSource>(spec->width, spec->height, spec->pixel,
                                                  spec->elevation,
                                                  Vec2d{spec->center_x, spec->center_y});
        }
        std::vector<LayerPtr> layers;
        if (spec->craters) {
            CraterLayerOptions options;
            options.seed = spec->seed;
            if (spec->models != 0) {
                options.generation.minDiameter = spec->min_d;
                options.generation.maxDiameter = spec->max_d;
            }
            options.apply.irregularity.enabled = spec->irregular != 0;
            options.apply.peaks.enabled = spec->peaks != 0;
            options.apply.terraces.enabled = spec->terraces != 0;
            if (spec->terraces != 0 && spec->terrace_full > spec->terrace_onset) {
                options.apply.terraces.onsetDiameter = spec->terrace_onset;
                options.apply.terraces.fullDiameter = spec->terrace_full;
            }
            options.apply.rimSmoothing.sigmaRadii = spec->rim_sigma;
            options.apply.noise.amplitude = spec->noise_amplitude;
            options.apply.noise.rimAmplitude = spec->noise_rim_amplitude;
            options.apply.texture.wallStreaks = spec->wall_streaks;
            options.apply.texture.ejecta = spec->ejecta;
            auto layer = std::make_shared<CraterLayer>(ModelsFor(*spec), options);
            out->craters = layer;
            layers.push_back(layer);
        }
        out->terrain = Terrain(Surface{}, source, layers);
        return out.release();
    } catch (const std::exception& e) {
        gError = e.what();
        return nullptr;
    }
}

void ref_case_close(void* handle) {
    delete static_cast<Case*>(handle);
}

// The reference's CPU evaluation of the whole terrain over the case's grid.
int ref_case_evaluate(void* handle, float* heights) {
    Case& c = *static_cast<Case*>(handle);
    try {
        EvaluationOptions options;
        options.On(Backend::Cpu).Threads(c.spec.threads);
        const Raster<float> raster = c.terrain.Evaluate(c.grid, options);
        std::memcpy(heights, raster.Values().data(), raster.Size() * sizeof(float));
        return 0;
    } catch (const std::exception& e) {
        gError = e.what();
        return 1;
    }
}

// The craters the layer applies to the grid (CraterLayer::ApplySamples, crater_layer.cpp:337-361),
// in generation order; returns the count, and fills the columns when `x` is not null.
long ref_case_craters(void* handle, double* x, double* y, double* d, double* age,
                      std::uint64_t* id) {
    Case& c = *static_cast<Case*>(handle);
    try {
        if (!c.craters) {
            return 0;
        }
        const double spacing = c.spec.pixel;
        const Bounds2d bounds = c.grid.Bounds();
        c.list = c.craters->Craters(Surface{}.Expand(bounds, c.craters->GroundReach(spacing)),
                                    spacing, 0.0, std::numeric_limits<double>::infinity(),
                                    c.spec.threads);
        if (x != nullptr) {
            for (std::size_t n = 0; n < c.list.Size(); ++n) {
                const Crater k = c.list.At(n);
                x[n] = k.x;
                y[n] = k.y;
                d[n] = k.diameter;
                age[n] = k.ageGyr;
                id[n] = k.id;
            }
        }
        return long(c.list.Size());
    } catch (const std::exception& e) {
        gError = e.what();
        return -1;
    }
}

// The kernel struct, records and data the reference builds for every prepared crater of the list
// (PrepareCraters then MakeCraterKernelBatch over all of them, in apply order). Call
// ref_case_craters first. Returns the record count; `data_count` receives the data length.
long ref_case_batch(void* handle, void* kernel188, long* dataCount) {
    Case& c = *static_cast<Case*>(handle);
    try {
        const CraterLayerOptions& options = c.craters->Options();
        const CraterShapeModels models = c.craters->Models().Shape();
        const std::vector<PreparedCrater> prepared =
            PrepareCraters(c.list, options.apply, c.spec.pixel, models, c.spec.threads, {}, {});
        c.kernel = MakeCraterShapeKernel(*models.profile, options.apply);
        std::vector<std::size_t> all(prepared.size());
        for (std::size_t n = 0; n < all.size(); ++n) {
            all[n] = n;
        }
        CraterKernelBatch batch =
            MakeCraterKernelBatch(c.kernel, *models.profile, prepared, all, options.apply);
        c.records = std::move(batch.records);
        c.data = std::move(batch.data);
        std::memcpy(kernel188, &c.kernel, sizeof(CraterShapeKernel));
        *dataCount = long(c.data.size());
        return long(c.records.size());
    } catch (const std::exception& e) {
        gError = e.what();
        return -1;
    }
}

void ref_case_batch_copy(void* handle, void* records, float* data) {
    Case& c = *static_cast<Case*>(handle);
    std::memcpy(records, c.records.data(), c.records.size() * sizeof(CraterRecord));
    std::memcpy(data, c.data.data(), c.data.size() * sizeof(float));
}

// One (cell, crater) pair through the reference's ApplyCraterKernel (craters.h:1186-1212).
float ref_apply_crater(const void* kernel188, const float* data, const void* record144,
                       float height, float dx, float dy) {
    CraterShapeKernel k;
    CraterRecord r;
    std::memcpy(&k, kernel188, sizeof k);
    std::memcpy(&r, record144, sizeof r);
    return ApplyCraterKernel(k, CraterData{data}, r, height, dx, dy);
}

}  // extern "C"
