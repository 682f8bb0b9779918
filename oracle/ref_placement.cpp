// SPDX-License-Identifier: BSD-3-Clause
//
// ORACLE — test infrastructure only (see noise_oracle.cpp). A driver of the reference's placement,
// built into oracle/_ref/libtgref.so from its sources: crater generation on a sphere
// (GenerateCraters, src/layers/craters/generation/population.cpp:414-424, with CubeSphereCells, a
// power-law density and uniform ages, over a longitude-latitude region), and the boulders a
// BoulderLayer of power-law sizes lists over a region for a spacing (BoulderLayer::Boulders,
// src/layers/boulders/boulder_layer.cpp:390-400). They return the lists the reference draws, for
// fixtures and for live comparison. Nothing of the reference is copied here.

#include "terraingen/distribution/power_law.h"
#include "terraingen/layers/boulders/boulder_layer.h"
#include "terraingen/geometry/cells.h"
#include "terraingen/geometry/surface.h"
#include "terraingen/layers/craters/ages/age_model.h"
#include "terraingen/layers/craters/generation/population.h"

#include <cstdint>
#include <exception>
#include <string>

using namespace terraingen;

namespace {

thread_local std::string gPlacementError;

}  // namespace

extern "C" {

typedef struct ref_sphere_generation {
    double radius;                              // meters
    double min_lon, min_lat, max_lon, max_lat;  // degrees
    double sfd_coefficient, sfd_exponent;       // N(>=D) = c D^-e per m^2
    double min_d, max_d, margin_radii, min_resolved, lifetime;
    std::uint32_t seed, threads;
} ref_sphere_generation;

const char* ref_placement_error() {
    return gPlacementError.c_str();
}

// The reference's list: its size, and with `capacity` at least that, its columns.
long ref_generate_sphere(const ref_sphere_generation* g, double* x, double* y, double* d,
                         double* age, std::uint64_t* id, long capacity) {
    try {
        const CubeSphereCells cells(g->radius);
        const PowerLawSfd density(g->sfd_coefficient, g->sfd_exponent);
        const UniformAges ages(g->lifetime);
        CraterGenerationOptions options;
        options.minDiameter = g->min_d;
        options.maxDiameter = g->max_d;
        options.marginRadii = g->margin_radii;
        options.minResolvedDiameter = g->min_resolved;
        options.threads = g->threads;
        options.surface = Surface::Sphere(g->radius);  // regions are tested on the sphere
        const CraterList list = GenerateCraters(
            cells, {g->min_lon, g->min_lat, g->max_lon, g->max_lat}, density, ages, g->seed,
            options);
        const long n = static_cast<long>(list.Size());
        if (capacity >= n) {
            for (long k = 0; k < n; ++k) {
                x[k] = list.X()[k];
                y[k] = list.Y()[k];
                d[k] = list.Diameters()[k];
                age[k] = list.AgesGyr()[k];
                id[k] = list.Ids()[k];
            }
        }
        return n;
    } catch (const std::exception& e) {
        gPlacementError = e.what();
        return -1;
    }
}

typedef struct ref_boulder_layer {
    double min_x, min_y, max_x, max_y;  // meters, or degrees on a sphere
    double sphere_radius;               // 0: a plane
    double spacing;
    double sfd_coefficient, sfd_exponent, min_d, max_d;
    double min_height_ratio, max_height_ratio, min_burial, max_burial, max_lean_degrees;
    std::uint32_t seed, threads;
} ref_boulder_layer;

// The boulders' x, y, diameter, height, burial, yaw, lean, wear (eight doubles each), IDs and model
// flags, with `capacity` at least their count; returns the count.
long ref_place_boulders(const ref_boulder_layer* g, double* values, std::uint64_t* ids,
                        std::uint8_t* models, long capacity) {
    try {
        BoulderLayerOptions options;
        options.seed = g->seed;
        options.sizes.distribution = std::make_shared<PowerLawSfd>(g->sfd_coefficient,
                                                                   g->sfd_exponent);
        options.sizes.minDiameter = g->min_d;
        options.sizes.maxDiameter = g->max_d;
        options.pose.minHeightRatio = g->min_height_ratio;
        options.pose.maxHeightRatio = g->max_height_ratio;
        options.pose.minBurial = g->min_burial;
        options.pose.maxBurial = g->max_burial;
        options.pose.maxLeanDegrees = g->max_lean_degrees;
        const Surface surface =
            g->sphere_radius > 0.0 ? Surface::Sphere(g->sphere_radius) : Surface::Plane();
        const BoulderLayer layer(options, std::shared_ptr<const CraterLayer>{},
                                 TerrainBase(surface));
        const BoulderList list = layer.Boulders({g->min_x, g->min_y, g->max_x, g->max_y},
                                                g->spacing, nullptr);
        const long n = static_cast<long>(list.Size());
        if (capacity >= n) {
            for (long k = 0; k < n; ++k) {
                const Boulder b = list.At(static_cast<std::size_t>(k));
                double* v = values + 8 * k;
                v[0] = b.x;
                v[1] = b.y;
                v[2] = b.diameter;
                v[3] = b.height;
                v[4] = b.burial;
                v[5] = b.yaw;
                v[6] = b.lean;
                v[7] = b.wear;
                ids[k] = b.id;
                models[k] = b.model ? 1 : 0;
            }
        }
        return n;
    } catch (const std::exception& e) {
        gPlacementError = e.what();
        return -1;
    }
}

}  // extern "C"
