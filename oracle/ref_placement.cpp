// SPDX-License-Identifier: BSD-3-Clause
//
// ORACLE — test infrastructure only (see noise_oracle.cpp). A driver of the reference's crater
// generation on a sphere, built into oracle/_ref/libtgref.so from its sources: GenerateCraters
// (src/layers/craters/generation/population.cpp:414-424) with CubeSphereCells on a sphere, a
// power-law density and uniform ages, over a longitude-latitude region. It returns the list the
// reference draws, for fixtures and for live comparison. Nothing of the reference is copied here.

#include "terraingen/distribution/power_law.h"
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

}  // extern "C"
