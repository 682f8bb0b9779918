// SPDX-License-Identifier: BSD-3-Clause
//
// The reference's default crater models (DefaultCraterModels, crater_layer.cpp:194-203) on the
// host (lunar_models.cpp):
//   density  the lunar equilibrium surface: the smaller of the extended Neukum production function
//            over 3.5 Gyr of the Neukum chronology and Hartmann's equilibrium power law
//            (steady_state.cpp:61-65, production_surface.cpp, spliced.cpp, neukum.cpp,
//            log_polynomial.cpp, equilibrium.cpp, minimum.cpp);
//   ages     steady-state ages: each crater's lifetime is the equilibrium over the production
//            rate, capped at 3.5 Gyr, and its age a formation time drawn from the chronology
//            (steady_state.cpp:45-59);
//   profile  the Martin-Parkes-Dunstan profile diffused in time: tables of radial heights solved
//            by implicit diffusion steps, one per fresh morphometry, at 161 diffusion ages
//            (diffusion.cpp:20-212);
//   wear     matched diffusion: each crater diffused so that it reaches the visibility depth
//            ratio at the end of its lifetime (diffusivities.cpp:28-63, diffusion.cpp:288-318).
#pragma once

#include "random.h"

#include <cstdint>
#include <map>
#include <memory>
#include <utility>
#include <vector>

namespace tgh {

// EquilibriumSurface{SplicedSfd(ExtendedNeukumProductionFunction()),
// PowerLawSfd(HartmannEquilibrium()), 3.5, NeukumLunarChronology}: cumulative densities per m^2
// without range checks (the reference's CumulativeAt), and the checked forms it exposes.
class LunarSurface {
  public:
    LunarSurface();

    // The production rate's cumulative and incremental (per m^2 per Gyr-normalized N1 unit).
    double Production(double d) const;
    double ProductionIncremental(double d) const;
    // Hartmann's equilibrium density.
    double Equilibrium(double d) const;
    double EquilibriumIncremental(double d) const;
    // MinimumSfd of the production surface and the equilibrium.
    double Density(double d) const;
    // The density's Cumulative with its range check: [2.5 m, 300 km].
    double DensityChecked(double d) const;
    // SteadyStateAges::EquilibriumLifetimeGyr and LifetimeAt (capped at the surface age).
    double EquilibriumLifetimeGyr(double d) const;
    double LifetimeGyr(double d) const;

    static constexpr double kMinDiameter = 2.5;
    static constexpr double kMaxDiameter = 300000.0;
    static constexpr double kSurfaceAgeGyr = 3.5;

  private:
    double m_smallCoefficient = 0.0;  // the power law below 10 m, joined to the NPF there
    double m_scale = 0.0;             // N1(3.5 Gyr) over the production's N(>=1 km)
    double m_hartmann = 0.0;          // Hartmann's coefficient in meters
};

// NeukumLunarChronology (neukum.cpp:21-23): N1 per m^2 after `ageGyr`, within [0, 4.6] Gyr.
double NeukumChronology(double ageGyr);

// The diffusion ages of the tables, 0 then 160 geometric ages from 1e-5 to 0.5 (diffusion.cpp:33-47).
const std::vector<double>& DiffusionAges();

struct Morphometry2 {
    double depthRatio, rimHeightRatio;
};

// DiffusedProfile over the fresh Martin-Parkes-Dunstan profile with the default reach (3 radii)
// and cells (600): each table solved once per fresh morphometry and kept.
class DiffusionTables {
  public:
    static constexpr double kReachRadii = 3.0;
    static constexpr std::uint32_t kCells = 600;

    struct Table {
        std::vector<float> heights;             // one row of kCells + 1 per age
        std::vector<Morphometry2> morphometry;  // per age: rim - floor, rim (over D)
    };

    // The table of craters with fresh morphometry `fresh` (simple craters: one table per
    // morphometry, solved at the 1 km reference diameter).
    const Table& TableFor(const Morphometry2& fresh);
    // DiffusedProfile::Morphometry at diffusion age `tau` (KernelAgeRowAt in FP32).
    Morphometry2 MorphometryAt(const Morphometry2& fresh, double tau);
    // DiffusedProfile::AgeAtDepthRatio.
    double AgeAtDepthRatio(const Morphometry2& fresh, double depthRatio);

  private:
    std::map<std::pair<double, double>, std::unique_ptr<Table>> m_tables;
};

// KernelAgeRowAt (craters.h:536-561) over the ages in FP32.
struct AgeRow {
    std::uint32_t row;
    float weight;
};
AgeRow AgeRowAt(float tau);

}  // namespace tgh
