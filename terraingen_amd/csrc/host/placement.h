// SPDX-License-Identifier: BSD-3-Clause
//
// Host placement of craters and boulder candidates (placement.cpp).
#pragma once

#include "cube_cells.h"
#include "lunar_models.h"
#include "random.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <functional>
#include <stdexcept>
#include <vector>

namespace tgh {

// A cumulative crater density: the expected count per area in a diameter interval, and a
// diameter drawn from it (SizeFrequencyDistribution::ExpectedCount and SampleDiameter,
// size_frequency.cpp:116-140).
class Density {
  public:
    virtual ~Density() = default;
    virtual double ExpectedCount(double area, double minD, double maxD) const = 0;
    virtual double SampleDiameter(Stream& random, double minD, double maxD) const = 0;
};

// N(>=D) = coefficient D^-exponent per m^2 (PowerLawSegment, power_law.h:15-33).
struct PowerLawDensity final : Density {
    double coefficient = 1.0;
    double exponent = 2.0;

    PowerLawDensity() = default;
    PowerLawDensity(double c, double e) : coefficient(c), exponent(e) {}

    double Cumulative(double d) const {
        if (!(d > 0.0 && std::isfinite(d))) {
            throw std::domain_error("Diameter must be positive and finite");
        }
        return coefficient * std::pow(d, -exponent);
    }
    double Invert(double target) const {
        return std::pow(target / coefficient, -1.0 / exponent);
    }
    // PowerLawSegment::Incremental (power_law.h:26-28): -dN/dD = c e D^(-e-1).
    double Incremental(double d) const {
        return coefficient * exponent * std::pow(d, -exponent - 1.0);
    }
    double ExpectedCount(double area, double minD, double maxD) const override;
    double SampleDiameter(Stream& random, double minD, double maxD) const override;
};

// A weighted sum of power laws, N(>=D) = sum of w c D^-e (SummedSfd, src/distribution/
// summed.cpp:35-49, over PowerLawSfd terms), sampled as any distribution without a closed-form
// inverse is (SizeFrequencyDistribution::SampleDiameter and InvertCumulative,
// size_frequency.cpp:95-140): the boulder layer's candidate sizes.
class SummedPowerLaws final : public Density {
  public:
    struct Term {
        double weight;
        PowerLawDensity law;
    };
    explicit SummedPowerLaws(std::vector<Term> terms) : m_terms(std::move(terms)) {}
    double Cumulative(double d) const;
    // IncrementalAt: the sum of the terms' incremental densities, -dN/dD.
    double Incremental(double d) const;
    double ExpectedCount(double area, double minD, double maxD) const override;
    double SampleDiameter(Stream& random, double minD, double maxD) const override;

  private:
    std::vector<Term> m_terms;
};

// SizeFrequencyDistribution::InvertCumulative (size_frequency.cpp:95-113): the diameter in
// [minD, maxD] where `cumulative` falls to `target`, by Illinois steps on log N over log D
// (InvertBySecant, :34-61), else by bisection.
double InvertCumulative(const std::function<double(double)>& cumulative, double target,
                        double minD, double maxD);

// The density of the reference's default crater models (lunar_models.h).
class LunarDensity final : public Density {
  public:
    double ExpectedCount(double area, double minD, double maxD) const override;
    double SampleDiameter(Stream& random, double minD, double maxD) const override;
    const LunarSurface& Surface() const {
        return m_surface;
    }

  private:
    // SizeFrequencyDistribution::InvertCumulative (size_frequency.cpp:95-113).
    double Invert(double target, double minD, double maxD) const;
    LunarSurface m_surface;
};

// Crater ages (CraterAgeModel, age_model.cpp:9-51): uniform over a lifetime in Gyr (UniformAges),
// all fresh at 0 (FreshAges, a lifetime of 0), or the lunar steady state (SteadyStateAges of the
// default models, when `surface` is set).
struct Ages {
    double lifetimeGyr = 0.0;
    const LunarSurface* surface = nullptr;

    // CraterAgeModel::LifetimeGyr.
    double Lifetime(double diameter) const;
    // CraterAgeModel::SampleAgeGyr.
    double Sample(Stream& random, double diameter) const;
    // CraterAgeModel::RemainingLifetime.
    double Remaining(double diameter, double ageGyr) const {
        const double lifetime = Lifetime(diameter);
        if (lifetime == 0.0) {
            return 1.0;
        }
        return std::clamp(1.0 - ageGyr / lifetime, 0.0, 1.0);
    }
};

struct Band {
    int exponent;
    double low, high, cellSize;
};

std::vector<Band> CraterBands(double minD, double maxD, double cellSizeFactor);

struct GenerationOptions {
    double minDiameter = 0.0;
    double maxDiameter = 0.0;
    double marginRadii = 0.0;
    double cellSizeFactor = 16.0;
    double minResolvedDiameter = 0.0;
    double maxExpectedCraters = 5e7;
    unsigned threads = 1;  // cells drawn on this many threads (0: every hardware thread)
};

struct Crater {
    double x, y, diameter, ageGyr;
    std::uint64_t id;
};

std::vector<Crater> GenerateCraters(double minX, double minY, double maxX, double maxY,
                                    const Density& density, const Ages& ages,
                                    std::uint32_t seed, const GenerationOptions& options);

// The same on a sphere of `radius` meters, over longitude-latitude regions (degrees), in
// cube-sphere cells (GenerateCraters with CubeSphereCells, population.cpp:414-424): for one
// region the reference's list; for several, the union of theirs, each cell drawn once.
std::vector<Crater> GenerateCratersOnSphere(double radius, const std::vector<Box>& regions,
                                            const Density& density, const Ages& ages,
                                            std::uint32_t seed, const GenerationOptions& options);

}  // namespace tgh
