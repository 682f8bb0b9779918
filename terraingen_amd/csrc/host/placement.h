// SPDX-License-Identifier: BSD-3-Clause
//
// Host placement of craters and boulder candidates (placement.cpp).
#pragma once

#include "random.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <vector>

namespace tgh {

// N(>=D) = coefficient D^-exponent per m^2 (PowerLawSegment, power_law.h:15-33).
struct PowerLawDensity {
    double coefficient = 1.0;
    double exponent = 2.0;

    double Cumulative(double d) const {
        if (!(d > 0.0 && std::isfinite(d))) {
            throw std::domain_error("Diameter must be positive and finite");
        }
        return coefficient * std::pow(d, -exponent);
    }
    double Invert(double target) const {
        return std::pow(target / coefficient, -1.0 / exponent);
    }
    double ExpectedCount(double area, double minD, double maxD) const;
    double SampleDiameter(Stream& random, double minD, double maxD) const;
};

// Crater ages: uniform over a lifetime in Gyr (UniformAges), or all fresh at 0 (FreshAges).
struct Ages {
    double lifetimeGyr = 0.0;
    double Sample(Stream& random) const;
    // CraterAgeModel::RemainingLifetime (src/layers/craters/ages/age_model.cpp).
    double Remaining(double ageGyr) const {
        if (lifetimeGyr == 0.0) {
            return 1.0;
        }
        return std::clamp(1.0 - ageGyr / lifetimeGyr, 0.0, 1.0);
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
};

struct Crater {
    double x, y, diameter, ageGyr;
    std::uint64_t id;
};

std::vector<Crater> GenerateCraters(double minX, double minY, double maxX, double maxY,
                                    const PowerLawDensity& density, const Ages& ages,
                                    std::uint32_t seed, const GenerationOptions& options);

}  // namespace tgh
