// SPDX-License-Identifier: BSD-3-Clause
//
// The reference's default crater models on the host; see lunar_models.h. Every quantity is
// computed with the reference's operations in its order (FP64 with the C library's pow, log10,
// expm1; the diffusion tables' heights from the FP32 kernel law of the fresh profile), so the
// densities, lifetimes, tables and morphometries are the reference's bits.

#include "lunar_models.h"

#include "prepare.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <stdexcept>
#include <string>

namespace tgh {

namespace {

// NeukumProductionFunction (neukum.cpp:29-33): log10 N(>=D) per km^2 as a polynomial in
// log10(D / 1 km), valid from 10 m to 300 km. The published coefficients (Neukum et al. 2001).
constexpr double kNpf[] = {-3.076756, -3.557528, 0.781027,  1.021521,  -0.156012, -0.444058,
                           0.019977,  0.086850,  -0.005874, -0.006809, 8.25e-4,   5.54e-5};
constexpr std::size_t kNpfTerms = sizeof(kNpf) / sizeof(kNpf[0]);
constexpr double kJoin = 10.0;          // ExtendedNeukumProductionFunction's splice (neukum.cpp:36)
constexpr double kSmallSlope = 2.85;    // and its slope below the splice
constexpr double kHartmannSlope = 1.83;  // HartmannEquilibrium (equilibrium.cpp:13-15)

// LogPolynomialSfd::Polynomial and Derivative (log_polynomial.cpp:37-52).
double NpfPolynomial(double x) {
    double sum = 0.0;
    for (std::size_t i = kNpfTerms; i-- > 0;) {
        sum = sum * x + kNpf[i];
    }
    return sum;
}

double NpfDerivative(double x) {
    double sum = 0.0;
    for (std::size_t i = kNpfTerms - 1; i > 0; --i) {
        sum = sum * x + static_cast<double>(i) * kNpf[i];
    }
    return sum;
}

// LogPolynomialSfd::CumulativeAt and IncrementalAt with a 1 km diameter unit and km^2 area unit.
double NpfCumulative(double d) {
    return std::pow(10.0, NpfPolynomial(std::log10(d / 1000.0))) / 1e6;
}

double NpfIncremental(double d) {
    return -NpfCumulative(d) * NpfDerivative(std::log10(d / 1000.0)) / d;
}

void CheckRange(double d, double low, double high) {
    if (!(d > 0.0 && std::isfinite(d))) {
        throw std::domain_error("Diameter must be positive and finite");
    }
    if (!(d >= low && d <= high)) {
        throw std::domain_error("Diameter " + std::to_string(d) +
                                " m is outside the distribution's valid range");
    }
}

// The Neukum chronology's terms (neukum.cpp:17-18), per km^2.
constexpr double kBurst = 5.44e-14, kDecay = 6.93, kLinear = 8.38e-4, kMaxAgeGyr = 4.6;

// The diffusion tables' ages (diffusion.cpp:28-47).
constexpr double kFirstAge = 1e-5;
constexpr double kLastAge = 0.5;
constexpr std::size_t kAgeSteps = 160;
constexpr double kReferenceDiameter = 1000.0;

// The Thomas algorithm over rows 0 to n - 1 with x[n] fixed (diffusion.cpp:59-82).
void SolveTridiagonal(const std::vector<double>& lower, const std::vector<double>& diag,
                      const std::vector<double>& upper, std::vector<double>& x,
                      std::vector<double>& c, std::vector<double>& d) {
    const std::size_t n = diag.size();
    c[0] = upper[0] / diag[0];
    d[0] = x[0] / diag[0];
    for (std::size_t i = 1; i < n; ++i) {
        const double m = diag[i] - lower[i] * c[i - 1];
        c[i] = upper[i] / m;
        d[i] = (x[i] - lower[i] * d[i - 1]) / m;
    }
    for (std::size_t i = n; i-- > 0;) {
        x[i] = d[i] - c[i] * x[i + 1];
    }
}

}  // namespace

LunarSurface::LunarSurface() {
    // ExtendedNeukumProductionFunction (neukum.cpp:35-41): N(>=10 m) of the NPF carried down
    // with slope 2.85 (the NPF's Cumulative at 10 m is within its range).
    m_smallCoefficient = NpfCumulative(kJoin) * std::pow(kJoin, kSmallSlope);
    // ProductionSurface (production_surface.cpp:11-23): N1(3.5 Gyr) / N_production(>=1 km).
    m_scale = NeukumChronology(kSurfaceAgeGyr) / Production(1000.0);
    // PowerLawCoefficientFromKilometers(10^-1.14, 1.83) (power_law.cpp:32-34).
    m_hartmann = std::pow(10.0, -1.14) * std::pow(1000.0, kHartmannSlope) / 1e6;
}

// SplicedSfd::CumulativeAt (spliced.cpp:44-50): the power law up to and including 10 m.
double LunarSurface::Production(double d) const {
    return d > kJoin ? NpfCumulative(d) : m_smallCoefficient * std::pow(d, -kSmallSlope);
}

double LunarSurface::ProductionIncremental(double d) const {
    return d > kJoin ? NpfIncremental(d)
                     : m_smallCoefficient * kSmallSlope * std::pow(d, -kSmallSlope - 1.0);
}

double LunarSurface::Equilibrium(double d) const {
    return m_hartmann * std::pow(d, -kHartmannSlope);
}

double LunarSurface::EquilibriumIncremental(double d) const {
    return m_hartmann * kHartmannSlope * std::pow(d, -kHartmannSlope - 1.0);
}

// MinimumSfd::LowestAt (minimum.cpp:23-35): the first distribution unless another is lower.
double LunarSurface::Density(double d) const {
    const double production = m_scale * Production(d);
    const double equilibrium = Equilibrium(d);
    return equilibrium < production ? equilibrium : production;
}

double LunarSurface::DensityChecked(double d) const {
    CheckRange(d, kMinDiameter, kMaxDiameter);
    return Density(d);
}

// SteadyStateAges::EquilibriumLifetimeGyr (steady_state.cpp:45-47): both incrementals checked
// against their distributions' ranges.
double LunarSurface::EquilibriumLifetimeGyr(double d) const {
    CheckRange(d, 0.0, std::numeric_limits<double>::infinity());
    const double equilibrium = EquilibriumIncremental(d);
    CheckRange(d, kMinDiameter, kMaxDiameter);
    return equilibrium / ProductionIncremental(d);
}

double LunarSurface::LifetimeGyr(double d) const {
    return std::fmin(EquilibriumLifetimeGyr(d), kSurfaceAgeGyr);
}

double NeukumChronology(double ageGyr) {
    if (!(ageGyr >= 0.0 && ageGyr <= kMaxAgeGyr)) {
        throw std::domain_error("Lunar surface age must be between 0 and 4.6 Gyr");
    }
    return (kBurst * std::expm1(kDecay * ageGyr) + kLinear * ageGyr) / 1e6;
}

const std::vector<double>& DiffusionAges() {
    static const std::vector<double> ages = [] {
        std::vector<double> values{0.0};
        const double ratio =
            std::pow(kLastAge / kFirstAge, 1.0 / static_cast<double>(kAgeSteps - 1));
        for (std::size_t k = 0; k < kAgeSteps; ++k) {
            values.push_back(k + 1 == kAgeSteps
                                 ? kLastAge
                                 : kFirstAge * std::pow(ratio, static_cast<double>(k)));
        }
        return values;
    }();
    return ages;
}

AgeRow AgeRowAt(float tau) {
    static const std::vector<float> ages = [] {
        std::vector<float> values;
        for (const double age : DiffusionAges()) {
            values.push_back(static_cast<float>(age));
        }
        return values;
    }();
    AgeRow age{0u, 0.0f};
    const std::uint32_t last = static_cast<std::uint32_t>(ages.size()) - 1u;
    const float first = ages[1];
    if (tau >= ages[last]) {
        age.row = last;
    } else if (tau < first) {
        age.weight = tau / first;
    } else {
        std::uint32_t low = 1u;
        std::uint32_t high = last;
        while (high - low > 1u) {
            const std::uint32_t middle = (low + high) / 2u;
            if (ages[middle] <= tau) {
                low = middle;
            } else {
                high = middle;
            }
        }
        age.row = low;
        const float below = ages[low];
        const float above = ages[high];
        age.weight = std::log(tau / below) / std::log(above / below);
    }
    return age;
}

// DiffusedProfile::Solve (diffusion.cpp:184-212) with RadialDiffusion (:84-124) and
// AppendHeights (:126-144), for a simple crater at the reference diameter.
const DiffusionTables::Table& DiffusionTables::TableFor(const Morphometry2& fresh) {
    const auto key = std::make_pair(fresh.depthRatio, fresh.rimHeightRatio);
    const auto found = m_tables.find(key);
    if (found != m_tables.end()) {
        return *found->second;
    }
    const std::size_t n = kCells;
    const double extent = 0.5 * kReachRadii;
    const double dx = extent / static_cast<double>(n);
    const double diameter = kReferenceDiameter;
    // The fresh profile's kernel law (MartinParkesDunstanProfile::Height) in its fresh state.
    tg_crater_shape mpd{};
    mpd.profile.simple = TG_PROFILE_MPD;
    std::vector<double> h(n + 1, 0.0);
    for (std::size_t i = 0; i < n; ++i) {
        h[i] = ProfileHeight(mpd, static_cast<double>(i) * dx * diameter, diameter,
                             fresh.depthRatio, fresh.rimHeightRatio, fresh.depthRatio,
                             fresh.rimHeightRatio) /
               diameter;
    }
    std::vector<double> below(n + 1, 0.0), above(n + 1, 0.0);
    for (std::size_t i = 1; i < n; ++i) {
        const double x = static_cast<double>(i) * dx;
        below[i] = (x - 0.5 * dx) / (x * dx * dx);
        above[i] = (x + 0.5 * dx) / (x * dx * dx);
    }
    above[0] = 4.0 / (dx * dx);
    std::vector<double> lower(n), diag(n), upper(n), c(n), d(n);
    auto table = std::make_unique<Table>();
    const std::vector<double>& ages = DiffusionAges();
    table->heights.reserve(ages.size() * (n + 1));
    double tau = 0.0;
    for (const double target : ages) {
        while (tau < target) {
            const double dt = std::fmin(target - tau, std::fmax(dx * dx, 0.002 * tau));
            for (std::size_t i = 0; i < n; ++i) {
                lower[i] = -dt * below[i];
                diag[i] = 1.0 + dt * (below[i] + above[i]);
                upper[i] = -dt * above[i];
            }
            h[n] = 0.0;
            SolveTridiagonal(lower, diag, upper, h, c, d);
            tau += dt;
        }
        double rim = -std::numeric_limits<double>::infinity();
        double floor = std::numeric_limits<double>::infinity();
        for (std::size_t i = 0; i < h.size(); ++i) {
            const double x = static_cast<double>(i) * dx;
            if (x >= 0.25) {
                rim = std::fmax(rim, h[i]);
            }
            if (x <= 0.5) {
                floor = std::fmin(floor, h[i]);
            }
            table->heights.push_back(static_cast<float>(h[i]));
        }
        table->morphometry.push_back({rim - floor, rim});
    }
    return *m_tables.emplace(key, std::move(table)).first->second;
}

// DiffusedProfile::Morphometry (diffusion.cpp:250-265).
Morphometry2 DiffusionTables::MorphometryAt(const Morphometry2& fresh, double tau) {
    if (!(tau > 0.0)) {
        return fresh;
    }
    const Table& table = TableFor(fresh);
    const AgeRow age = AgeRowAt(static_cast<float>(tau));
    Morphometry2 result = table.morphometry[age.row];
    if (age.weight > 0.0f) {
        const auto weight = static_cast<double>(age.weight);
        const Morphometry2& next = table.morphometry[age.row + 1];
        result.depthRatio += weight * (next.depthRatio - result.depthRatio);
        result.rimHeightRatio += weight * (next.rimHeightRatio - result.rimHeightRatio);
    }
    return result;
}

// DiffusedProfile::AgeAtDepthRatio (diffusion.cpp:267-285).
double DiffusionTables::AgeAtDepthRatio(const Morphometry2& fresh, double depthRatio) {
    const std::vector<Morphometry2>& rows = TableFor(fresh).morphometry;
    const std::vector<double>& ages = DiffusionAges();
    if (rows.front().depthRatio <= depthRatio) {
        return 0.0;
    }
    for (std::size_t r = 1; r < rows.size(); ++r) {
        const double above = rows[r - 1].depthRatio;
        const double below = rows[r].depthRatio;
        if (below <= depthRatio) {
            const double weight = (above - depthRatio) / (above - below);
            return r == 1 ? weight * ages[1] : ages[r - 1] * std::pow(ages[r] / ages[r - 1], weight);
        }
    }
    return ages.back();
}

}  // namespace tgh
