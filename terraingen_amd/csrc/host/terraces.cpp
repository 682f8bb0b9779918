// SPDX-License-Identifier: BSD-3-Clause
//
// Wall terraces on the host: the reference's AppendKernelTerraces
// (src/layers/craters/shaping/crater_kernel.cpp:312-459) for the closed-form profiles this host
// prepares. A terraced crater's data block is its header (inner wall edge, slope share, strength,
// scarp rounding, arc cut, arc width), five arc harmonics, the rim scallop terms, then per inner
// scarp its base and scallop, wobble and block terms, then per step its bench level, scarp share
// and kept share, every value drawn and derived as the reference does so the block is byte-equal.
//
// The wall is sampled through CraterProfile::Height, which for these profiles is the FP32 kernel
// law evaluated on the host with float arguments and the C++ float overloads of pow and sqrt
// (martin_parkes_dunstan.cpp:9-19, fassett_thomson.cpp:14-18, complex.cpp:69-77 over
// include/terraingen/layers/kernels/craters.h:306-507): the host functions below restate those
// laws in the same FP32 operation order.

#include "prepare.h"
#include "random.h"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace tgh {

namespace {

constexpr double kPi = 3.141592653589793238462643383279502884;
constexpr double kDegree = kPi / 180.0;

// The constants of the terrace preparation (crater_kernel.cpp:26-60).
constexpr double kFullTerraceCover = 0.9;
constexpr double kTerraceArcWidth = 0.3;
constexpr int kTerraceWallSamples = 256;
constexpr double kNarrowestTerraceStep = 0.25;
constexpr double kTerraceScarpRounding = 0.4;
constexpr double kMostTerraceScarpShare = 0.9;
constexpr double kSteepTerraceStep = 0.7;
constexpr double kTerraceScarpSqueeze = 0.75;
constexpr int kTerraceMeanSamples = 64;
constexpr int kTerraceLevelIterations = 30;
constexpr double kTerraceLevelMargin = 0.1;
constexpr double kTerraceScarpRoom = 0.4;
constexpr double kTerraceScarpReach = 1.2;
constexpr double kMinTerraceScallops = 6.0;
constexpr double kMaxTerraceScallops = 12.0;
constexpr double kTerraceRimScallop = 0.5;

// The record fields the profile laws read (KernelStateRecord, profile.h:34-43).
struct ShapeState {
    float freshDepth, freshRim, currentDepth, currentRim;
};

// KernelEjectaScale (craters.h:309-311).
float EjectaScale(float radius) {
    return 0.14f * std::pow(radius, 0.74f);
}

struct Point {
    float height, slope;
};

// MartinParkesDunstanExterior (craters.h:376-387).
Point MpdExterior(float x, float rim, float ejectaScale) {
    const float offset = rim - ejectaScale;
    const float gradient = -ejectaScale / 8.0f - offset;
    const float r = 1.0f / (x + 1.0f);
    const float law = ejectaScale * r * r * r + offset + gradient * x;
    const float lawSlope = -3.0f * ejectaScale * r * r * r * r + gradient;
    const float taper = 1.0f - x * x;
    return {law * taper, lawSlope * taper - 2.0f * x * law};
}

// MartinParkesDunstanHeight (craters.h:389-419).
float MpdHeight(float distance, float diameter, const ShapeState& c) {
    const float depth = c.currentDepth * diameter;
    const float rim = c.currentRim * diameter;
    const float freshRim = c.freshRim * diameter;
    const float radius = 0.5f * diameter;
    const float x = distance / radius - 1.0f;
    if (x >= 1.0f) {
        return 0.0f;
    }
    const float k = freshRim - rim + depth;
    const float alpha = (rim - freshRim) / k;
    if (x <= alpha) {
        return (k * x + 2.0f * k) * x + freshRim;
    }
    if (x <= 0.0f) {
        return k * (alpha + 1.0f) / alpha * x * x + rim;
    }
    const float ejectaScale =
        freshRim > 0.0f ? EjectaScale(radius) * rim / freshRim : EjectaScale(radius);
    const float beta = std::min(1.0f, 1.5f * (freshRim - rim) / k);
    if (x <= beta) {
        const Point join = MpdExterior(beta, rim, ejectaScale);
        const float drop = rim - join.height;
        const float slope = drop > 0.0f ? std::max(join.slope, -3.0f * drop / beta) : join.slope;
        const float t = x / beta;
        return rim + (join.height - rim) * t * t * (3.0f - 2.0f * t) +
               beta * slope * t * t * (t - 1.0f);
    }
    return MpdExterior(x, rim, ejectaScale).height;
}

// FassettThomsonExterior and FassettThomsonShape (craters.h:330-362).
float FtExterior(float x, float start) {
    const float polynomial =
        0.188253307f + x * (-0.187050452f + x * (0.01844746f + x * 0.01505647f));
    const float u = (x - start) / (1.5f - start);
    return polynomial * (1.0f - u * u);
}

float FtShape(float x, std::uint32_t rim, float freshRim, float currentDepth) {
    if (x > 1.5f) {
        return 0.0f;
    }
    const float interior =
        -0.228809953f + x * (0.227533882f + x * (0.083116795f - x * 0.039499407f));
    float h;
    if (rim == 2u) {
        if (x <= 0.98f) {
            h = interior;
        } else if (x <= 1.02f) {
            h = freshRim;
        } else {
            h = FtExterior(x, 1.02f);
        }
    } else {
        const float edge = rim == 0u ? 0.98f : 1.0f;
        if (x <= 0.2f) {
            h = -0.181f;
        } else if (x <= edge) {
            h = interior;
        } else {
            h = FtExterior(x, edge);
        }
    }
    return std::max(h, freshRim - currentDepth);
}

// KilometerLaw (craters.h:432-435).
float KilometerLaw(float coefficient, float exponent, float diameter, bool volume) {
    const float kilometers = coefficient * std::pow(diameter / 1000.0f, exponent);
    return volume ? kilometers * 1e9f : kilometers * 1000.0f;
}

// ComplexProfileHeight with KernelComplexPeak, KernelPeakDome and KernelKeptDepth
// (craters.h:444-494).
float ComplexHeight(const tg_crater_profile& p, float distance, float diameter,
                    const ShapeState& c) {
    const float radius = 0.5f * diameter;
    const float x = distance / radius;
    const float depth = c.currentDepth * diameter;
    const float rim = c.currentRim * diameter;
    if (x >= 1.0f) {
        if (x >= p.complex_reach_radii) {
            return 0.0f;
        }
        const float edge = std::pow(p.complex_reach_radii, -p.ejecta_exponent);
        return rim * (std::pow(x, -p.ejecta_exponent) - edge) / (1.0f - edge);
    }
    const float floorShare = std::min(
        0.9f, KilometerLaw(p.floor_coefficient, p.floor_exponent, diameter, false) / diameter);
    float height = rim - depth;
    if (x > floorShare) {
        const float t = (x - floorShare) / (1.0f - floorShare);
        height += depth * t * t;
    }
    // The fresh central peak's height and base, scaled past the laws' fitted range.
    const float fitted =
        p.peak_fit_max_diameter > 0.0f ? std::min(diameter, p.peak_fit_max_diameter) : diameter;
    const float peakHeight =
        KilometerLaw(p.peak_height_coefficient, p.peak_height_exponent, fitted, false);
    float base;
    if (p.peak_from_volume == 1u) {
        const float volume =
            KilometerLaw(p.peak_width_coefficient, p.peak_width_exponent, fitted, true);
        base = std::sqrt(3.0f * volume / (3.14159265358979f * peakHeight));
    } else {
        base = 0.5f * KilometerLaw(p.peak_width_coefficient, p.peak_width_exponent, fitted, false);
    }
    base *= diameter / fitted;
    float dome = 0.0f;
    if (distance < base) {
        const float u = 1.0f - (distance / base) * (distance / base);
        dome = peakHeight * u * u;
    }
    const float kept = c.freshDepth > 0.0f ? c.currentDepth / c.freshDepth : 1.0f;
    return height + kept * dome;
}

// KernelTerraceEase (craters.h:968-974).
float TerraceEase(float v, float rounding) {
    const float steepest = 1.0f / (1.0f - 0.5f * rounding);
    if (v < rounding) {
        return steepest * v * v / (2.0f * rounding);
    }
    return steepest * (v - 0.5f * rounding);
}

double SmoothstepD(double t) {
    t = std::clamp(t, 0.0, 1.0);
    return t * t * (3.0 - 2.0 * t);
}

}  // namespace

double ProfileHeight(const tg_crater_shape& k, double distance, double diameter,
                     double freshDepth, double freshRim, double currentDepth, double currentRim) {
    const ShapeState s{static_cast<float>(freshDepth), static_cast<float>(freshRim),
                       static_cast<float>(currentDepth), static_cast<float>(currentRim)};
    const float d = static_cast<float>(distance);
    const float D = static_cast<float>(diameter);
    const tg_crater_profile& p = k.profile;
    if (p.complex_ == 1u && diameter >= double(p.transition_diameter)) {
        return static_cast<double>(ComplexHeight(p, d, D, s));
    }
    if (p.simple == TG_PROFILE_MPD) {
        if (!(currentDepth > 0.0 && currentRim >= 0.0 && currentRim <= freshRim)) {
            throw std::invalid_argument(
                "Crater state needs a positive depth and 0 <= rim height <= fresh rim height");
        }
        return static_cast<double>(MpdHeight(d, D, s));
    }
    return static_cast<double>(FtShape(d / (0.5f * D), p.rim, s.freshRim, s.currentDepth) * D);
}

void AppendTerraces(std::uint64_t id, double diameter, double freshDepth, double freshRim,
                    double currentDepth, double currentRim, const Terraces& terraces,
                    const tg_crater_shape& kernel, tg_crater_record& record,
                    std::vector<float>& data) {
    if (!terraces.enabled || diameter < terraces.onsetDiameter) {
        return;
    }
    const double freshness = freshDepth > 0.0 ? currentDepth / freshDepth : 1.0;
    const double strength =
        SmoothstepD((freshness - terraces.minFreshness) / (1.0 - terraces.minFreshness));
    if (!(strength > 0.0)) {
        return;
    }
    const auto height = [&](double distance) {
        return ProfileHeight(kernel, distance, diameter, freshDepth, freshRim, currentDepth,
                             currentRim);
    };
    // The floor edge: the complex shape's floor law, which for the lunar shape is the default
    // (0.187 km D^1.249).
    const tg_crater_profile& p = kernel.profile;
    const float floorCoefficient = p.complex_ == 1u ? p.floor_coefficient : 0.187f;
    const float floorExponent = p.complex_ == 1u ? p.floor_exponent : 1.249f;
    const double floor = static_cast<double>(
        KilometerLaw(floorCoefficient, floorExponent, static_cast<float>(diameter), false));
    const double inner = std::clamp(floor / diameter, 0.2, 0.9);
    const double radius = 0.5 * diameter;
    const double drop = height(radius) - height(inner * radius);
    const double meanSlope = drop / ((1.0 - inner) * radius);
    const double benchSlope = std::tan(terraces.benchSlopeDegrees * kDegree);
    const double scarpSlope = std::tan(terraces.scarpSlopeDegrees * kDegree) *
                              (1.0 - 0.5 * kTerraceScarpRounding) * kTerraceScarpSqueeze;
    const double slopeShare =
        meanSlope > 0.0 ? std::clamp(benchSlope / meanSlope, 0.02, 0.5) : 0.5;
    const double size = std::clamp((diameter - terraces.onsetDiameter) /
                                       (terraces.fullDiameter - terraces.onsetDiameter),
                                   0.0, 1.0);
    Stream random(terraces.seed, kCraterTerraceStream, std::uint32_t(id),
                  std::uint32_t(id >> 32));
    const double spread = terraces.maxBenches - terraces.minBenches;
    const int benches = std::clamp(
        static_cast<int>(std::lround(terraces.minBenches + spread * size + random.NextDouble() -
                                     0.5)),
        terraces.minBenches, terraces.maxBenches);
    // Five arc harmonics, orders 2-6, weights 1/k normalized, random phases.
    double arcs[10];
    {
        double total = 0.0;
        for (int k = 2; k < 7; ++k) {
            total += 1.0 / k;
        }
        for (int k = 2; k < 7; ++k) {
            arcs[2 * (k - 2)] = 1.0 / k / total;
            arcs[2 * (k - 2) + 1] = 2.0 * kPi * random.NextDouble();
        }
    }
    const double cover =
        terraces.arcShare >= 1.0
            ? 1.0
            : terraces.arcShare + std::max(0.0, kFullTerraceCover - terraces.arcShare) * size;
    double cut = -2.0;
    if (cover < 1.0) {
        // The (1 - cover) quantile of the arc sum over 360 bearings.
        constexpr std::size_t kSamples = 360;
        std::vector<double> values(kSamples);
        for (std::size_t n = 0; n < kSamples; ++n) {
            const double theta = 2.0 * kPi * static_cast<double>(n) / double{kSamples};
            double sum = 0.0;
            for (std::size_t k = 0; k < 5; ++k) {
                sum += arcs[2 * k] *
                       std::cos((static_cast<double>(k) + 2.0) * theta + arcs[2 * k + 1]);
            }
            values[n] = sum;
        }
        const std::size_t at = static_cast<std::size_t>((1.0 - cover) * double{kSamples - 1});
        std::nth_element(values.begin(), values.begin() + std::ptrdiff_t(at), values.end());
        cut = values[at];
    }
    record.terrace_count = static_cast<std::uint32_t>(benches);
    record.terrace_offset = static_cast<std::uint32_t>(data.size());
    for (const double v : {inner, slopeShare, strength, kTerraceScarpRounding, cut,
                           kTerraceArcWidth}) {
        data.push_back(static_cast<float>(v));
    }
    for (const double v : arcs) {
        data.push_back(static_cast<float>(v));
    }
    const auto scallopCount = [&random] {
        return kMinTerraceScallops +
               std::floor((kMaxTerraceScallops - kMinTerraceScallops + 1.0) * random.NextDouble());
    };
    const double rimScallop = kTerraceRimScallop * terraces.rimVariation * strength *
                              (0.6 + 0.8 * random.NextDouble());
    // Drawn in the order written: the scallop count, then the phase.
    const double rimScallops = scallopCount();
    const double rimPhase = random.NextDouble();
    for (const double v : {rimScallop, rimScallops, rimPhase, terraces.blockGaps}) {
        data.push_back(static_cast<float>(v));
    }
    // The wall's height from the floor edge to the rim crest, made non-decreasing.
    std::vector<double> wall(kTerraceWallSamples + 1);
    for (std::size_t n = 0; n < wall.size(); ++n) {
        const double t = static_cast<double>(n) / kTerraceWallSamples;
        const double h = height((inner + t * (1.0 - inner)) * radius);
        wall[n] = n == 0 ? h : std::max(h, wall[n - 1]);
    }
    const double rise = wall.back() - wall.front();
    // The share of the wall's width below which it reaches `share` of its rise.
    const auto across = [&](double share) {
        if (!(rise > 0.0)) {
            return share;
        }
        const double target = wall.front() + share * rise;
        const auto above = std::lower_bound(wall.begin(), wall.end(), target);
        if (above == wall.begin()) {
            return 0.0;
        }
        if (above == wall.end()) {
            return 1.0;
        }
        const auto n = static_cast<double>(above - wall.begin());
        const double gap = *above - *(above - 1);
        const double part = gap > 0.0 ? (*above - target) / gap : 0.0;
        return (n - part) / kTerraceWallSamples;
    };
    // Step heights: random rises, normalized; their bases across the wall, kept apart.
    std::vector<double> heights{0.0};
    double total = 0.0;
    for (int i = 0; i < benches; ++i) {
        total += 0.5 + random.NextDouble();
        heights.push_back(total);
    }
    std::vector<double> bases;
    for (double& h : heights) {
        h /= total;
        bases.push_back(across(h));
    }
    bases.back() = 1.0;
    const double narrowest = kNarrowestTerraceStep / benches;
    for (std::size_t k = 1; k + 1 < bases.size(); ++k) {
        bases[k] = std::max(bases[k], bases[k - 1] + narrowest);
    }
    for (std::size_t k = bases.size() - 2; k > 0; --k) {
        bases[k] = std::min(bases[k], bases[k + 1] - narrowest);
    }
    // Each inner scarp: base, scallops, wobble and block offset within its room.
    const double variation = terraces.rimVariation / (1.0 - inner);
    for (std::size_t k = 1; k < bases.size() - 1; ++k) {
        const double room =
            kTerraceScarpRoom * std::min(bases[k] - bases[k - 1], bases[k + 1] - bases[k]);
        double scallop = std::min(variation * (0.6 + 0.8 * random.NextDouble()), 0.9 * room);
        const double scallops = scallopCount();
        const double phase = random.NextDouble();
        double wobble = std::min(0.5 * variation * random.NextDouble(), 0.4 * room);
        const double wobblePhase = 2.0 * kPi * random.NextDouble();
        double block = terraces.blockOffset * room;
        const double reach = 2.0 / 3.0 * scallop + wobble + block;
        if (reach > kTerraceScarpReach * room) {
            const double fit = kTerraceScarpReach * room / reach;
            scallop *= fit;
            wobble *= fit;
            block *= fit;
        }
        const double key = static_cast<double>(random.NextU32() >> 8u);
        for (const double v : {bases[k], scallop, scallops, phase, wobble, wobblePhase, block, key}) {
            data.push_back(static_cast<float>(v));
        }
    }
    const auto wallAt = [&](double t) {
        const double at = std::clamp(t, 0.0, 1.0) * kTerraceWallSamples;
        const auto n = std::min(static_cast<std::size_t>(at), wall.size() - 2);
        return wall[n] + (at - static_cast<double>(n)) * (wall[n + 1] - wall[n]);
    };
    // A step's stair at `within` its width: the bench at `level`, eased into its scarps.
    const auto stairAt = [&](double within, double level, double scarp) {
        const auto rounding = static_cast<float>(kTerraceScarpRounding);
        const double lower = scarp * level;
        const double upper = scarp * (1.0 - level);
        double stair = level;
        if (within < lower) {
            stair = level * (1.0 - double{TerraceEase(static_cast<float>(1.0 - within / lower),
                                                      rounding)});
        } else if (within > 1.0 - upper) {
            stair = level + (1.0 - level) *
                                double{TerraceEase(
                                    static_cast<float>((within - (1.0 - upper)) / upper),
                                    rounding)};
        }
        return slopeShare * within + (1.0 - slopeShare) * stair;
    };
    // The area-weighted mean height of a step, stepped at `level` (unstepped below 0).
    const auto meanHeight = [&](double from, double to, double level, double scarp) {
        double weighted = 0.0;
        double weights = 0.0;
        for (int n = 0; n < kTerraceMeanSamples; ++n) {
            const double within = (n + 0.5) / kTerraceMeanSamples;
            const double r = inner + (from + within * (to - from)) * (1.0 - inner);
            const double shaped = level < 0.0 ? within : stairAt(within, level, scarp);
            weighted += r * wallAt(from + shaped * (to - from));
            weights += r;
        }
        return weighted / weights;
    };
    // Each step: the scarp share its slope needs, the share of its bench kept, and the bench
    // level that keeps its mean height, by bisection.
    for (std::size_t k = 0; k + 1 < bases.size(); ++k) {
        const double from = bases[k];
        const double to = bases[k + 1];
        const double stepSlope =
            (wallAt(to) - wallAt(from)) / ((to - from) * (1.0 - inner) * radius);
        const double needed =
            scarpSlope > benchSlope ? (stepSlope - benchSlope) / (scarpSlope - benchSlope) : 1.0;
        const double scarp = std::clamp(needed, terraces.scarpShare, kMostTerraceScarpShare);
        const double kept =
            1.0 - SmoothstepD((needed - kSteepTerraceStep) / (1.0 - kSteepTerraceStep));
        const double target = meanHeight(from, to, -1.0, scarp);
        double low = kTerraceLevelMargin;
        double high = 1.0 - kTerraceLevelMargin;
        for (int i = 0; i < kTerraceLevelIterations && high - low > 0.0; ++i) {
            const double middle = 0.5 * (low + high);
            (!(meanHeight(from, to, middle, scarp) < target) ? high : low) = middle;
        }
        for (const double v : {0.5 * (low + high), scarp, kept}) {
            data.push_back(static_cast<float>(v));
        }
    }
}

}  // namespace tgh
