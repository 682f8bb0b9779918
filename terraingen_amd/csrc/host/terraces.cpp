// SPDX-License-Identifier: BSD-3-Clause
//
// Wall terraces: the data block a terraced crater's kernel reads (KernelTerracedRadii,
// include/terraingen/layers/kernels/craters.h:986-1030), prepared on the host as the reference's
// AppendKernelTerraces prepares it (src/layers/craters/shaping/crater_kernel.cpp:312-459).
//
// The preparation runs in two phases:
//   1. TerraceDraws takes every value the crater's terrace stream gives, in the stream's order:
//      the bench count's jitter, five arc phases, the rim scallop's three draws, one rise per
//      step, then per inner scarp five doubles and a block key. Nothing drawn depends on anything
//      derived except the bench count, which the first draw settles.
//   2. The derivation works from those draws and the crater's wall: WallTable samples the
//      profile from the floor edge to the rim crest (monotone), maps a share of the wall's rise to
//      a share of its width and back; the step bases split the rise by the drawn rises; each step
//      gets the scarp share its slope needs and the bench level that keeps its area-weighted mean
//      height (bisection).
// The block is then written section by section: header, arc harmonics, rim scallop, scarps,
// steps. Byte-equal data needs the draws in stream order and each FP64 expression evaluated as
// the reference evaluates it before the narrowing to float; both are kept here.

#include "prepare.h"
#include "profile_laws.h"
#include "random.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <stdexcept>

namespace tgh {

namespace {

constexpr double kPi = 3.141592653589793238462643383279502884;

// Tuning of the terrace preparation. The values are the reference's (crater_kernel.cpp:26-60);
// the kernel's reading of the block depends on them.
namespace tune {
constexpr double kFullCover = 0.9;          // arc share reached at the full diameter
constexpr double kArcWidth = 0.3;           // written to the header for the kernel
constexpr int kWallSamples = 256;           // the wall table's intervals
constexpr double kNarrowestStep = 0.25;     // of the wall width, divided by the bench count
constexpr double kScarpRounding = 0.4;      // the ease of a scarp's foot and crest
constexpr double kMostScarpShare = 0.9;
constexpr double kSteepStep = 0.7;          // above this scarp need a bench fades
constexpr double kScarpSqueeze = 0.75;
constexpr int kMeanSamples = 64;            // midpoint samples of a step's mean height
constexpr int kLevelIterations = 30;
constexpr double kLevelMargin = 0.1;        // bench levels stay inside [0.1, 0.9]
constexpr double kScarpRoom = 0.4;          // of the narrower neighbouring step
constexpr double kScarpReach = 1.2;         // of the room: scallop, wobble and block together
constexpr double kFewestScallops = 6.0;
constexpr double kMostScallops = 12.0;
constexpr double kRimScallop = 0.5;
constexpr int kArcHarmonics = 5;            // orders 2..6
constexpr std::size_t kArcBearings = 360;   // bearings of the arc cut's quantile
}  // namespace tune

double Clamp01Smooth(double t) {
    t = std::clamp(t, 0.0, 1.0);
    return t * t * (3.0 - 2.0 * t);
}

// The drawn part of one crater's terraces (phase 1).
struct TerraceDraws {
    int benches = 0;
    std::array<double, tune::kArcHarmonics> arcPhase{};
    double rimAmount = 0.0, rimCount = 0.0, rimPhase = 0.0;
    std::vector<double> rises;  // one per step, each in [0.5, 1.5)
    struct Scarp {
        double amount, count, phase, wobble, wobblePhase;
        std::uint32_t key;
    };
    std::vector<Scarp> scarps;  // one per inner scarp: benches - 1
};

// A scallop count: an integer from the fewest to the most, uniform.
double ScallopCount(double u) {
    return tune::kFewestScallops +
           std::floor((tune::kMostScallops - tune::kFewestScallops + 1.0) * u);
}

TerraceDraws DrawTerraces(Stream& s, const Terraces& opts, double size) {
    TerraceDraws d;
    const double spread = opts.maxBenches - opts.minBenches;
    const double jitter = s.NextDouble();
    d.benches = std::clamp(
        static_cast<int>(std::lround(opts.minBenches + spread * size + jitter - 0.5)),
        opts.minBenches, opts.maxBenches);
    for (double& phase : d.arcPhase) {
        phase = 2.0 * kPi * s.NextDouble();
    }
    d.rimAmount = 0.6 + 0.8 * s.NextDouble();
    d.rimCount = ScallopCount(s.NextDouble());
    d.rimPhase = s.NextDouble();
    for (int i = 0; i < d.benches; ++i) {
        d.rises.push_back(0.5 + s.NextDouble());
    }
    for (int i = 1; i < d.benches; ++i) {
        TerraceDraws::Scarp c{};
        c.amount = 0.6 + 0.8 * s.NextDouble();
        c.count = ScallopCount(s.NextDouble());
        c.phase = s.NextDouble();
        c.wobble = s.NextDouble();
        c.wobblePhase = 2.0 * kPi * s.NextDouble();
        c.key = s.NextU32() >> 8u;
        d.scarps.push_back(c);
    }
    return d;
}

// The crater's wall from the floor edge (t = 0) to the rim crest (t = 1), sampled at
// tune::kWallSamples + 1 points and made non-decreasing.
class WallTable {
  public:
    template <class Height>
    WallTable(double inner, double radius, Height&& height) : h_(tune::kWallSamples + 1) {
        for (std::size_t n = 0; n < h_.size(); ++n) {
            const double t = static_cast<double>(n) / tune::kWallSamples;
            const double v = height((inner + t * (1.0 - inner)) * radius);
            h_[n] = n == 0 ? v : std::max(v, h_[n - 1]);
        }
    }

    // Height at the share t of the wall's width, linear between samples.
    double At(double t) const {
        const double at = std::clamp(t, 0.0, 1.0) * tune::kWallSamples;
        const std::size_t n = std::min(static_cast<std::size_t>(at), h_.size() - 2);
        return h_[n] + (at - static_cast<double>(n)) * (h_[n + 1] - h_[n]);
    }

    // The share of the wall's width at which it has risen `share` of its rise: the first sample
    // not below the target, interpolated back to the crossing.
    double WidthFor(double share) const {
        const double rise = h_.back() - h_.front();
        if (!(rise > 0.0)) {
            return share;
        }
        const double target = h_.front() + share * rise;
        std::size_t lo = 0, hi = h_.size();  // h_[lo..hi) holds the first sample >= target
        while (lo < hi) {
            const std::size_t mid = lo + (hi - lo) / 2;
            if (h_[mid] < target) {
                lo = mid + 1;
            } else {
                hi = mid;
            }
        }
        if (lo == 0) {
            return 0.0;
        }
        if (lo == h_.size()) {
            return 1.0;
        }
        const double gap = h_[lo] - h_[lo - 1];
        const double part = gap > 0.0 ? (h_[lo] - target) / gap : 0.0;
        return (static_cast<double>(lo) - part) / tune::kWallSamples;
    }

  private:
    std::vector<double> h_;
};

// One step of the wall, from base `from` to base `to` (shares of the wall's width), shaped as a
// bench at `level` between eased scarps, tilted by the slope share.
struct Step {
    const WallTable& wall;
    double inner, slopeShare, from, to;

    // The stair's height share at `within` the step (0..1) for bench level and scarp share.
    double Stair(double within, double level, double scarp) const {
        const auto rounding = static_cast<float>(tune::kScarpRounding);
        const double lower = scarp * level;
        const double upper = scarp * (1.0 - level);
        double stair = level;
        if (within < lower) {
            const auto v = static_cast<float>(1.0 - within / lower);
            stair = level * (1.0 - double{tg::TerraceEase(v, rounding)});
        } else if (within > 1.0 - upper) {
            const auto v = static_cast<float>((within - (1.0 - upper)) / upper);
            stair = level + (1.0 - level) * double{tg::TerraceEase(v, rounding)};
        }
        return slopeShare * within + (1.0 - slopeShare) * stair;
    }

    // Area-weighted (by radius) mean height over the step, stepped at `level`; the unstepped
    // wall for a negative level.
    double MeanHeight(double level, double scarp) const {
        double weighted = 0.0;
        double weights = 0.0;
        for (int n = 0; n < tune::kMeanSamples; ++n) {
            const double within = (n + 0.5) / tune::kMeanSamples;
            const double r = inner + (from + within * (to - from)) * (1.0 - inner);
            const double shaped = level < 0.0 ? within : Stair(within, level, scarp);
            weighted += r * wall.At(from + shaped * (to - from));
            weights += r;
        }
        return weighted / weights;
    }

    // The bench level in [margin, 1 - margin] whose stair keeps the wall's mean height.
    double Level(double scarp) const {
        const double target = MeanHeight(-1.0, scarp);
        double low = tune::kLevelMargin;
        double high = 1.0 - tune::kLevelMargin;
        for (int i = 0; i < tune::kLevelIterations && high - low > 0.0; ++i) {
            const double middle = 0.5 * (low + high);
            if (MeanHeight(middle, scarp) < target) {
                low = middle;
            } else {
                high = middle;
            }
        }
        return 0.5 * (low + high);
    }
};

// The arc cut: the value the arc harmonics' sum exceeds over a share `cover` of the bearings
// (-2, below any sum, when every bearing is terraced).
double ArcCut(const std::array<double, 2 * tune::kArcHarmonics>& arcs, double cover) {
    if (!(cover < 1.0)) {
        return -2.0;
    }
    std::vector<double> sums(tune::kArcBearings);
    for (std::size_t n = 0; n < sums.size(); ++n) {
        const double theta = 2.0 * kPi * static_cast<double>(n) / double{tune::kArcBearings};
        double sum = 0.0;
        for (std::size_t k = 0; k < tune::kArcHarmonics; ++k) {
            sum += arcs[2 * k] * std::cos((static_cast<double>(k) + 2.0) * theta + arcs[2 * k + 1]);
        }
        sums[n] = sum;
    }
    std::sort(sums.begin(), sums.end());
    return sums[static_cast<std::size_t>((1.0 - cover) * double{tune::kArcBearings - 1})];
}

void Put(std::vector<float>& data, std::initializer_list<double> values) {
    for (const double v : values) {
        data.push_back(static_cast<float>(v));
    }
}

}  // namespace

double ProfileHeight(const tg_crater_shape& k, double distance, double diameter,
                     double freshDepth, double freshRim, double currentDepth, double currentRim) {
    const auto D = static_cast<float>(diameter);
    const tg::ProfileState s{D, static_cast<float>(freshDepth), static_cast<float>(freshRim),
                             static_cast<float>(currentDepth), static_cast<float>(currentRim)};
    const auto d = static_cast<float>(distance);
    const tg_crater_profile& p = k.profile;
    // CraterProfile::Height of the closed-form profiles (complex.cpp:69-77,
    // martin_parkes_dunstan.cpp:9-19, fassett_thomson.cpp:14-18): the kernel's FP32 laws.
    if (p.complex_ == 1u && diameter >= double(p.transition_diameter)) {
        return static_cast<double>(tg::ComplexProfileHeight(p, d, D, s));
    }
    if (p.simple == TG_PROFILE_MPD) {
        if (!(currentDepth > 0.0 && currentRim >= 0.0 && currentRim <= freshRim)) {
            throw std::invalid_argument(
                "Crater state needs a positive depth and 0 <= rim height <= fresh rim height");
        }
        return static_cast<double>(tg::MpdHeight(d, D, s));
    }
    return static_cast<double>(
        tg::FassettThomsonShape(d / (0.5f * D), p.rim, s.freshRim, s.currentDepth) * D);
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
        Clamp01Smooth((freshness - terraces.minFreshness) / (1.0 - terraces.minFreshness));
    if (!(strength > 0.0)) {
        return;
    }
    const auto height = [&](double distance) {
        return ProfileHeight(kernel, distance, diameter, freshDepth, freshRim, currentDepth,
                             currentRim);
    };
    const double radius = 0.5 * diameter;
    // The floor edge, as a share of the radius: the complex shape's floor law, or the lunar one
    // (0.187 km D^1.249) for simple profiles.
    const tg_crater_profile& p = kernel.profile;
    const bool complex = p.complex_ == 1u;
    const double floorWidth = static_cast<double>(
        tg::KilometerLaw(complex ? p.floor_coefficient : 0.187f,
                         complex ? p.floor_exponent : 1.249f, static_cast<float>(diameter),
                         false));
    const double inner = std::clamp(floorWidth / diameter, 0.2, 0.9);
    // Slopes: the wall's mean, a bench's and a rounded scarp's; benches take a share of the mean.
    const double drop = height(radius) - height(inner * radius);
    const double meanSlope = drop / ((1.0 - inner) * radius);
    const double degree = kPi / 180.0;
    const double benchSlope = std::tan(terraces.benchSlopeDegrees * degree);
    const double scarpSlope = std::tan(terraces.scarpSlopeDegrees * degree) *
                              (1.0 - 0.5 * tune::kScarpRounding) * tune::kScarpSqueeze;
    const double slopeShare =
        meanSlope > 0.0 ? std::clamp(benchSlope / meanSlope, 0.02, 0.5) : 0.5;
    const double size = std::clamp((diameter - terraces.onsetDiameter) /
                                       (terraces.fullDiameter - terraces.onsetDiameter),
                                   0.0, 1.0);

    // Phase 1: every draw of the crater's terrace stream.
    Stream stream(terraces.seed, kCraterTerraceStream, std::uint32_t(id),
                  std::uint32_t(id >> 32));
    const TerraceDraws draws = DrawTerraces(stream, terraces, size);

    // Phase 2: the block. Header and arc harmonics (weights 1/k normalized, k = 2..6).
    std::array<double, 2 * tune::kArcHarmonics> arcs{};
    double weightSum = 0.0;
    for (int k = 2; k < 2 + tune::kArcHarmonics; ++k) {
        weightSum += 1.0 / k;
    }
    for (int k = 2; k < 2 + tune::kArcHarmonics; ++k) {
        arcs[2 * (k - 2)] = 1.0 / k / weightSum;
        arcs[2 * (k - 2) + 1] = draws.arcPhase[k - 2];
    }
    const double cover =
        terraces.arcShare >= 1.0
            ? 1.0
            : terraces.arcShare + std::max(0.0, tune::kFullCover - terraces.arcShare) * size;
    record.terrace_count = static_cast<std::uint32_t>(draws.benches);
    record.terrace_offset = static_cast<std::uint32_t>(data.size());
    Put(data, {inner, slopeShare, strength, tune::kScarpRounding, ArcCut(arcs, cover),
               tune::kArcWidth});
    for (const double v : arcs) {
        data.push_back(static_cast<float>(v));
    }
    // The rim crest's scallops.
    Put(data, {tune::kRimScallop * terraces.rimVariation * strength * draws.rimAmount,
               draws.rimCount, draws.rimPhase, terraces.blockGaps});

    // Step bases: the wall's rise split by the drawn rises, mapped to widths and kept at least
    // the narrowest step apart (forward, then backward).
    const WallTable wall(inner, radius, height);
    std::vector<double> bases{wall.WidthFor(0.0)};
    double risen = 0.0;
    for (const double r : draws.rises) {
        risen += r;
    }
    {
        double cumulative = 0.0;
        for (const double r : draws.rises) {
            cumulative += r;
            bases.push_back(wall.WidthFor(cumulative / risen));
        }
    }
    bases.back() = 1.0;
    const double narrowest = tune::kNarrowestStep / draws.benches;
    for (std::size_t k = 1; k + 1 < bases.size(); ++k) {
        bases[k] = std::max(bases[k], bases[k - 1] + narrowest);
    }
    for (std::size_t k = bases.size() - 2; k > 0; --k) {
        bases[k] = std::min(bases[k], bases[k + 1] - narrowest);
    }

    // Inner scarps: base, scallops, wobble and block offset, scaled together to fit the room
    // the neighbouring steps leave.
    const double variation = terraces.rimVariation / (1.0 - inner);
    for (std::size_t k = 1; k + 1 < bases.size(); ++k) {
        const TerraceDraws::Scarp& c = draws.scarps[k - 1];
        const double room =
            tune::kScarpRoom * std::min(bases[k] - bases[k - 1], bases[k + 1] - bases[k]);
        double scallop = std::min(variation * c.amount, 0.9 * room);
        double wobble = std::min(0.5 * variation * c.wobble, 0.4 * room);
        double block = terraces.blockOffset * room;
        const double reach = 2.0 / 3.0 * scallop + wobble + block;
        if (reach > tune::kScarpReach * room) {
            const double fit = tune::kScarpReach * room / reach;
            scallop *= fit;
            wobble *= fit;
            block *= fit;
        }
        Put(data, {bases[k], scallop, c.count, c.phase, wobble, c.wobblePhase, block,
                   static_cast<double>(c.key)});
    }

    // Steps: the scarp share the step's slope needs, the share of its bench kept, its level.
    for (std::size_t k = 0; k + 1 < bases.size(); ++k) {
        const Step step{wall, inner, slopeShare, bases[k], bases[k + 1]};
        const double slope = (wall.At(step.to) - wall.At(step.from)) /
                             ((step.to - step.from) * (1.0 - inner) * radius);
        const double needed =
            scarpSlope > benchSlope ? (slope - benchSlope) / (scarpSlope - benchSlope) : 1.0;
        const double scarp = std::clamp(needed, terraces.scarpShare, tune::kMostScarpShare);
        const double kept =
            1.0 - Clamp01Smooth((needed - tune::kSteepStep) / (1.0 - tune::kSteepStep));
        Put(data, {step.Level(scarp), scarp, kept});
    }
}

}  // namespace tgh
