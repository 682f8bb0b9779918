// SPDX-License-Identifier: BSD-3-Clause
//
// Crater preparation on the host: the reference's PrepareCraters and MakeCraterKernelBatch
// (src/layers/craters/shaping/{prepared,crater_kernel,irregularity,apply_options}.cpp) for
// closed-form profiles (Martin-Parkes-Dunstan or Fassett-Thomson, optionally complex from a
// transition diameter) under freshness degradation. It yields the kernel struct, one record per
// crater in apply order, the variable-length data, and each crater's footprint for binning:
//   order       oldest first, then larger, then smaller ID (apply_options.cpp:96-128);
//   resolution  1 - Falloff(D / spacing, 2, 4, smoothstep), zero-weight craters dropped
//               (src/layers/fade.cpp:18-52, prepared.cpp:60-87);
//   outline     ellipse and power-law rim harmonics drawn from the crater's ID
//               (irregularity.cpp:124-150,240-284);
//   state       FreshnessDegradation with the lunar fresh morphometry
//               (freshness_degradation.cpp:26-32, morphometry.cpp);
//   records     KernelStateRecord, split centre, outline data, noise and texture keys, wall
//               terraces (terraces.cpp), central peak massifs
//               (crater_kernel.cpp:181-199,312-519,581-606).

#include "common.h"
#include "lunar_models.h"
#include "placement.h"
#include "prepare.h"
#include "random.h"

#include <algorithm>
#include <cstdlib>
#include <limits>
#include <map>
#include <memory>
#include <numeric>

namespace tgh {

namespace {

constexpr double kPi = 3.141592653589793238462643383279502884;
constexpr double kDegree = kPi / 180.0;

double SmoothstepD(double t) {
    t = std::clamp(t, 0.0, 1.0);
    return t * t * (3.0 - 2.0 * t);
}

// Falloff with the smoothstep shape (fade.cpp:23-46).
double FalloffSmoothstep(double x, double start, double end) {
    if (x <= start) {
        return 1.0;
    }
    if (x >= end) {
        return 0.0;
    }
    const double t = (x - start) / (end - start);
    return 1.0 - SmoothstepD(t);
}

struct Morphometry {
    double depthRatio, rimHeightRatio;
};

// LunarFreshCraterDepthRatio and LunarFreshCraterMorphometry (morphometry.cpp:10-38); the complex
// laws above 15 km are evaluated through the FP32 kilometer law as the reference does
// (KilometerPowerLaw::At, complex.cpp:9-13).
float KilometerLawF(float coefficient, float exponent, float diameter) {
    return coefficient * std::pow(diameter / 1000.0f, exponent) * 1000.0f;
}

Morphometry LunarFresh(double d) {
    if (d >= 15000.0 && std::isfinite(d)) {
        const double depth = double(KilometerLawF(1.044f, 0.301f, float(d)));
        const double rim = double(KilometerLawF(0.236f, 0.399f, float(d)));
        return {depth / d, rim / d};
    }
    double ratio = 0.21;
    if (d <= 10.0) {
        ratio = 0.10;
    } else if (d < 40.0) {
        ratio = 0.11;
    } else if (d < 100.0) {
        ratio = 0.13;
    } else if (d < 200.0) {
        ratio = 0.15;
    } else if (d < 400.0) {
        ratio = 0.17;
    }
    return {ratio, 0.036822095};
}

struct Outline {
    double ellipticity = 1.0;
    double heading = 0.0;
    std::vector<double> amplitudes, phases;
    int firstHarmonic = 2;

    double MaxScale() const {
        double sum = 1.0;
        for (const double a : amplitudes) {
            sum += std::fabs(a);
        }
        return std::sqrt(ellipticity) * sum;
    }
};

// IrregularOutline with the power-law rim model (irregularity.cpp:240-284).
Outline IrregularOutline(const tgh::Crater& c, const Irregularity& irr) {
    Outline outline;
    if (!irr.enabled) {
        return outline;
    }
    Stream random(irr.seed, kCraterIrregularityStream, std::uint32_t(c.id),
                  std::uint32_t(c.id >> 32));
    double draws[9];
    for (double& d : draws) {
        d = random.NextDouble();
    }
    const double rough = draws[0], elongated = draws[1], r1 = draws[2], s1 = draws[4],
                 e1 = draws[6], h = draws[8];
    if (elongated < irr.elongatedShare) {
        outline.ellipticity = irr.ellipticityLow + (irr.ellipticityHigh - irr.ellipticityLow) * e1;
        outline.heading = kPi * h;
    }
    if (rough < irr.roughShare) {
        const double roughness = irr.roughnessLow + (irr.roughnessHigh - irr.roughnessLow) * r1;
        (void)s1;  // the spectral slope is a constant distribution: its draw is unused
        const double slope = irr.spectralSlope;
        double power = 0.0;
        for (int k = irr.minHarmonic; k <= irr.maxHarmonic; ++k) {
            power += std::pow(static_cast<double>(k), -2.0 * slope);
        }
        const double norm = roughness * std::sqrt(2.0 / power);
        outline.firstHarmonic = irr.minHarmonic;
        if (roughness > 0.0) {
            for (int k = irr.minHarmonic; k <= irr.maxHarmonic; ++k) {
                outline.amplitudes.push_back(norm * std::pow(static_cast<double>(k), -slope));
                outline.phases.push_back(2.0 * kPi * random.NextDouble());
            }
        }
    }
    return outline;
}

// AppendKernelPeaks (crater_kernel.cpp:461-519): a complex crater's central peak massifs.
void AppendPeaks(const tgh::Crater& c, const Morphometry& fresh, const Morphometry& current,
                 const Peaks& peaks, double transitionDiameter, tg_crater_record& r,
                 std::vector<float>& data) {
    if (!peaks.enabled || c.diameter < transitionDiameter) {
        return;
    }
    const double freshness =
        fresh.depthRatio > 0.0 ? std::clamp(current.depthRatio / fresh.depthRatio, 0.0, 1.0)
                               : 1.0;
    Stream random(peaks.seed, kCraterPeakStream, std::uint32_t(c.id), std::uint32_t(c.id >> 32));
    const int massifs =
        peaks.minMassifs +
        std::min(peaks.maxMassifs - peaks.minMassifs,
                 static_cast<int>((peaks.maxMassifs - peaks.minMassifs + 1) * random.NextDouble()));
    r.peak_count = static_cast<std::uint32_t>(massifs);
    r.peak_offset = static_cast<std::uint32_t>(data.size());
    const double key = static_cast<double>(random.NextU32() >> 8u);
    for (const double v : {peaks.roughness * freshness, peaks.roughnessWavelength, peaks.warp, key,
                           peaks.heightScale, peaks.widthScale}) {
        data.push_back(static_cast<float>(v));
    }
    const double chain = 2.0 * kPi * random.NextDouble();
    const double room = 2.3 - peaks.warp;
    for (int n = 0; n < massifs; ++n) {
        const bool main = n == 0;
        const double side = n % 2 == 0 ? kPi : 0.0;
        const double bearing = main ? 2.0 * kPi * random.NextDouble()
                                    : chain + side + peaks.chainSpreadDegrees * kDegree *
                                                         (random.NextDouble() - 0.5);
        const double reach = peaks.spreadRadii * (main ? 0.25 * random.NextDouble()
                                                       : 0.35 + 0.65 * random.NextDouble());
        const double aspect = 1.0 + (peaks.elongation - 1.0) * random.NextDouble();
        const double stretch = std::sqrt(aspect);
        const double crowding = massifs == 1 ? 1.0 : std::sqrt(2.0 / massifs);
        double radius = crowding * (main ? 1.0 : 0.6 * (0.5 + random.NextDouble()));
        radius = std::max(0.1, std::min(radius, (room - reach) / stretch));
        const double heading =
            main ? 2.0 * kPi * random.NextDouble() : bearing + 0.5 * (random.NextDouble() - 0.5);
        const double height = main ? 1.0 : 1.0 - peaks.heightVariation * random.NextDouble();
        const double exponent = std::clamp(
            (2.0 - 1.5 * peaks.sharpness) * (0.8 + 0.4 * random.NextDouble()), 0.4, 2.5);
        const double pull = 0.45 * peaks.lopsidedness * std::sqrt(random.NextDouble());
        const double pullBearing = 2.0 * kPi * random.NextDouble();
        for (const double v :
             {reach * std::cos(bearing), reach * std::sin(bearing), radius, stretch,
              std::cos(heading), std::sin(heading), height, exponent, pull * std::cos(pullBearing),
              pull * std::sin(pullBearing), 2.0 - peaks.sharpness}) {
            data.push_back(static_cast<float>(v));
        }
    }
}

}  // namespace

tg_crater_shape DefaultShapeKernel() {
    tg_crater_shape k{};
    k.profile.ejecta_exponent = 3.0f;
    k.profile.complex_reach_radii = 3.0f;
    k.profile.table_extent = 1.5f;
    k.reach_radii = 1.5f;
    k.profile_reach_radii = 1.5f;
    k.blend_start = 0.8f;
    k.blend_end = 1.2f;
    k.replace_start = 1.0f;
    k.replace_end = 1.5f;
    k.ejecta_reach_radii = 3.0f;
    k.rim_inner = 0.8f;
    k.rim_outer = 1.2f;
    k.rim_strength = 1.0f;
    k.noise_wavelength = 0.5f;
    k.noise_octaves = 4;
    k.noise_gain = 0.5f;
    k.noise_rim_wavelength = 0.1f;
    k.noise_rim_inner = 0.8f;
    k.noise_rim_outer = 1.2f;
    k.wall_streak_width = 0.15f;
    k.ejecta_lineation = 0.15f;
    k.min_freshness = 0.5f;
    return k;
}

double ProfileReachRadii(const CraterModels& m) {
    if (m.diffused) {
        return DiffusionTables::kReachRadii;  // DiffusedProfile::ReachRadii (diffusion.h:65-67)
    }
    const double simple = m.profile == TG_PROFILE_MPD ? 2.0 : 1.5;
    return m.complex_ ? std::fmax(simple, 3.0) : simple;
}

double CraterReachRadii(const CraterModels& m, const ApplyOptions& o) {
    const double reach = ProfileReachRadii(m);
    return (o.ejecta > 0.0 || o.mantling > 0.0) ? std::fmax(reach, o.ejectaReachRadii) : reach;
}

double MaxOutlineScale(const Irregularity& irr) {
    if (!irr.enabled) {
        return 1.0;
    }
    const int harmonics = irr.maxHarmonic - irr.minHarmonic + 1;
    const double rough =
        irr.roughShare > 0.0 ? 1.0 + irr.roughnessHigh * std::sqrt(2.0 * harmonics) : 1.0;
    const double elongated = irr.elongatedShare > 0.0 ? std::sqrt(irr.ellipticityHigh) : 1.0;
    return rough * elongated;
}

// SeededCraterApply (apply_options.cpp:58-72): per-purpose seeds from the layer's seed.
ApplyOptions SeededApply(ApplyOptions o, std::uint32_t layerSeed) {
    o.irregularity.seed = SubstreamSeed(layerSeed, kCraterIrregularityStream);
    o.noiseSeed = SubstreamSeed(layerSeed, kCraterNoiseStream);
    o.textureSeed = SubstreamSeed(layerSeed, kCraterTextureStream);
    o.peaks.seed = SubstreamSeed(layerSeed, kCraterPeakStream);
    o.terraces.seed = SubstreamSeed(layerSeed, kCraterTerraceStream);
    return o;
}

tg_crater_shape MakeShapeKernel(const CraterModels& m, const ApplyOptions& o) {
    tg_crater_shape k = DefaultShapeKernel();
    k.profile.simple = m.profile;
    if (m.diffused) {
        // DiffusedProfile's kernel block (diffusion.cpp:176-181): tables of 600 cells over half
        // the reach, the ages first in the data.
        k.profile.diffused = 1u;
        k.profile.table_extent = static_cast<float>(0.5 * DiffusionTables::kReachRadii);
        k.profile.table_cells = DiffusionTables::kCells;
        k.profile.ages_offset = 0u;
        k.profile.age_count = static_cast<std::uint32_t>(DiffusionAges().size());
    }
    if (m.complex_) {
        // SetKernelComplexShape (complex.cpp:29-43) with the lunar complex shape's laws.
        k.profile.complex_ = 1u;
        k.profile.transition_diameter = static_cast<float>(m.transitionDiameter);
        k.profile.floor_coefficient = 0.187f;
        k.profile.floor_exponent = 1.249f;
        k.profile.peak_height_coefficient = static_cast<float>(0.589e-3);
        k.profile.peak_height_exponent = 1.969f;
        k.profile.peak_width_coefficient = static_cast<float>(0.987e-7);
        k.profile.peak_width_exponent = 5.078f;
        k.profile.peak_from_volume = 1u;
        k.profile.peak_fit_max_diameter = 51000.0f;
        k.profile.ejecta_exponent = 3.0f;
        k.profile.complex_reach_radii = 3.0f;
    }
    // SetShapeOptions (crater_kernel.cpp:106-139).
    const double reach = ProfileReachRadii(m);
    k.reach_radii = static_cast<float>(CraterReachRadii(m, o));
    k.profile_reach_radii = static_cast<float>(reach);
    k.blend_start = 0.8f;
    k.blend_end = 1.2f;
    k.blend_shape = 2u;
    k.interior_replacement = 0.0f;
    k.replace_start = 1.0f;
    k.replace_end = static_cast<float>(reach);
    k.mantling = static_cast<float>(o.mantling);
    k.ejecta_reach_radii = static_cast<float>(o.ejectaReachRadii);
    k.rim_sigma = static_cast<float>(o.rimSigma);
    k.rim_inner = 0.8f;
    k.rim_outer = 1.2f;
    k.rim_strength = 1.0f;
    k.rim_shape = 2u;
    k.noise_amplitude = static_cast<float>(o.noiseAmplitude);
    k.noise_wavelength = 0.5f;
    k.noise_octaves = 4;
    k.noise_gain = 0.5f;
    k.noise_rim_amplitude = static_cast<float>(o.noiseRimAmplitude);
    k.noise_rim_wavelength = 0.1f;
    k.noise_rim_inner = 0.8f;
    k.noise_rim_outer = 1.2f;
    k.wall_streaks = static_cast<float>(o.wallStreaks);
    k.wall_streak_width = 0.15f;
    k.ejecta = static_cast<float>(o.ejecta);
    k.ejecta_lineation = 0.15f;
    k.min_freshness = 0.5f;
    return k;
}

std::vector<std::size_t> ApplyOrder(const std::vector<tgh::Crater>& craters, unsigned threads) {
    struct Key {
        double age, diameter;
        std::uint64_t id;
        std::size_t index;
    };
    std::vector<Key> keys(craters.size());
    for (std::size_t k = 0; k < craters.size(); ++k) {
        keys[k] = {craters[k].ageGyr, craters[k].diameter, craters[k].id, k};
    }
    ParallelSort(keys, [](const Key& p, const Key& q) {
        if (p.age != q.age) {
            return p.age > q.age;
        }
        if (p.diameter != q.diameter) {
            return p.diameter > q.diameter;
        }
        return p.id != q.id ? p.id < q.id : p.index < q.index;
    }, threads);
    std::vector<std::size_t> index(craters.size());
    for (std::size_t k = 0; k < keys.size(); ++k) {
        index[k] = keys[k].index;
    }
    return index;
}

namespace {

// One crater's record and data, appended to `b` (CraterLayer's preparation of a crater:
// PrepareCraters, prepared.cpp:60-87, then MakeCraterKernelBatch's record and data,
// crater_kernel.cpp:94-283). `tables`, `tableOffsets` and `surface` are the default models' shared
// diffusion state (null otherwise).
struct CraterPrep {
    const CraterModels& m;
    const ApplyOptions& o;
    double spacing, reachRadii;
    LunarSurface* surface;
    DiffusionTables* tables;
    std::map<std::pair<double, double>, std::uint32_t>* tableOffsets;

    // The crater's resolution weight at the spacing: 0 drops it.
    double Resolution(const tgh::Crater& c) const {
        return 1.0 - FalloffSmoothstep(c.diameter / spacing, 2.0, 4.0);
    }

    void Append(const tgh::Crater& c, double resolution, Batch& b) const {
        const Outline outline = IrregularOutline(c, o.irregularity);
        const Morphometry fresh = LunarFresh(c.diameter);
        Morphometry current;
        double tau = 0.0;
        if (m.diffused) {
            // DiffusionDegradation::WornState (diffusion.cpp:299-316) with the matched diffusion
            // ages (EquilibriumDiffusionAges, diffusivities.cpp:28-46): worn to the visibility
            // depth ratio at the end of the crater's equilibrium lifetime.
            if (!(c.ageGyr >= 0.0 && std::isfinite(c.ageGyr))) {
                throw std::domain_error("Crater age must be finite and not negative");
            }
            const Morphometry2 f{fresh.depthRatio, fresh.rimHeightRatio};
            const double cut = tables->AgeAtDepthRatio(f, m.cutoffDepthRatio);
            const double kappaT = 1.0 * (cut * c.diameter * c.diameter * c.ageGyr /
                                         surface->EquilibriumLifetimeGyr(c.diameter));
            if (!(kappaT >= 0.0 && std::isfinite(kappaT))) {
                throw std::domain_error(
                    "A crater's diffusion age must be finite and not negative");
            }
            tau = kappaT / (c.diameter * c.diameter);
            const Morphometry2 worn = tables->MorphometryAt(f, tau);
            current = {worn.depthRatio, worn.rimHeightRatio};
        } else {
            // FreshnessDegradation::State.
            const double floor = std::fmin(m.cutoffDepthRatio / fresh.depthRatio, 1.0);
            const double erosion =
                floor + (1.0 - floor) * m.ages.Remaining(c.diameter, c.ageGyr);
            current = {erosion * fresh.depthRatio, erosion * fresh.rimHeightRatio};
        }
        // CheckProfileState (crater_kernel.cpp:172-184): a crater the MPD law shapes directly,
        // in the fresh state under a diffused profile.
        const Morphometry& shaped = m.diffused ? fresh : current;
        if (m.profile == TG_PROFILE_MPD && !(m.complex_ && c.diameter >= m.transitionDiameter) &&
            !(m.diffused && tau > 0.0) &&
            !(shaped.depthRatio > 0.0 && shaped.rimHeightRatio >= 0.0 &&
              shaped.rimHeightRatio <= fresh.rimHeightRatio)) {
            throw std::invalid_argument(
                "Crater state needs a positive depth and 0 <= rim height <= fresh rim height");
        }
        tg_crater_record r{};
        r.diameter = static_cast<float>(c.diameter);
        r.fresh_depth = static_cast<float>(fresh.depthRatio);
        r.fresh_rim = static_cast<float>(fresh.rimHeightRatio);
        r.current_depth = static_cast<float>(current.depthRatio);
        r.current_rim = static_cast<float>(current.rimHeightRatio);
        r.diffusion = static_cast<float>(tau);
        r.resolution = static_cast<float>(resolution);
        r.ellipticity = 1.0f;
        r.first_harmonic = 2u;
        r.noise_key = HashWord64(c.id, o.noiseSeed);
        r.texture_key = HashWord64(c.id, SubstreamSeed(o.textureSeed, kCraterTextureStream));
        r.x = static_cast<float>(c.x);
        r.x_low = static_cast<float>(c.x - static_cast<double>(r.x));
        r.y = static_cast<float>(c.y);
        r.y_low = static_cast<float>(c.y - static_cast<double>(r.y));
        // AppendKernelOutline (irregularity.cpp:286-318) for harmonic outlines.
        r.ellipticity = static_cast<float>(outline.ellipticity);
        r.heading = static_cast<float>(outline.heading);
        r.harmonic_offset = static_cast<std::uint32_t>(b.data.size());
        r.first_harmonic = static_cast<std::uint32_t>(outline.firstHarmonic);
        r.harmonic_count = static_cast<std::uint32_t>(outline.amplitudes.size());
        for (std::size_t n = 0; n < outline.amplitudes.size(); ++n) {
            b.data.push_back(static_cast<float>(outline.amplitudes[n]));
            b.data.push_back(static_cast<float>(outline.phases[n]));
        }
        r.table_count = 0;
        r.table_offset = static_cast<std::uint32_t>(b.data.size());
        r.mark_count = 0;
        r.mark_offset = static_cast<std::uint32_t>(b.data.size());
        r.terrace_offset = kUnset;
        r.peak_offset = kUnset;
        AppendTerraces(c.id, c.diameter, fresh.depthRatio, fresh.rimHeightRatio,
                       current.depthRatio, current.rimHeightRatio, o.terraces, b.kernel, r,
                       b.data);
        if (m.complex_) {
            AppendPeaks(c, fresh, current, o.peaks, m.transitionDiameter, r, b.data);
        }
        if (m.diffused && tau > 0.0) {
            // ProfileTables::Append (crater_kernel.cpp:224-254): the crater's table, appended
            // the first time a crater uses it.
            const Morphometry2 f{fresh.depthRatio, fresh.rimHeightRatio};
            const auto key = std::make_pair(f.depthRatio, f.rimHeightRatio);
            const auto [at, added] =
                tableOffsets->try_emplace(key, static_cast<std::uint32_t>(b.data.size()));
            if (added) {
                const std::vector<float>& heights = tables->TableFor(f).heights;
                b.data.insert(b.data.end(), heights.begin(), heights.end());
            }
            r.profile_offset = at->second;
        }
        b.records.push_back(r);
        b.cx.push_back(c.x);
        b.cy.push_back(c.y);
        b.reach.push_back(outline.MaxScale() * reachRadii * (0.5 * c.diameter));
        b.diameter.push_back(c.diameter);
    }

    // An offset nothing set: terraces and peaks set theirs only when the crater has them.
    static constexpr std::uint32_t kUnset = 0xFFFFFFFFu;
};

// Craters per chunk of a parallel preparation.
constexpr std::size_t kPrepareChunk = 4096;

}  // namespace

Batch PrepareBatch(const std::vector<tgh::Crater>& craters, const CraterModels& m,
                   const ApplyOptions& o, double spacing, unsigned threads) {
    if (!(spacing > 0.0)) {
        throw std::invalid_argument("Sample spacing must be positive");
    }
    Batch b;
    b.kernel = MakeShapeKernel(m, o);
    // The default models' steady state and diffusion tables (ProfileTables::AppendAges puts the
    // tables' ages first in the data, crater_kernel.cpp:215-222).
    std::unique_ptr<LunarSurface> surface;
    DiffusionTables tables;
    std::map<std::pair<double, double>, std::uint32_t> tableOffsets;
    if (m.diffused) {
        surface = std::make_unique<LunarSurface>();
        for (const double age : DiffusionAges()) {
            b.data.push_back(static_cast<float>(age));
        }
    }
    const CraterPrep prep{m, o, spacing, CraterReachRadii(m, o), surface.get(), &tables,
                          &tableOffsets};
    std::vector<std::size_t> order;
    for (const std::size_t k : ApplyOrder(craters, threads)) {
        if (prep.Resolution(craters[k]) != 0.0) {
            order.push_back(k);
        }
    }
    const auto finish = [](tg_crater_record& r, std::uint32_t base) {
        r.harmonic_offset += base;
        r.table_offset += base;
        r.mark_offset += base;
        r.terrace_offset = r.terrace_offset == CraterPrep::kUnset ? 0u : r.terrace_offset + base;
        r.peak_offset = r.peak_offset == CraterPrep::kUnset ? 0u : r.peak_offset + base;
    };
    // The default models share diffusion tables in first-use order: one crater after another.
    // Otherwise each crater's record and data depend on it alone, so chunks of the apply order
    // are prepared on threads and joined in order, their data offsets moved past what precedes.
    if (m.diffused || order.size() <= kPrepareChunk || threads == 1) {
        for (const std::size_t k : order) {
            prep.Append(craters[k], prep.Resolution(craters[k]), b);
            finish(b.records.back(), 0u);
        }
    } else {
        const std::size_t chunks = (order.size() + kPrepareChunk - 1) / kPrepareChunk;
        std::vector<Batch> parts(chunks);
        ForEachIndex(chunks, threads, [&](std::size_t n) {
            Batch& part = parts[n];
            part.kernel = b.kernel;
            const std::size_t end = std::min(order.size(), (n + 1) * kPrepareChunk);
            for (std::size_t i = n * kPrepareChunk; i < end; ++i) {
                const tgh::Crater& c = craters[order[i]];
                prep.Append(c, prep.Resolution(c), part);
            }
        });
        // Each part's place in the joined batch, then the parts copied there on threads.
        std::vector<std::size_t> recordAt(chunks + 1, 0), dataAt(chunks + 1, b.data.size());
        for (std::size_t n = 0; n < chunks; ++n) {
            recordAt[n + 1] = recordAt[n] + parts[n].records.size();
            dataAt[n + 1] = dataAt[n] + parts[n].data.size();
        }
        if (dataAt[chunks] > std::numeric_limits<std::uint32_t>::max()) {
            throw std::length_error("crater data exceeds 2^32 floats");
        }
        b.records.resize(recordAt[chunks]);
        b.data.resize(dataAt[chunks]);
        for (std::vector<double>* column : {&b.cx, &b.cy, &b.reach, &b.diameter}) {
            column->resize(recordAt[chunks]);
        }
        ForEachIndex(chunks, threads, [&](std::size_t n) {
            Batch& part = parts[n];
            const auto base = static_cast<std::uint32_t>(dataAt[n]);
            const auto at = static_cast<std::ptrdiff_t>(recordAt[n]);
            for (tg_crater_record& r : part.records) {
                finish(r, base);
            }
            std::copy(part.records.begin(), part.records.end(), b.records.begin() + at);
            std::copy(part.data.begin(), part.data.end(),
                      b.data.begin() + static_cast<std::ptrdiff_t>(dataAt[n]));
            std::copy(part.cx.begin(), part.cx.end(), b.cx.begin() + at);
            std::copy(part.cy.begin(), part.cy.end(), b.cy.begin() + at);
            std::copy(part.reach.begin(), part.reach.end(), b.reach.begin() + at);
            std::copy(part.diameter.begin(), part.diameter.end(), b.diameter.begin() + at);
            part = Batch{};
        });
    }
    if (b.data.empty()) {
        b.data.push_back(0.0f);
    }
    return b;
}

}  // namespace tgh

extern "C" {

int tgx_prepare_craters(const tgx_crater* craters, size_t n, const tgx_crater_models* models,
                        const tgx_apply* apply, uint32_t layer_seed, double spacing,
                        tg_crater_shape* kernel, tg_crater_record** records, size_t* record_count,
                        float** data, size_t* data_count, double** footprint,
                        char* err, size_t errLen) {
    return tgh::Guard(err, errLen, [&] {
        if (models == nullptr || apply == nullptr || kernel == nullptr || records == nullptr ||
            data == nullptr || (n > 0 && craters == nullptr)) {
            throw std::invalid_argument("tgx_prepare_craters: null argument");
        }
        std::vector<tgh::Crater> list(n);
        for (std::size_t i = 0; i < n; ++i) {
            list[i] = {craters[i].x, craters[i].y, craters[i].diameter, craters[i].age_gyr,
                       craters[i].id};
        }
        const tgh::CraterModels m = tgh::ModelsFrom(*models);
        const tgh::ApplyOptions o = tgh::SeededApply(tgh::ApplyFrom(*apply), layer_seed);
        const tgh::Batch b = tgh::PrepareBatch(list, m, o, spacing, 0);
        *kernel = b.kernel;
        *records = tgh::Export(b.records);
        *record_count = b.records.size();
        *data = tgh::Export(b.data);
        *data_count = b.data.size();
        if (footprint != nullptr) {
            std::vector<double> f;
            f.reserve(4 * b.records.size());
            for (std::size_t i = 0; i < b.records.size(); ++i) {
                f.push_back(b.cx[i]);
                f.push_back(b.cy[i]);
                f.push_back(b.reach[i]);
                f.push_back(b.diameter[i]);
            }
            *footprint = tgh::Export(f);
        }
    });
}

double tgx_margin_radii(const tgx_crater_models* models, const tgx_apply* apply) {
    const tgh::CraterModels m = tgh::ModelsFrom(*models);
    const tgh::ApplyOptions o = tgh::ApplyFrom(*apply);
    return tgh::CraterReachRadii(m, o) * tgh::MaxOutlineScale(o.irregularity);
}

}  // extern "C"
