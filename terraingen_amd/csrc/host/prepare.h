// SPDX-License-Identifier: BSD-3-Clause
//
// Crater preparation (prepare.cpp): models, apply options and the kernel batch.
#pragma once

#include "placement.h"
#include "terraingen_amd.h"

#include <cstdint>
#include <vector>

namespace tgh {

// CraterIrregularity (include/terraingen/layers/craters/shaping/irregularity.h:34-70) with its
// defaults, power-law rim model.
struct Irregularity {
    bool enabled = false;
    std::uint32_t seed = 0;
    double roughShare = 1.0;
    double roughnessLow = 0.01, roughnessHigh = 0.03;
    double spectralSlope = 1.5;
    int minHarmonic = 2, maxHarmonic = 12;
    double elongatedShare = 0.054;
    double ellipticityLow = 1.1, ellipticityHigh = 1.5;
};

// CraterPeaks (apply_options.h:241-262) with its defaults.
struct Peaks {
    bool enabled = false;
    std::uint32_t seed = 0;
    int minMassifs = 1, maxMassifs = 4;
    double spreadRadii = 0.8, elongation = 2.0, chainSpreadDegrees = 70.0, sharpness = 0.6,
           lopsidedness = 0.5, heightVariation = 0.5, roughness = 0.35,
           roughnessWavelength = 0.35, warp = 0.12, heightScale = 1.0, widthScale = 1.0;
};

// CraterTerraces (apply_options.h:184-231) with its defaults.
struct Terraces {
    bool enabled = false;
    std::uint32_t seed = 0;
    double onsetDiameter = 20000.0, fullDiameter = 65000.0;
    int minBenches = 1, maxBenches = 4;
    double benchSlopeDegrees = 3.0, scarpShare = 0.25, scarpSlopeDegrees = 35.0, arcShare = 0.6,
           rimVariation = 0.08, blockOffset = 0.6, blockGaps = 0.25, minFreshness = 0.3;
};

struct ApplyOptions {
    Irregularity irregularity;
    Peaks peaks;
    Terraces terraces;
    std::uint32_t noiseSeed = 0, textureSeed = 0;
    double rimSigma = 0.0, noiseAmplitude = 0.0, noiseRimAmplitude = 0.0, wallStreaks = 0.0,
           ejecta = 0.0, mantling = 0.0, ejectaReachRadii = 3.0;
};

struct CraterModels {
    std::uint32_t profile = TG_PROFILE_MPD;  // the simple profile: MPD or Fassett-Thomson
    bool complex_ = false;
    double transitionDiameter = 15000.0;
    double cutoffDepthRatio = 0.04;  // kVisibilityDepthRatio (degradation/state.h:13)
    Ages ages;
    // DefaultCraterModels: the MPD profile diffused in time (DiffusedProfile) under matched
    // diffusion over the lunar steady-state ages, in place of freshness degradation.
    bool diffused = false;
};

struct Batch {
    tg_crater_shape kernel{};
    std::vector<tg_crater_record> records;
    std::vector<float> data;
    std::vector<double> cx, cy, reach, diameter;  // footprint: centre, half-width; D (m)
};

inline CraterModels ModelsFrom(const tgx_crater_models& m) {
    CraterModels out;
    if (m.model > 1u) {
        throw std::invalid_argument("unknown crater model");
    }
    if (m.model == 1u) {
        out.profile = TG_PROFILE_MPD;
        out.diffused = true;
        return out;
    }
    out.profile = m.profile;
    out.complex_ = m.complex_ != 0;
    out.transitionDiameter = m.transition_diameter;
    out.ages.lifetimeGyr = m.lifetime_gyr;
    if (m.profile != TG_PROFILE_MPD && m.profile != TG_PROFILE_FASSETT_THOMSON) {
        throw std::invalid_argument("only Martin-Parkes-Dunstan and Fassett-Thomson profiles");
    }
    return out;
}

inline ApplyOptions ApplyFrom(const tgx_apply& a) {
    ApplyOptions out;
    out.irregularity.enabled = a.irregular != 0;
    out.peaks.enabled = a.peaks != 0;
    out.rimSigma = a.rim_sigma;
    out.noiseAmplitude = a.noise_amplitude;
    out.noiseRimAmplitude = a.noise_rim_amplitude;
    out.wallStreaks = a.wall_streaks;
    out.ejecta = a.ejecta;
    out.terraces.enabled = a.terraces != 0;
    if (a.terrace_onset > 0.0 || a.terrace_full > 0.0) {
        if (!(a.terrace_onset > 0.0 && a.terrace_full > a.terrace_onset)) {
            throw std::invalid_argument("terraces need 0 < onset diameter < full diameter");
        }
        out.terraces.onsetDiameter = a.terrace_onset;
        out.terraces.fullDiameter = a.terrace_full;
    }
    return out;
}

tg_crater_shape MakeShapeKernel(const CraterModels& m, const ApplyOptions& o);
ApplyOptions SeededApply(ApplyOptions o, std::uint32_t layerSeed);
double CraterReachRadii(const CraterModels& m, const ApplyOptions& o);
double MaxOutlineScale(const Irregularity& irr);
// The craters prepared in apply order (a crater the spacing does not resolve is dropped), on
// `threads` threads (0: every hardware thread; the batch is the same on any).
Batch PrepareBatch(const std::vector<Crater>& craters, const CraterModels& m,
                   const ApplyOptions& o, double spacing, unsigned threads = 0);

// CraterProfile::Height of the kernel's closed-form profile (terraces.cpp).
double ProfileHeight(const tg_crater_shape& k, double distance, double diameter,
                     double freshDepth, double freshRim, double currentDepth, double currentRim);
// AppendKernelTerraces (crater_kernel.cpp:312-459) for crater `id` (terraces.cpp).
void AppendTerraces(std::uint64_t id, double diameter, double freshDepth, double freshRim,
                    double currentDepth, double currentRim, const Terraces& terraces,
                    const tg_crater_shape& kernel, tg_crater_record& record,
                    std::vector<float>& data);

}  // namespace tgh
