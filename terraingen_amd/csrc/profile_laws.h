// SPDX-License-Identifier: BSD-3-Clause
//
// The closed-form crater profile laws in FP32, one source for the kernels (hipcc, host and device)
// and for the host's crater preparation (g++), as the reference keeps one kernel header for its
// GPU kernels and its CPU path (include/terraingen/layers/kernels/craters.h:306-503,975-981):
// ejecta scale, Fassett-Thomson, Martin-Parkes-Dunstan, the complex profile with its central peak
// dome, and the terrace ease. The host samples walls and initial diffusion profiles through them
// (CraterProfile::Height evaluates the same FP32 laws on the CPU), so the same operations in the
// same order give the reference's bits there; on the device they are the per-pair arithmetic.
#pragma once

#include "tg_layouts.h"

#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define TG_LAW __host__ __device__ __forceinline__
#else
#define TG_LAW inline
#endif

namespace tg {

constexpr float kLawPi = 3.14159265358979f;

TG_LAW float LawMax(float a, float b) {
    return a < b ? b : a;
}

TG_LAW float LawMin(float a, float b) {
    return b < a ? b : a;
}

// KernelEjectaScale (craters.h:306-311).
TG_LAW float EjectaScale(float radius) {
    return 0.14f * powf(radius, 0.74f);
}

// FassettThomsonExterior / FassettThomsonShape (craters.h:326-362).
TG_LAW float FassettThomsonExterior(float x, float start) {
    const float polynomial =
        0.188253307f + x * (-0.187050452f + x * (0.01844746f + x * 0.01505647f));
    const float u = (x - start) / (1.5f - start);
    return polynomial * (1.0f - u * u);
}

TG_LAW float FassettThomsonShape(float x, uint32_t rim, float freshRim,
                                 float currentDepth) {
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
            h = FassettThomsonExterior(x, 1.02f);
        }
    } else {
        const float edge = rim == 0u ? 0.98f : 1.0f;
        if (x <= 0.2f) {
            h = -0.181f;
        } else if (x <= edge) {
            h = interior;
        } else {
            h = FassettThomsonExterior(x, edge);
        }
    }
    return LawMax(h, freshRim - currentDepth);
}

struct ProfilePoint {
    float height, slope;
};

// MartinParkesDunstanExterior (craters.h:374-386).
TG_LAW ProfilePoint MpdExterior(float x, float rim, float ejectaScale) {
    const float offset = rim - ejectaScale;
    const float gradient = -ejectaScale / 8.0f - offset;
    const float r = 1.0f / (x + 1.0f);
    const float law = ejectaScale * r * r * r + offset + gradient * x;
    const float lawSlope = -3.0f * ejectaScale * r * r * r * r + gradient;
    const float taper = 1.0f - x * x;
    ProfilePoint p;
    p.height = law * taper;
    p.slope = lawSlope * taper - 2.0f * x * law;
    return p;
}

// The record fields a profile reads; a diffused crater's fresh state swaps the current ones.
struct ProfileState {
    float diameter, freshDepth, freshRim, currentDepth, currentRim;
};

// MartinParkesDunstanHeight (craters.h:388-416).
TG_LAW float MpdHeight(float distance, float diameter, const ProfileState& c) {
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
    const float beta = LawMin(1.0f, 1.5f * (freshRim - rim) / k);
    if (x <= beta) {
        const ProfilePoint join = MpdExterior(beta, rim, ejectaScale);
        const float drop = rim - join.height;
        const float slope = drop > 0.0f ? LawMax(join.slope, -3.0f * drop / beta) : join.slope;
        const float t = x / beta;
        return rim + (join.height - rim) * t * t * (3.0f - 2.0f * t) +
               beta * slope * t * t * (t - 1.0f);
    }
    return MpdExterior(x, rim, ejectaScale).height;
}

// SimpleProfileHeight (craters.h:418-427).
TG_LAW float SimpleProfileHeight(const tg_crater_profile& p, float distance,
                                 float diameter, const ProfileState& c) {
    if (p.simple == TG_PROFILE_MPD) {
        return MpdHeight(distance, diameter, c);
    }
    return FassettThomsonShape(distance / (0.5f * diameter), p.rim, c.freshRim, c.currentDepth) *
           diameter;
}

// KilometerLaw (craters.h:429-433).
TG_LAW float KilometerLaw(float coefficient, float exponent, float diameter,
                          bool volume) {
    const float kilometers = coefficient * powf(diameter / 1000.0f, exponent);
    return volume ? kilometers * 1e9f : kilometers * 1000.0f;
}

struct PeakSize {
    float height, base;
};

// KernelComplexPeak (craters.h:441-456).
TG_LAW PeakSize ComplexPeak(const tg_crater_profile& p, float diameter) {
    const float fitted =
        p.peak_fit_max_diameter > 0.0f ? LawMin(diameter, p.peak_fit_max_diameter) : diameter;
    PeakSize size;
    size.height = KilometerLaw(p.peak_height_coefficient, p.peak_height_exponent, fitted, false);
    if (p.peak_from_volume == 1u) {
        const float volume =
            KilometerLaw(p.peak_width_coefficient, p.peak_width_exponent, fitted, true);
        size.base = sqrtf(3.0f * volume / (kLawPi * size.height));
    } else {
        size.base =
            0.5f * KilometerLaw(p.peak_width_coefficient, p.peak_width_exponent, fitted, false);
    }
    size.base *= diameter / fitted;
    return size;
}

// KernelPeakDome (craters.h:458-465).
TG_LAW float PeakDome(PeakSize size, float distance) {
    if (!(distance < size.base)) {
        return 0.0f;
    }
    const float u = 1.0f - (distance / size.base) * (distance / size.base);
    return size.height * u * u;
}

// KernelKeptDepth (craters.h:467-470).
TG_LAW float KeptDepth(float freshDepth, float currentDepth) {
    return freshDepth > 0.0f ? currentDepth / freshDepth : 1.0f;
}

// ComplexProfileHeight (craters.h:472-494).
TG_LAW float ComplexProfileHeight(const tg_crater_profile& p, float distance,
                                  float diameter, const ProfileState& c) {
    const float radius = 0.5f * diameter;
    const float x = distance / radius;
    const float depth = c.currentDepth * diameter;
    const float rim = c.currentRim * diameter;
    if (x >= 1.0f) {
        if (x >= p.complex_reach_radii) {
            return 0.0f;
        }
        const float edge = powf(p.complex_reach_radii, -p.ejecta_exponent);
        return rim * (powf(x, -p.ejecta_exponent) - edge) / (1.0f - edge);
    }
    const float floorShare = LawMin(
        0.9f, KilometerLaw(p.floor_coefficient, p.floor_exponent, diameter, false) / diameter);
    float height = rim - depth;
    if (x > floorShare) {
        const float t = (x - floorShare) / (1.0f - floorShare);
        height += depth * t * t;
    }
    return height +
           KeptDepth(c.freshDepth, c.currentDepth) * PeakDome(ComplexPeak(p, diameter), distance);
}

// FreshProfileHeight (craters.h:496-503).
TG_LAW float FreshProfileHeight(const tg_crater_profile& p, float distance,
                                float diameter, const ProfileState& c) {
    if (p.complex_ == 1u && diameter >= p.transition_diameter) {
        return ComplexProfileHeight(p, distance, diameter, c);
    }
    return SimpleProfileHeight(p, distance, diameter, c);
}

// KernelTerraceEase (craters.h:975-981).
TG_LAW float TerraceEase(float v, float rounding) {
    const float steepest = 1.0f / (1.0f - 0.5f * rounding);
    if (v < rounding) {
        return steepest * v * v / (2.0f * rounding);
    }
    return steepest * (v - 0.5f * rounding);
}

}  // namespace tg
