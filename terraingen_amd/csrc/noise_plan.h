// SPDX-License-Identifier: BSD-3-Clause
//
// The noise plan: a fractal stack with every cell-invariant quantity of its octave loops worked
// out once on the host, in FP32 and in the reference's order of operations, so kernels read them
// instead of recomputing them per cell. For each octave the reference derives, per cell
// (include/terraingen/noise/fractal.h:219-239):
//   seed, offsetX, offsetY  from Pcg3dHash(OctaveKey(seed, octave, 0))          (hash.h:40-62)
//   frequency               1 / wavelength times lacunarity per octave (FP32)
//   turnCos, turnSin        the golden-angle recurrence in FP32 (fractal.h:142-144,234-236)
//   weight, totalWeight     gain^i and their running sum in FP32 (fractal.h:194-196)
// Each is a pure function of the fractal's parameters, so the plan holds the same bits the
// reference computes in every cell. Warp fields (fractals.h:98-104) get their own octave tables.
#pragma once

#include "tg_layouts.h"

#include <cstdint>

namespace tg {

// One octave: 8 words.
struct OctavePlan {
    uint32_t seed;      // the octave's lattice seed, Pcg3dHash(key).x
    uint32_t seedLcg;   // seed * 1664525 + 1013904223, the first step of the corner hash
    float offsetX;      // HashToUnitFloat(Pcg3dHash(key).y)
    float offsetY;      // HashToUnitFloat(Pcg3dHash(key).z)
    float frequency;
    float turnCos;
    float turnSin;
    float weight;       // gain^octave as the FP32 running product
};
static_assert(sizeof(OctavePlan) == 32, "OctavePlan is 8 words");

// The octave loop of one fractal (its own or a warp field's).
struct FractalLoopPlan {
    uint32_t octaves;
    uint32_t basis;
    uint32_t type;
    float amplitude;    // FinishFractal's amplitude (1 for a stack member's unit value)
    float offset;       // ridge offset
    float totalWeight;  // sum of the weights, accumulated in octave order in FP32
    float gain;
    uint32_t pad;
    OctavePlan octave[TG_MAX_OCTAVES];
};

// One octave of a fractal on a body (Fractal3, fractal.h:284-317), in FP64 where the reference is:
// the octave's turn (the golden angle about (1, 2, 3), composed once per octave) and frequency.
struct BodyOctave {
    double turn[9];      // row-major
    double frequency;    // 1 / wavelength times lacunarity per octave
    double offsetX;      // HashToUnitFloat(Pcg3dHash(seed, octave, 0).y), widened exactly
    double offsetY;      // HashToUnitFloat(Pcg3dHash(seed, octave, 0).z)
    double offsetZ;      // HashToUnitFloat(Pcg3dHash(seed, octave, 1).x)
    uint32_t zKey;       // PcgHash(Pcg3dHash(seed, octave, 0).x): PerlinLatticeKey3's z term
    uint32_t pad;
};
static_assert(sizeof(BodyOctave) == 112, "BodyOctave is 14 doubles");

// One enabled fractal of a stack.
struct FractalPlan {
    uint32_t blend;
    float amplitude;      // scales the unit value into meters
    uint32_t warped;      // 1 when warpAmplitude > 0
    uint32_t pad;
    FractalLoopPlan unit;   // the fractal at amplitude 1
    FractalLoopPlan warpX;  // first warp field
    FractalLoopPlan warpY;  // second warp field
};

struct NoisePlan {
    uint32_t count;  // enabled fractals, at most TG_MAX_FRACTALS
    uint32_t maxOctaves;
    uint32_t allPerlinFbm;  // 1 when every loop is Perlin fBm without warp (the fast path)
    uint32_t bodyPerlin;    // 1 when every fractal is an unwarped Perlin one
    FractalPlan fractal[TG_MAX_FRACTALS];
    BodyOctave body[TG_MAX_FRACTALS][TG_MAX_OCTAVES];  // each fractal's octaves on a body
    // The octaves on a body of a warped fractal's three warp fields (DomainWarp3,
    // fractals.h:221-230); their sums and weights are warpX's.
    BodyOctave warpBody[TG_MAX_FRACTALS][3][4];
};

// Pcg3dHash of (seed, octave, 0): include/terraingen/noise/hash.h:40-57, lattice.h:126-129.
inline void HostPcg3(uint32_t v[3]) {
    for (int i = 0; i < 3; ++i) {
        v[i] = v[i] * 1664525u + 1013904223u;
    }
    v[0] += v[1] * v[2];
    v[1] += v[2] * v[0];
    v[2] += v[0] * v[1];
    for (int i = 0; i < 3; ++i) {
        v[i] ^= v[i] >> 16u;
    }
    v[0] += v[1] * v[2];
    v[1] += v[2] * v[0];
    v[2] += v[0] * v[1];
}

inline uint32_t HostPcg(uint32_t v) {
    const uint32_t state = v * 747796405u + 2891336453u;
    const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

// Fills `loop` for a fractal; must be built with FP contraction off, as the reference is.
inline void PlanLoop(FractalLoopPlan& loop, uint32_t seed, float wavelength, float amplitude,
                     int octaves, float lacunarity, float gain, uint32_t basis, uint32_t type,
                     float offset) {
    const float kTurnCos = -0.7373688780783197f;
    const float kTurnSin = 0.6754902942615238f;
    loop.octaves = uint32_t(octaves < 0 ? 0 : (octaves > TG_MAX_OCTAVES ? TG_MAX_OCTAVES : octaves));
    loop.basis = basis;
    loop.type = type;
    loop.amplitude = amplitude;
    loop.offset = offset;
    loop.gain = gain;
    loop.pad = 0;
    float frequency = 1.0f / wavelength;
    float c = 1.0f, s = 0.0f, weight = 1.0f, total = 0.0f;
    for (uint32_t o = 0; o < loop.octaves; ++o) {
        uint32_t h[3] = {seed, o, 0u};
        HostPcg3(h);
        OctavePlan& q = loop.octave[o];
        q.seed = h[0];
        q.seedLcg = h[0] * 1664525u + 1013904223u;
        q.offsetX = float(h[1] >> 8u) * 0x1p-24f;
        q.offsetY = float(h[2] >> 8u) * 0x1p-24f;
        q.frequency = frequency;
        q.turnCos = c;
        q.turnSin = s;
        q.weight = weight;
        total += weight;
        weight *= gain;
        frequency *= lacunarity;
        const float nc = c * kTurnCos - s * kTurnSin;
        s = s * kTurnCos + c * kTurnSin;
        c = nc;
    }
    loop.totalWeight = total;
}

// The plan of a noise-layer kernel: its enabled fractals in order (noise_layer.h:41-50 keeps only
// enabled ones; a disabled one leaves the stack unchanged, so skipping it is exact).
// Fractal3's per-octave constants (fractal.h:284-317), with the reference's FP64 recurrences;
// must be built with FP contraction off, as the reference is.
inline void PlanBody(BodyOctave* octaves, uint32_t count, uint32_t seed, float wavelength,
                     float lacunarity) {
    static constexpr double kTurn[3][3] = {
        {-0.6132711010727254, -0.293401570821183, 0.7333580809050306},
        {0.7897926788435601, -0.24097777005594262, 0.5640542870894418},
        {0.011228581128535142, 0.9251190369776896, 0.37951111497202883}};
    double frequency = 1.0 / static_cast<double>(wavelength);
    double turn[3][3] = {{1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}};
    for (uint32_t o = 0; o < count; ++o) {
        uint32_t h[3] = {seed, o, 0u};
        uint32_t g[3] = {seed, o, 1u};
        HostPcg3(h);
        HostPcg3(g);
        BodyOctave& b = octaves[o];
        for (int r = 0; r < 3; ++r) {
            for (int c = 0; c < 3; ++c) {
                b.turn[3 * r + c] = turn[r][c];
            }
        }
        b.frequency = frequency;
        b.offsetX = double(float(h[1] >> 8u) * 0x1p-24f);
        b.offsetY = double(float(h[2] >> 8u) * 0x1p-24f);
        b.offsetZ = double(float(g[0] >> 8u) * 0x1p-24f);
        b.zKey = HostPcg(h[0]);
        b.pad = 0u;
        frequency *= static_cast<double>(lacunarity);
        double next[3][3] = {};
        for (int r = 0; r < 3; ++r) {
            for (int c = 0; c < 3; ++c) {
                next[r][c] = kTurn[r][0] * turn[0][c] + kTurn[r][1] * turn[1][c] +
                             kTurn[r][2] * turn[2][c];
            }
        }
        for (int r = 0; r < 3; ++r) {
            for (int c = 0; c < 3; ++c) {
                turn[r][c] = next[r][c];
            }
        }
    }
}

inline void MakeNoisePlan(const tg_noise_layer_kernel& k, NoisePlan& plan) {
    plan.count = 0;
    plan.maxOctaves = 0;
    plan.allPerlinFbm = 1;
    plan.bodyPerlin = 1;
    const uint32_t n = k.count < TG_MAX_FRACTALS ? k.count : TG_MAX_FRACTALS;
    for (uint32_t i = 0; i < n; ++i) {
        const tg_fractal& f = k.fractals[i];
        if (!f.enabled) {
            continue;
        }
        FractalPlan& p = plan.fractal[plan.count++];
        p.blend = f.blend;
        p.amplitude = f.amplitude;
        p.warped = f.warp_amplitude > 0.0f ? 1u : 0u;
        p.pad = 0;
        PlanLoop(p.unit, f.seed, f.wavelength, 1.0f, f.octaves, f.lacunarity, f.gain, f.basis,
                 f.type, f.offset);
        if (p.warped) {
            // WarpParams: 4-octave Perlin fBm, lacunarity 2, gain 0.5 (fractals.h:98-104).
            const uint32_t seedX = HostPcg(f.seed ^ 0x5bd1e995u);
            const uint32_t seedY = HostPcg(seedX);
            PlanLoop(p.warpX, seedX, f.warp_wavelength, f.warp_amplitude, 4, 2.0f, 0.5f,
                     TG_BASIS_PERLIN, TG_FRACTAL_FBM, 1.0f);
            PlanLoop(p.warpY, seedY, f.warp_wavelength, f.warp_amplitude, 4, 2.0f, 0.5f,
                     TG_BASIS_PERLIN, TG_FRACTAL_FBM, 1.0f);
            plan.allPerlinFbm = 0;
        } else {
            p.warpX.octaves = 0;
            p.warpY.octaves = 0;
        }
        if (p.unit.basis != TG_BASIS_PERLIN || p.unit.type != TG_FRACTAL_FBM) {
            plan.allPerlinFbm = 0;
        }
        if (p.warped || p.unit.basis != TG_BASIS_PERLIN) {
            plan.bodyPerlin = 0;
        }
        PlanBody(plan.body[plan.count - 1], p.unit.octaves, f.seed, f.wavelength, f.lacunarity);
        if (p.warped) {
            // Each further 3D warp field is seeded with PcgHash of the one before.
            uint32_t seed = HostPcg(f.seed ^ 0x5bd1e995u);
            for (uint32_t w = 0; w < 3u; ++w) {
                PlanBody(plan.warpBody[plan.count - 1][w], 4u, seed, f.warp_wavelength, 2.0f);
                seed = HostPcg(seed);
            }
        }
        if (p.unit.octaves > plan.maxOctaves) {
            plan.maxOctaves = p.unit.octaves;
        }
    }
}

// A bound on the magnitude of every lattice coordinate the plan's body noise reaches at points
// within `radius` of the origin: the turns are rotations, the offsets below 1 and the floor one
// more, and a warp moves a point by at most its amplitude along each axis (so by less than twice
// it).
inline double BodyLatticeBound(const NoisePlan& plan, double radius) {
    double bound = 0.0;
    for (uint32_t i = 0; i < plan.count; ++i) {
        const FractalPlan& f = plan.fractal[i];
        const double reach = f.warped ? radius + 2.0 * double(f.warpX.amplitude) : radius;
        for (uint32_t o = 0; o < f.unit.octaves; ++o) {
            const double c = reach * plan.body[i][o].frequency;
            bound = c > bound ? c : bound;
        }
        if (f.warped) {
            for (uint32_t o = 0; o < 4u; ++o) {
                const double c = radius * plan.warpBody[i][0][o].frequency;
                bound = c > bound ? c : bound;
            }
        }
    }
    return bound * (1.0 + 1e-9) + 2.0;
}

}  // namespace tg
