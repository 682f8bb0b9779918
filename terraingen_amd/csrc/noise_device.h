// SPDX-License-Identifier: BSD-3-Clause
//
// 2D noise for gfx950, one lane per cell, in VALU. Every floating-point operation is the
// reference's, in its order, and the library is built with -ffp-contract=off, so a cell's noise
// has the same bits as the reference's CPU backend (include/terraingen/noise/{hash,lattice,
// perlin,simplex,value,worley,fractal,fractals}.h). Cell-invariant octave quantities come from a
// NoisePlan (noise_plan.h).
//
// The Perlin corner hash is the x word of pcg3d(x, y, seed) (perlin.h:31-33, hash.h:40-57). Its
// first LCG step is linear in the corner, so the four corners of a cell share work exactly in
// 32-bit wrap-around arithmetic: lcg(x + 1) = lcg(x) + 1664525, the seed's step is per octave
// (NoisePlan), and the first cross product y * z depends only on the corner's row.
#pragma once

#include "noise_plan.h"

#include <hip/hip_runtime.h>

namespace tg {

constexpr uint32_t kLcgMul = 1664525u;
constexpr uint32_t kLcgAdd = 1013904223u;

struct U3 {
    uint32_t x, y, z;
};

__device__ __forceinline__ uint32_t Pcg(uint32_t v) {
    const uint32_t state = v * 747796405u + 2891336453u;
    const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

__device__ __forceinline__ U3 Pcg3(U3 v) {
    v.x = v.x * kLcgMul + kLcgAdd;
    v.y = v.y * kLcgMul + kLcgAdd;
    v.z = v.z * kLcgMul + kLcgAdd;
    v.x += v.y * v.z;
    v.y += v.z * v.x;
    v.z += v.x * v.y;
    v.x ^= v.x >> 16u;
    v.y ^= v.y >> 16u;
    v.z ^= v.z >> 16u;
    v.x += v.y * v.z;
    v.y += v.z * v.x;
    v.z += v.x * v.y;
    return v;
}

__device__ __forceinline__ float UnitFloat(uint32_t h) {
    return float(h >> 8u) * 0x1p-24f;
}

// std::max, std::min and std::clamp as the reference's C++ evaluates them, so NaN and signed zeros
// come out as on the CPU (fmaxf/fminf may pick either zero).
__device__ __forceinline__ float StdMax(float a, float b) {
    return a < b ? b : a;
}
__device__ __forceinline__ float StdMin(float a, float b) {
    return b < a ? b : a;
}
__device__ __forceinline__ float StdClamp(float v, float lo, float hi) {
    return v < lo ? lo : (hi < v ? hi : v);
}
__device__ __forceinline__ float Clamp01(float v) {
    return StdClamp(v, 0.0f, 1.0f);
}

// The x word of pcg3d from the corner's LCG'd x and y, the seed's LCG'd word and yz = y' * z'.
__device__ __forceinline__ uint32_t CornerX(uint32_t ax, uint32_t ay, uint32_t az, uint32_t yz) {
    uint32_t vx = ax + yz;
    uint32_t vy = ay + az * vx;
    uint32_t vz = az + vx * vy;
    vx ^= vx >> 16u;
    vy ^= vy >> 16u;
    vz ^= vz >> 16u;
    return vx + vy * vz;
}

// PerlinGradient2 (perlin.h:17-26). Both cases are computed and one is picked with a bitwise
// select, so a wave whose lanes hold different hashes does not split into two branches; negation
// is the sign-bit flip that unary minus is, so every result has the reference's bits.
__device__ __forceinline__ float Grad2(uint32_t hash, float x, float y) {
    const float kInvSqrt2 = 0.707106781186547524f;
    const uint32_t h = hash & 7u;
    // h < 4: the x or y axis, negated for odd h.
    const float axis = (h & 2u) != 0u ? y : x;
    const uint32_t signedAxis = __float_as_uint(axis) ^ (h << 31u);
    // h >= 4: the diagonals.
    const float sx = h == 5u ? -x : x;
    const float sy = h == 6u ? -y : y;
    const float sum = sx + sy;
    const float diagonal = (h == 7u ? -sum : sum) * kInvSqrt2;
    const uint32_t pickAxis = 0u - uint32_t(h < 4u);
    return __uint_as_float((signedAxis & pickAxis) | (__float_as_uint(diagonal) & ~pickAxis));
}

__device__ __forceinline__ float Fade(float t) {
    return t * t * t * (t * (t * 6.0f - 15.0f) + 10.0f);
}

__device__ __forceinline__ float Mix(float t, float a, float b) {
    return a + t * (b - a);
}

// PerlinNoise2 at lattice point (x, y) (perlin.h:72-98) with the octave's seed and LCG'd seed.
__device__ __forceinline__ float Perlin2(float x, float y, uint32_t seed, uint32_t seedLcg) {
    (void)seed;
    const float cx = floorf(x);
    const float cy = floorf(y);
    const float fx = x - cx;
    const float fy = y - cy;
    const uint32_t ix = uint32_t(int32_t(cx));
    const uint32_t iy = uint32_t(int32_t(cy));
    const uint32_t ax0 = ix * kLcgMul + kLcgAdd;
    const uint32_t ax1 = ax0 + kLcgMul;
    const uint32_t ay0 = iy * kLcgMul + kLcgAdd;
    const uint32_t ay1 = ay0 + kLcgMul;
    const uint32_t yz0 = ay0 * seedLcg;
    const uint32_t yz1 = ay1 * seedLcg;
    const float n00 = Grad2(CornerX(ax0, ay0, seedLcg, yz0), fx, fy);
    const float n10 = Grad2(CornerX(ax1, ay0, seedLcg, yz0), fx - 1.0f, fy);
    const float n01 = Grad2(CornerX(ax0, ay1, seedLcg, yz1), fx, fy - 1.0f);
    const float n11 = Grad2(CornerX(ax1, ay1, seedLcg, yz1), fx - 1.0f, fy - 1.0f);
    const float u = Fade(fx);
    const float v = Fade(fy);
    return 1.41421356237309505f * Mix(v, Mix(u, n00, n10), Mix(u, n01, n11));
}

// Perlin2 for a wave whose 64 lanes are the cells of an 8 x 8 square, every lane active: lane & 7
// the cell's column, lane >> 3 its row, the cell centre's x non-decreasing along the columns and
// its y non-increasing along the rows. The lattice coordinates (x cos - y sin, x sin + y cos)
// times the frequency plus an offset are then monotone along the columns and along the rows (each
// correctly rounded operation is monotone in its operands), so their floors are least and greatest
// at corner lanes picked by the signs of cos and sin. When the square's lattice cells span at most
// 7 x 7, its 8 x 8 lattice corners are hashed once each, one per lane, and every lane gathers its
// four corner hashes with ds_bpermute: a quarter of the hashing, whose 32-bit multiplies are the
// costliest VALU work of the noise. The hashes, and so the noise, are identical to Perlin2's.
// `stepX` and `stepY` are the lane's column and row times the LCG multiplier (SquareSteps).
__device__ __forceinline__ float Perlin2Square(float x, float y, uint32_t seedLcg, float turnCos,
                                               float turnSin, uint32_t stepX, uint32_t stepY) {
    const float cx = floorf(x);
    const float cy = floorf(y);
    const float fx = x - cx;
    const float fy = y - cy;
    const int32_t ix = int32_t(cx);
    const int32_t iy = int32_t(cy);
    // x' grows along the columns when cos >= 0 and along the rows (y falling) when sin >= 0;
    // y' grows along the columns when sin >= 0 and along the rows when cos < 0.
    const uint32_t lowX = (turnCos >= 0.0f ? 0u : 7u) + (turnSin >= 0.0f ? 0u : 56u);
    const uint32_t lowY = (turnSin >= 0.0f ? 0u : 7u) + (turnCos >= 0.0f ? 56u : 0u);
    const int32_t xMin = __builtin_amdgcn_readlane(ix, int(lowX));
    const int32_t xMax = __builtin_amdgcn_readlane(ix, int(63u - lowX));
    const int32_t yMin = __builtin_amdgcn_readlane(iy, int(lowY));
    const int32_t yMax = __builtin_amdgcn_readlane(iy, int(63u - lowY));
    uint32_t h00, h10, h01, h11;
    if (uint32_t(xMax) - uint32_t(xMin) <= 6u && uint32_t(yMax) - uint32_t(yMin) <= 6u) {
        // This lane hashes lattice corner (xMin + lane & 7, yMin + lane >> 3); the LCG step of a
        // corner is linear in it, so its words are the least corner's plus the lane's steps.
        const uint32_t ax = uint32_t(xMin) * kLcgMul + kLcgAdd + stepX;
        const uint32_t ay = uint32_t(yMin) * kLcgMul + kLcgAdd + stepY;
        const int32_t mine = int32_t(CornerX(ax, ay, seedLcg, ay * seedLcg));
        const int32_t at = int32_t(((uint32_t(iy - yMin) << 3u) + uint32_t(ix - xMin)) << 2u);
        h00 = uint32_t(__builtin_amdgcn_ds_bpermute(at, mine));
        h10 = uint32_t(__builtin_amdgcn_ds_bpermute(at + 4, mine));
        h01 = uint32_t(__builtin_amdgcn_ds_bpermute(at + 32, mine));
        h11 = uint32_t(__builtin_amdgcn_ds_bpermute(at + 36, mine));
    } else {
        const uint32_t ax0 = uint32_t(ix) * kLcgMul + kLcgAdd;
        const uint32_t ax1 = ax0 + kLcgMul;
        const uint32_t ay0 = uint32_t(iy) * kLcgMul + kLcgAdd;
        const uint32_t ay1 = ay0 + kLcgMul;
        const uint32_t yz0 = ay0 * seedLcg;
        const uint32_t yz1 = ay1 * seedLcg;
        h00 = CornerX(ax0, ay0, seedLcg, yz0);
        h10 = CornerX(ax1, ay0, seedLcg, yz0);
        h01 = CornerX(ax0, ay1, seedLcg, yz1);
        h11 = CornerX(ax1, ay1, seedLcg, yz1);
    }
    const float n00 = Grad2(h00, fx, fy);
    const float n10 = Grad2(h10, fx - 1.0f, fy);
    const float n01 = Grad2(h01, fx, fy - 1.0f);
    const float n11 = Grad2(h11, fx - 1.0f, fy - 1.0f);
    const float u = Fade(fx);
    const float v = Fade(fy);
    return 1.41421356237309505f * Mix(v, Mix(u, n00, n10), Mix(u, n01, n11));
}

// The lane's column and row in its 8 x 8 square times the LCG multiplier, for Perlin2Square. The
// empty asm hides their origin, so the compiler adds them to the least corner's LCG word (one add
// per coordinate) instead of re-multiplying the corner's index every octave.
struct SquareSteps {
    uint32_t x, y;
};
__device__ __forceinline__ SquareSteps MakeSquareSteps(uint32_t lane) {
    uint32_t x = (lane & 7u) * kLcgMul;
    uint32_t y = (lane >> 3u) * kLcgMul;
    asm("" : "+v"(x), "+v"(y));
    return {x, y};
}

// SimplexNoise2 (simplex.h:18-59).
__device__ __forceinline__ float SimplexCorner(uint32_t ci, uint32_t cj, float dx, float dy,
                                               uint32_t seed) {
    const float t = 0.5f - dx * dx - dy * dy;
    const float t2 = t * t;
    const float corner = t2 * t2 * Grad2(Pcg3({ci, cj, seed}).x, dx, dy);
    return !(t <= 0.0f) ? corner : 0.0f;
}

__device__ __forceinline__ float Simplex2(float x, float y, uint32_t seed) {
    const float kSkew = 0.366025403784438647f;
    const float kUnskew = 0.211324865405187118f;
    const float skew = (x + y) * kSkew;
    const float cellX = floorf(x + skew);
    const float cellY = floorf(y + skew);
    const float unskew = (cellX + cellY) * kUnskew;
    const float x0 = x - (cellX - unskew);
    const float y0 = y - (cellY - unskew);
    const uint32_t stepX = x0 > y0 ? 1u : 0u;
    const uint32_t stepY = 1u - stepX;
    const float x1 = x0 - float(stepX) + kUnskew;
    const float y1 = y0 - float(stepY) + kUnskew;
    const float x2 = x0 - 1.0f + 2.0f * kUnskew;
    const float y2 = y0 - 1.0f + 2.0f * kUnskew;
    const uint32_t i = uint32_t(int32_t(cellX));
    const uint32_t j = uint32_t(int32_t(cellY));
    return 99.0f * (SimplexCorner(i, j, x0, y0, seed) +
                    SimplexCorner(i + stepX, j + stepY, x1, y1, seed) +
                    SimplexCorner(i + 1u, j + 1u, x2, y2, seed));
}

// ValueNoise2 (value.h:19-45).
__device__ __forceinline__ float LatticeValue(uint32_t x, uint32_t y, uint32_t seed) {
    return 2.0f * UnitFloat(Pcg3({x, y, seed}).y) - 1.0f;
}

__device__ __forceinline__ float Value2(float x, float y, uint32_t seed) {
    const float cx = floorf(x);
    const float cy = floorf(y);
    const float fx = x - cx;
    const float fy = y - cy;
    const uint32_t x0 = uint32_t(int32_t(cx));
    const uint32_t y0 = uint32_t(int32_t(cy));
    const float v00 = LatticeValue(x0, y0, seed);
    const float v10 = LatticeValue(x0 + 1u, y0, seed);
    const float v01 = LatticeValue(x0, y0 + 1u, seed);
    const float v11 = LatticeValue(x0 + 1u, y0 + 1u, seed);
    const float u = Fade(fx);
    const float v = Fade(fy);
    return Mix(v, Mix(u, v00, v10), Mix(u, v01, v11));
}

struct Worley {
    float f1, f2, value;
};

// Worley2 (worley.h:38-63).
__device__ __forceinline__ Worley Worley2(float x, float y, uint32_t seed) {
    const float cx = floorf(x);
    const float cy = floorf(y);
    const float fx = x - cx;
    const float fy = y - cy;
    const int32_t ix = int32_t(cx);
    const int32_t iy = int32_t(cy);
    float d1 = 1e30f;
    float d2 = 1e30f;
    uint32_t nearest = 0u;
    for (int32_t dy = -2; dy <= 2; ++dy) {
        for (int32_t dx = -2; dx <= 2; ++dx) {
            const U3 h = Pcg3({uint32_t(ix + dx), uint32_t(iy + dy), seed});
            const float px = float(dx) + UnitFloat(h.y) - fx;
            const float py = float(dy) + UnitFloat(h.z) - fy;
            const float d = px * px + py * py;
            if (d < d1) {
                d2 = d1;
                d1 = d;
                nearest = h.x;
            } else if (d < d2) {
                d2 = d;
            }
        }
    }
    return {sqrtf(d1), sqrtf(d2), 2.0f * UnitFloat(nearest) - 1.0f};
}

// BasisNoise2 (fractal.h:89-111).
__device__ __forceinline__ float Basis2(uint32_t basis, float x, float y, const OctavePlan& q) {
    const float kInvSqrt2 = 0.707106781186547524f;
    const float kInvSqrt5 = 0.447213595499957939f;
    switch (basis) {
    case TG_BASIS_SIMPLEX:
        return Simplex2(x, y, q.seed);
    case TG_BASIS_VALUE:
        return Value2(x, y, q.seed);
    case TG_BASIS_WORLEY_F1:
        return 2.0f * StdMin(Worley2(x, y, q.seed).f1 * kInvSqrt2, 1.0f) - 1.0f;
    case TG_BASIS_WORLEY_F2:
        return 2.0f * StdMin(Worley2(x, y, q.seed).f2 * kInvSqrt5, 1.0f) - 1.0f;
    case TG_BASIS_WORLEY_F2_MINUS_F1: {
        const Worley w = Worley2(x, y, q.seed);
        return 2.0f * StdMin((w.f2 - w.f1) * kInvSqrt5, 1.0f) - 1.0f;
    }
    case TG_BASIS_VORONOI:
        return Worley2(x, y, q.seed).value;
    default:
        return Perlin2(x, y, q.seed, q.seedLcg);
    }
}

// Fractal2 (fractal.h:219-239) over a planned octave loop; AddFractalOctave and FinishFractal
// (fractal.h:164-211) with the weights read from the plan.
// kSquare: the caller's wave is an 8 x 8 square of cells as Perlin2Square requires, and `lane` is
// the lane's index in it.
template <bool kSquare = false>
__device__ __forceinline__ float FractalLoop(const FractalLoopPlan& p, float x, float y,
                                             uint32_t lane = 0u) {
    const uint32_t type = p.type;
    const uint32_t basis = p.basis;
    const uint32_t octaves = p.octaves;
    float sum = 0.0f;
    float feedback = 1.0f;
    if (type == TG_FRACTAL_FBM && basis == TG_BASIS_PERLIN) {
        const SquareSteps steps = MakeSquareSteps(lane);
        for (uint32_t o = 0; o < octaves; ++o) {
            const OctavePlan& q = p.octave[o];
            const float rx = x * q.turnCos - y * q.turnSin;
            const float ry = x * q.turnSin + y * q.turnCos;
            const float px = rx * q.frequency + q.offsetX;
            const float py = ry * q.frequency + q.offsetY;
            float n;
            if constexpr (kSquare) {
                n = Perlin2Square(px, py, q.seedLcg, q.turnCos, q.turnSin, steps.x, steps.y);
            } else {
                n = Perlin2(px, py, q.seed, q.seedLcg);
            }
            sum += q.weight * n;
        }
        return p.amplitude * sum / p.totalWeight;
    }
    for (uint32_t o = 0; o < octaves; ++o) {
        const OctavePlan& q = p.octave[o];
        const float rx = x * q.turnCos - y * q.turnSin;
        const float ry = x * q.turnSin + y * q.turnCos;
        const float n = Basis2(basis, rx * q.frequency + q.offsetX, ry * q.frequency + q.offsetY, q);
        const float weight = q.weight;
        switch (type) {
        case TG_FRACTAL_BILLOW:
            sum += weight * (2.0f * fabsf(n) - 1.0f);
            break;
        case TG_FRACTAL_RIDGED: {
            float signal = p.offset - fabsf(n);
            signal *= signal * feedback;
            feedback = Clamp01(2.0f * signal);
            sum += weight * signal;
            break;
        }
        case TG_FRACTAL_HYBRID: {
            const float signal = (n + p.offset) * weight;
            if (o == 0) {
                sum = signal;
                feedback = signal;
            } else {
                feedback = Clamp01(feedback);
                sum += feedback * signal;
                feedback *= signal;
            }
            break;
        }
        default:
            sum += weight * n;
            break;
        }
    }
    switch (type) {
    case TG_FRACTAL_RIDGED: {
        const float ridge = StdMax(p.offset, 1.0f - p.offset);
        return p.amplitude * (2.0f * sum / (ridge * ridge * p.totalWeight) - 1.0f);
    }
    case TG_FRACTAL_HYBRID:
        return p.amplitude * sum / ((1.0f + p.offset) * p.totalWeight);
    default:
        return p.amplitude * sum / p.totalWeight;
    }
}

// EvaluateFractals (fractals.h:106-135,151-196): the stack's height at (x, y).
// kSquare: as for FractalLoop; a warped fractal's coordinates are not monotone across the square,
// so it always takes the per-lane path.
template <bool kSquare = false>
__device__ __forceinline__ float EvaluateStack(const NoisePlan& plan, float x, float y,
                                               uint32_t lane = 0u) {
    float height = 0.0f;
    const uint32_t count = plan.count;
    for (uint32_t i = 0; i < count; ++i) {
        const FractalPlan& f = plan.fractal[i];
        float value;
        if (f.warped) {
            const float dx = FractalLoop(f.warpX, x, y);
            const float dy = FractalLoop(f.warpY, x, y);
            value = FractalLoop(f.unit, x + dx, y + dy);
        } else {
            value = FractalLoop<kSquare>(f.unit, x, y, lane);
        }
        const float scaled = f.amplitude * value;
        if (i == 0) {
            height = scaled;
            continue;
        }
        switch (f.blend) {
        case TG_BLEND_MAX:
            height = StdMax(height, scaled);
            break;
        case TG_BLEND_MIN:
            height = StdMin(height, scaled);
            break;
        case TG_BLEND_MASK:
            height *= 0.5f * (value + 1.0f);
            break;
        default:
            height += scaled;
            break;
        }
    }
    return height;
}

// ---- Noise on a body ---------------------------------------------------------------------------

// PerlinGradient3 (perlin.h:38-43), branch-free: the two terms picked by selects and signed by the
// hash's low bits (negation is the sign-bit flip).
__device__ __forceinline__ float Grad3(uint32_t hash, float x, float y, float z) {
    const uint32_t h = hash & 15u;
    const float u = h < 8u ? x : y;
    const float v = h < 4u ? y : (h == 12u || h == 14u ? x : z);
    const float su = __uint_as_float(__float_as_uint(u) ^ ((h & 1u) << 31u));
    const float sv = __uint_as_float(__float_as_uint(v) ^ ((h & 2u) << 30u));
    return su + sv;
}

// PerlinNoise3 (perlin.h:56-63,113-128) at cell (ix, iy, iz) and fraction (fx, fy, fz). A corner's
// hash is the x word of pcg3d(x, y, z + PcgHash(seed)) (lattice.h:99-102); the LCG step of each
// coordinate is linear in it, so the eight corners share their three pairs of LCG words.
__device__ __forceinline__ float Perlin3(int32_t ix, int32_t iy, int32_t iz, float fx, float fy,
                                         float fz, uint32_t zKey) {
    const uint32_t ax0 = uint32_t(ix) * kLcgMul + kLcgAdd;
    const uint32_t ay0 = uint32_t(iy) * kLcgMul + kLcgAdd;
    const uint32_t az0 = (uint32_t(iz) + zKey) * kLcgMul + kLcgAdd;
    const uint32_t ax1 = ax0 + kLcgMul, ay1 = ay0 + kLcgMul, az1 = az0 + kLcgMul;
    const float c000 = Grad3(CornerX(ax0, ay0, az0, ay0 * az0), fx, fy, fz);
    const float c100 = Grad3(CornerX(ax1, ay0, az0, ay0 * az0), fx - 1.0f, fy, fz);
    const float c010 = Grad3(CornerX(ax0, ay1, az0, ay1 * az0), fx, fy - 1.0f, fz);
    const float c110 = Grad3(CornerX(ax1, ay1, az0, ay1 * az0), fx - 1.0f, fy - 1.0f, fz);
    const float c001 = Grad3(CornerX(ax0, ay0, az1, ay0 * az1), fx, fy, fz - 1.0f);
    const float c101 = Grad3(CornerX(ax1, ay0, az1, ay0 * az1), fx - 1.0f, fy, fz - 1.0f);
    const float c011 = Grad3(CornerX(ax0, ay1, az1, ay1 * az1), fx, fy - 1.0f, fz - 1.0f);
    const float c111 = Grad3(CornerX(ax1, ay1, az1, ay1 * az1), fx - 1.0f, fy - 1.0f, fz - 1.0f);
    const float u = Fade(fx);
    const float v = Fade(fy);
    const float w = Fade(fz);
    const float nearSide = Mix(v, Mix(u, c000, c100), Mix(u, c010, c110));
    const float farSide = Mix(v, Mix(u, c001, c101), Mix(u, c011, c111));
    return StdClamp(Mix(w, nearSide, farSide), -1.0f, 1.0f);
}

// The lane's point in a 4 x 4 x 4 block of lattice points (lane & 3, (lane >> 2) & 3, lane >> 4)
// times the LCG multiplier, for Perlin3Shared; the empty asm keeps the compiler from
// re-multiplying them every octave (as SquareSteps does in 2D).
struct CubeSteps {
    uint32_t x, y, z;
};
__device__ __forceinline__ CubeSteps MakeCubeSteps(uint32_t lane) {
    uint32_t x = (lane & 3u) * kLcgMul;
    uint32_t y = ((lane >> 2u) & 3u) * kLcgMul;
    uint32_t z = (lane >> 4u) * kLcgMul;
    asm("" : "+v"(x), "+v"(y), "+v"(z));
    return {x, y, z};
}

// Perlin3 for a wave of 64 active lanes whose cells lie close together (a tile's 8 x 8 square on
// a body). When every lane's lattice cell lies within one cell of lane 27's along each axis, the
// 4 x 4 x 4 lattice points from one below lane 27's cell hold every corner of the wave: each lane
// hashes one of them (one PCG hash instead of eight, whose 32-bit multiplies are the costliest
// instructions of the noise), the gradient codes (the hash's low four bits, all PerlinGradient3
// reads) of the four points of each point's cell face are packed into one word by two lane
// exchanges, and each lane reads its cell's two faces with two more. Otherwise every lane hashes
// its own eight corners. The hashes' low bits, and so the noise, are Perlin3's.
__device__ __forceinline__ float Perlin3Shared(int32_t ix, int32_t iy, int32_t iz, float fx,
                                               float fy, float fz, uint32_t zKey, uint32_t lane,
                                               const CubeSteps& steps) {
    const int32_t bx = __builtin_amdgcn_readlane(ix, 27) - 1;
    const int32_t by = __builtin_amdgcn_readlane(iy, 27) - 1;
    const int32_t bz = __builtin_amdgcn_readlane(iz, 27) - 1;
    const uint32_t rx = uint32_t(ix - bx), ry = uint32_t(iy - by), rz = uint32_t(iz - bz);
    if (__ballot(rx > 2u || ry > 2u || rz > 2u) != 0ull) {
        return Perlin3(ix, iy, iz, fx, fy, fz, zKey);
    }
    const uint32_t ax = uint32_t(bx) * kLcgMul + kLcgAdd + steps.x;
    const uint32_t ay = uint32_t(by) * kLcgMul + kLcgAdd + steps.y;
    const uint32_t az = (uint32_t(bz) + zKey) * kLcgMul + kLcgAdd + steps.z;
    // Nibble k of `codes`: the code of corner (k & 1, (k >> 1) & 1, k >> 2) of the lane's cell.
    // Lanes whose point has no cell in the block (an index of 3) pack words no lane reads.
    const auto from = [](uint32_t byteAddress, uint32_t value) {
        return uint32_t(__builtin_amdgcn_ds_bpermute(int32_t(byteAddress), int32_t(value)));
    };
    uint32_t word = CornerX(ax, ay, az, ay * az) & 15u;
    const uint32_t byteAt = lane << 2u;
    word |= from(byteAt + 4u, word) << 4u;    // the point one along x
    word |= from(byteAt + 16u, word) << 8u;   // one along y
    // The cell's four corners at its z and the four one along z, read independently.
    const uint32_t cellAt = (rx + 4u * ry + 16u * rz) << 2u;
    const uint32_t codes = from(cellAt, word) | (from(cellAt + 64u, word) << 16u);
    const float c000 = Grad3(codes, fx, fy, fz);
    const float c100 = Grad3(codes >> 4u, fx - 1.0f, fy, fz);
    const float c010 = Grad3(codes >> 8u, fx, fy - 1.0f, fz);
    const float c110 = Grad3(codes >> 12u, fx - 1.0f, fy - 1.0f, fz);
    const float c001 = Grad3(codes >> 16u, fx, fy, fz - 1.0f);
    const float c101 = Grad3(codes >> 20u, fx - 1.0f, fy, fz - 1.0f);
    const float c011 = Grad3(codes >> 24u, fx, fy - 1.0f, fz - 1.0f);
    const float c111 = Grad3(codes >> 28u, fx - 1.0f, fy - 1.0f, fz - 1.0f);
    const float u = Fade(fx);
    const float v = Fade(fy);
    const float w = Fade(fz);
    const float nearSide = Mix(v, Mix(u, c000, c100), Mix(u, c010, c110));
    const float farSide = Mix(v, Mix(u, c001, c101), Mix(u, c011, c111));
    return StdClamp(Mix(w, nearSide, farSide), -1.0f, 1.0f);
}

// SimplexNoise3 (simplex.h:27-35,74-114) at cell (ix, iy, iz) and fraction (fx, fy, fz): the skew
// split into the cell sum's integer third and a fraction, corner hashes as Perlin's (zKey =
// PcgHash(seed)). A corner beyond the kernel's radius is zero, picked by a select as in Simplex2.
__device__ __forceinline__ float SimplexCorner3(uint32_t ci, uint32_t cj, uint32_t ck, float dx,
                                                float dy, float dz, uint32_t zKey) {
    const float t = 0.5f - dx * dx - dy * dy - dz * dz;
    const float t2 = t * t;
    const float corner = t2 * t2 * Grad3(Pcg3({ci, cj, ck + zKey}).x, dx, dy, dz);
    return !(t <= 0.0f) ? corner : 0.0f;
}

__device__ __forceinline__ float Simplex3(int32_t ix, int32_t iy, int32_t iz, float fx, float fy,
                                          float fz, uint32_t zKey) {
    const int32_t cellSum = int32_t(uint32_t(ix) + uint32_t(iy) + uint32_t(iz));
    int32_t third = cellSum / 3;
    if (cellSum - 3 * third < 0) {
        third -= 1;
    }
    const float remainder = float(cellSum - 3 * third);
    const float skew = (remainder + fx + fy + fz) / 3.0f;
    const float stepX = floorf(fx + skew);
    const float stepY = floorf(fy + skew);
    const float stepZ = floorf(fz + skew);
    const float unskew = (remainder + stepX + stepY + stepZ) / 6.0f;
    const float x0 = fx - stepX + unskew;
    const float y0 = fy - stepY + unskew;
    const float z0 = fz - stepZ + unskew;
    const uint32_t i1 = x0 >= y0 && x0 >= z0 ? 1u : 0u;
    const uint32_t j1 = y0 > x0 && y0 >= z0 ? 1u : 0u;
    const uint32_t k1 = z0 > x0 && z0 > y0 ? 1u : 0u;
    const uint32_t i2 = x0 >= y0 || x0 >= z0 ? 1u : 0u;
    const uint32_t j2 = y0 > x0 || y0 >= z0 ? 1u : 0u;
    const uint32_t k2 = z0 > x0 || z0 > y0 ? 1u : 0u;
    const float kUnskew = 1.0f / 6.0f;
    const float x1 = x0 - float(i1) + kUnskew;
    const float y1 = y0 - float(j1) + kUnskew;
    const float z1 = z0 - float(k1) + kUnskew;
    const float x2 = x0 - float(i2) + 2.0f * kUnskew;
    const float y2 = y0 - float(j2) + 2.0f * kUnskew;
    const float z2 = z0 - float(k2) + 2.0f * kUnskew;
    const float x3 = x0 - 1.0f + 3.0f * kUnskew;
    const float y3 = y0 - 1.0f + 3.0f * kUnskew;
    const float z3 = z0 - 1.0f + 3.0f * kUnskew;
    const uint32_t i = uint32_t(ix) + uint32_t(third) + uint32_t(int32_t(stepX));
    const uint32_t j = uint32_t(iy) + uint32_t(third) + uint32_t(int32_t(stepY));
    const uint32_t k = uint32_t(iz) + uint32_t(third) + uint32_t(int32_t(stepZ));
    const float sum = SimplexCorner3(i, j, k, x0, y0, z0, zKey) +
                      SimplexCorner3(i + i1, j + j1, k + k1, x1, y1, z1, zKey) +
                      SimplexCorner3(i + i2, j + j2, k + k2, x2, y2, z2, zKey) +
                      SimplexCorner3(i + 1u, j + 1u, k + 1u, x3, y3, z3, zKey);
    return StdClamp(76.0f * sum, -1.0f, 1.0f);
}

// ValueNoise3 (value.h:22-26,57-73): each corner's value from pcg3d(x, y, z ^ key).y with key =
// PcgHash(seed) (ValueLatticeKey3, lattice.h:108-111), blended with the quintic fade.
__device__ __forceinline__ float Value3(int32_t ix, int32_t iy, int32_t iz, float fx, float fy,
                                        float fz, uint32_t key) {
    const uint32_t x0 = uint32_t(ix), y0 = uint32_t(iy), z0 = uint32_t(iz);
    const auto value = [key](uint32_t x, uint32_t y, uint32_t z) {
        return 2.0f * UnitFloat(Pcg3({x, y, z ^ key}).y) - 1.0f;
    };
    const float v000 = value(x0, y0, z0);
    const float v100 = value(x0 + 1u, y0, z0);
    const float v010 = value(x0, y0 + 1u, z0);
    const float v110 = value(x0 + 1u, y0 + 1u, z0);
    const float v001 = value(x0, y0, z0 + 1u);
    const float v101 = value(x0 + 1u, y0, z0 + 1u);
    const float v011 = value(x0, y0 + 1u, z0 + 1u);
    const float v111 = value(x0 + 1u, y0 + 1u, z0 + 1u);
    const float u = Fade(fx);
    const float v = Fade(fy);
    const float w = Fade(fz);
    const float nearSide = Mix(v, Mix(u, v000, v100), Mix(u, v010, v110));
    const float farSide = Mix(v, Mix(u, v001, v101), Mix(u, v011, v111));
    return Mix(w, nearSide, farSide);
}

// CellGapSquared (worley.h:24-32): the squared gap along one axis to a cell d cells away.
__device__ __forceinline__ float CellGap2(int32_t d, float f) {
    float gap = 0.0f;
    if (d > 0) {
        gap = float(d) - f;
    } else if (d < 0) {
        gap = f - float(d + 1);
    }
    return gap * gap;
}

// Worley3 (worley.h:75-114): the 3 x 3 x 3 cells around the point, then the outer shell of the
// 5 x 5 x 5 cells whose gap to the point is below F2; feature points from
// pcg3d(x ^ key, y, z) with key = PcgHash(seed) (CellularLatticeKey3, lattice.h:116-119), the
// nearest cell's value from PcgHash of its three words.
__device__ __forceinline__ Worley Worley3(int32_t ix, int32_t iy, int32_t iz, float fx, float fy,
                                          float fz, uint32_t key) {
    float d1 = 1e30f;
    float d2 = 1e30f;
    uint32_t nearest = 0u;
    const auto visit = [&](int32_t dx, int32_t dy, int32_t dz) {
        const U3 h = Pcg3({(uint32_t(ix) + uint32_t(dx)) ^ key, uint32_t(iy) + uint32_t(dy),
                           uint32_t(iz) + uint32_t(dz)});
        const float px = float(dx) + UnitFloat(h.x) - fx;
        const float py = float(dy) + UnitFloat(h.y) - fy;
        const float pz = float(dz) + UnitFloat(h.z) - fz;
        const float d = px * px + py * py + pz * pz;
        if (d < d1) {
            d2 = d1;
            d1 = d;
            nearest = Pcg(h.x ^ h.y ^ h.z);
        } else if (d < d2) {
            d2 = d;
        }
    };
    for (int32_t dz = -1; dz <= 1; ++dz) {
        for (int32_t dy = -1; dy <= 1; ++dy) {
            for (int32_t dx = -1; dx <= 1; ++dx) {
                visit(dx, dy, dz);
            }
        }
    }
    for (int32_t dz = -2; dz <= 2; ++dz) {
        for (int32_t dy = -2; dy <= 2; ++dy) {
            for (int32_t dx = -2; dx <= 2; ++dx) {
                if (abs(dx) <= 1 && abs(dy) <= 1 && abs(dz) <= 1) {
                    continue;
                }
                if (CellGap2(dx, fx) + CellGap2(dy, fy) + CellGap2(dz, fz) >= d2) {
                    continue;
                }
                visit(dx, dy, dz);
            }
        }
    }
    return {sqrtf(d1), sqrtf(d2), 2.0f * UnitFloat(nearest) - 1.0f};
}

// BasisNoise3 (fractal.h:115-136) for the bases other than Perlin, with key = PcgHash of the
// octave's seed; Worley distances over their 3D bounds sqrt(3) and sqrt(6).
__device__ __forceinline__ float Basis3(uint32_t basis, int32_t ix, int32_t iy, int32_t iz,
                                        float fx, float fy, float fz, uint32_t key) {
    const float kInvSqrt3 = 0.577350269189625765f;
    const float kInvSqrt6 = 0.408248290463863016f;
    switch (basis) {
    case TG_BASIS_SIMPLEX:
        return Simplex3(ix, iy, iz, fx, fy, fz, key);
    case TG_BASIS_VALUE:
        return Value3(ix, iy, iz, fx, fy, fz, key);
    case TG_BASIS_WORLEY_F1:
        return 2.0f * StdMin(Worley3(ix, iy, iz, fx, fy, fz, key).f1 * kInvSqrt3, 1.0f) - 1.0f;
    case TG_BASIS_WORLEY_F2:
        return 2.0f * StdMin(Worley3(ix, iy, iz, fx, fy, fz, key).f2 * kInvSqrt6, 1.0f) - 1.0f;
    case TG_BASIS_WORLEY_F2_MINUS_F1: {
        const Worley w = Worley3(ix, iy, iz, fx, fy, fz, key);
        return 2.0f * StdMin((w.f2 - w.f1) * kInvSqrt6, 1.0f) - 1.0f;
    }
    case TG_BASIS_VORONOI:
        return Worley3(ix, iy, iz, fx, fy, fz, key).value;
    default:
        return Perlin3(ix, iy, iz, fx, fy, fz, key);
    }
}

// Fractal3 (fractal.h:284-317) at body-fixed (x, y, z) meters: the octave's
// FP64 turn and frequency from the plan, lattice coordinates split in FP64 before narrowing
// (SplitLatticePrecise3, lattice.h:59-69), the octave sums as the 2D loop's.
// kShared: the caller's 64 lanes are all active and their cells close together, as
// Perlin3Shared requires, and `lane` and `steps` are the lane's.
// kSmall: every lattice coordinate lies within 2^31 of the origin (the caller has bounded the
// positions and frequencies), so int32_t(int64_t(c)) is one conversion of the double.
// kGeneral: any basis; when false the caller has checked that the basis is Perlin
// (NoisePlan::bodyPerlin), and the other bases' code stays out of the kernel.
template <bool kShared = false, bool kSmall = false, bool kGeneral = true>
__device__ __forceinline__ float FractalLoop3(const FractalLoopPlan& p, const BodyOctave* body,
                                              double x, double y, double z, uint32_t lane = 0u,
                                              CubeSteps steps = {}) {
    const auto index = [](double c) {
        if constexpr (kSmall) {
            return int32_t(c);
        } else {
            return int32_t(int64_t(c));
        }
    };
    const uint32_t type = p.type;
    float sum = 0.0f;
    float feedback = 1.0f;
    for (uint32_t o = 0; o < p.octaves; ++o) {
        const BodyOctave& b = body[o];
        const double rx = b.turn[0] * x + b.turn[1] * y + b.turn[2] * z;
        const double ry = b.turn[3] * x + b.turn[4] * y + b.turn[5] * z;
        const double rz = b.turn[6] * x + b.turn[7] * y + b.turn[8] * z;
        const double px = rx * b.frequency + b.offsetX;
        const double py = ry * b.frequency + b.offsetY;
        const double pz = rz * b.frequency + b.offsetZ;
        const double cx = floor(px);
        const double cy = floor(py);
        const double cz = floor(pz);
        float n;
        if (kGeneral && p.basis != TG_BASIS_PERLIN) {
            n = Basis3(p.basis, index(cx), index(cy), index(cz), float(px - cx), float(py - cy),
                       float(pz - cz), b.zKey);
        } else if constexpr (kShared) {
            n = Perlin3Shared(index(cx), index(cy), index(cz), float(px - cx), float(py - cy),
                              float(pz - cz), b.zKey, lane, steps);
        } else {
            n = Perlin3(index(cx), index(cy), index(cz), float(px - cx), float(py - cy),
                        float(pz - cz), b.zKey);
        }
        const float weight = p.octave[o].weight;
        switch (type) {
        case TG_FRACTAL_BILLOW:
            sum += weight * (2.0f * fabsf(n) - 1.0f);
            break;
        case TG_FRACTAL_RIDGED: {
            float signal = p.offset - fabsf(n);
            signal *= signal * feedback;
            feedback = Clamp01(2.0f * signal);
            sum += weight * signal;
            break;
        }
        case TG_FRACTAL_HYBRID: {
            const float signal = (n + p.offset) * weight;
            if (o == 0) {
                sum = signal;
                feedback = signal;
            } else {
                feedback = Clamp01(feedback);
                sum += feedback * signal;
                feedback *= signal;
            }
            break;
        }
        default:
            sum += weight * n;
            break;
        }
    }
    switch (type) {
    case TG_FRACTAL_RIDGED: {
        const float ridge = StdMax(p.offset, 1.0f - p.offset);
        return p.amplitude * (2.0f * sum / (ridge * ridge * p.totalWeight) - 1.0f);
    }
    case TG_FRACTAL_HYBRID:
        return p.amplitude * sum / ((1.0f + p.offset) * p.totalWeight);
    default:
        return p.amplitude * sum / p.totalWeight;
    }
}

// EvaluateFractals3 (fractals.h:221-269): the stack's height at body-fixed (x, y, z) meters, a
// warped fractal evaluated at the point moved by its three warp fields (DomainWarp3), each a
// Fractal3 whose FP32 value is added in FP64. kGeneral as for FractalLoop3: when false, every
// fractal is an unwarped Perlin one.
template <bool kShared = false, bool kSmall = false, bool kGeneral = true>
__device__ __forceinline__ float EvaluateStack3(const NoisePlan& plan, double x, double y,
                                                double z, uint32_t lane = 0u,
                                                CubeSteps steps = {}) {
    float height = 0.0f;
    const uint32_t count = plan.count;
    for (uint32_t i = 0; i < count; ++i) {
        const FractalPlan& f = plan.fractal[i];
        double wx = x, wy = y, wz = z;
        if (kGeneral && f.warped) {
            wx = x + double(FractalLoop3<kShared, kSmall>(f.warpX, plan.warpBody[i][0], x, y, z,
                                                          lane, steps));
            wy = y + double(FractalLoop3<kShared, kSmall>(f.warpX, plan.warpBody[i][1], x, y, z,
                                                          lane, steps));
            wz = z + double(FractalLoop3<kShared, kSmall>(f.warpX, plan.warpBody[i][2], x, y, z,
                                                          lane, steps));
        }
        const float value =
            FractalLoop3<kShared, kSmall, kGeneral>(f.unit, plan.body[i], wx, wy, wz, lane, steps);
        const float scaled = f.amplitude * value;
        if (i == 0) {
            height = scaled;
            continue;
        }
        switch (f.blend) {
        case TG_BLEND_MAX:
            height = StdMax(height, scaled);
            break;
        case TG_BLEND_MIN:
            height = StdMin(height, scaled);
            break;
        case TG_BLEND_MASK:
            height *= 0.5f * (value + 1.0f);
            break;
        default:
            height += scaled;
            break;
        }
    }
    return height;
}

}  // namespace tg
