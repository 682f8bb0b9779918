// SPDX-License-Identifier: BSD-3-Clause
//
// Crater shaping on gfx950: the height one crater gives a cell, in FP32, as the reference's
// kernel header computes it (include/terraingen/layers/kernels/craters.h:150-1212). Records, the
// shape block and the variable-length crater data use the reference's layouts (tg_layouts.h).
// Transcendentals are the HIP device library's; the reference's own bound for GPU vs CPU is
// 2e-3 m (tests/unit/layers/craters/shaping/crater_kernel_test.cpp:212-272) and its kernels keep
// series where GPU inverse trigonometry is too coarse (KernelArcsineRatio, craters.h:156-166).
#pragma once

#include "noise_device.h"
#include "profile_laws.h"
#include "tg_layouts.h"

#include <hip/hip_runtime.h>

namespace tg {

constexpr float kPi = 3.14159265358979f;
constexpr float kTwoPi = 6.28318530717959f;

struct Offset {
    float dx, dy;
};

// A thin view of the shape block so device code reads it by name.
struct CraterShape {
    const tg_crater_shape& k;
    __device__ explicit CraterShape(const tg_crater_shape& block) : k(block) {}
};

// SplitDifference (craters.h:152-154).
__device__ __forceinline__ float SplitDifference(float a, float aLow, float b, float bLow) {
    return (a - b) + (aLow - bLow);
}

// KernelArcsineRatio (craters.h:156-166): asin(s)/s by its series below 0.25.
__device__ __forceinline__ float ArcsineRatio(float s) {
    if (s < 0.25f) {
        const float s2 = s * s;
        return 1.0f + s2 * (1.0f / 6.0f +
                            s2 * (3.0f / 40.0f + s2 * (5.0f / 112.0f + s2 * (35.0f / 1152.0f))));
    }
    return atan2f(s, sqrtf(1.0f - s * s)) / s;
}

// KernelFrameOffset (craters.h:168-186).
__device__ __forceinline__ Offset FrameOffset(const tg_crater_frame& f, float dx, float dy,
                                              float dz) {
    Offset o;
    o.dx = dx * f.east_x + dy * f.east_y + dz * f.east_z;
    o.dy = dx * f.north_x + dy * f.north_y + dz * f.north_z;
    if (f.sphere_radius > 0.0f) {
        const float chord = sqrtf(dx * dx + dy * dy + dz * dz);
        const float arc = chord * ArcsineRatio(StdMin(0.5f * chord / f.sphere_radius, 1.0f));
        const float across = sqrtf(o.dx * o.dx + o.dy * o.dy);
        if (across > 0.0f) {
            o.dx *= arc / across;
            o.dy *= arc / across;
        } else {
            o.dy = arc;
        }
    }
    return o;
}

// BodySplitOffset (craters.h:204-209).
__device__ __forceinline__ Offset BodySplitOffset(const tg_crater_frame& f,
                                                  const tg_split_point& p) {
    return FrameOffset(f, SplitDifference(p.x, p.x_low, f.x, f.x_low),
                       SplitDifference(p.y, p.y_low, f.y, f.y_low),
                       SplitDifference(p.z, p.z_low, f.z, f.z_low));
}

// KernelFalloffShape / KernelFalloff (craters.h:241-270).
__device__ __forceinline__ float FalloffShape(float t, uint32_t shape) {
    if (shape == 0u) {
        return 1.0f - t;
    }
    if (shape == 1u) {
        return 1.0f - t * t * (3.0f - 2.0f * t);
    }
    if (shape == 2u) {
        return 1.0f - t * t * t * (t * (t * 6.0f - 15.0f) + 10.0f);
    }
    const float tail = expf(-4.5f);
    const float z = 3.0f * t;
    return (expf(-0.5f * z * z) - tail) / (1.0f - tail);
}

__device__ __forceinline__ float Falloff(float x, float start, float end, uint32_t shape) {
    if (x <= start) {
        return 1.0f;
    }
    if (x >= end) {
        return 0.0f;
    }
    return FalloffShape((x - start) / (end - start), shape);
}

// KernelSmoothstep (craters.h:272-276).
__device__ __forceinline__ float Smoothstep(float t) {
    const float u = StdClamp(t, 0.0f, 1.0f);
    return u * u * (3.0f - 2.0f * u);
}

// KernelAngle (craters.h:278-285).
__device__ __forceinline__ float Angle(float dy, float dx) {
    if (dx == 0.0f && dy == 0.0f) {
        return 0.0f;
    }
    return atan2f(dy, dx);
}

struct PeriodicCell {
    uint32_t index;
    float t;
};

// KernelPeriodicAt (craters.h:297-304).
__device__ __forceinline__ PeriodicCell PeriodicAt(float turns, uint32_t size) {
    const float position = (turns - floorf(turns)) * float(size);
    PeriodicCell cell;
    const uint32_t whole = uint32_t(position);
    cell.index = whole < size - 1u ? whole : size - 1u;
    cell.t = position - float(cell.index);
    return cell;
}

// KernelEjectaShare (craters.h:313-324).
__device__ __forceinline__ float EjectaShare(float radii, float reachRadii) {
    if (radii <= 1.0f) {
        return 1.0f;
    }
    if (radii >= reachRadii) {
        return 0.0f;
    }
    const float atReach = 1.0f / (reachRadii * reachRadii * reachRadii);
    return (1.0f / (radii * radii * radii) - atReach) / (1.0f - atReach);
}

struct TableCell {
    uint32_t index;
    float t;
};

// KernelTableAt (craters.h:511-518).
__device__ __forceinline__ TableCell TableAt(const tg_crater_profile& p, float x) {
    const float position = x / p.table_extent * float(p.table_cells);
    TableCell cell;
    const uint32_t whole = uint32_t(position);
    cell.index = whole < p.table_cells - 1u ? whole : p.table_cells - 1u;
    cell.t = position - float(cell.index);
    return cell;
}

// KernelLerp (craters.h:520-523).
__device__ __forceinline__ float DataLerp(const float* data, uint32_t at, float t) {
    return (1.0f - t) * data[at] + t * data[at + 1u];
}

struct AgeRow {
    uint32_t row;
    float weight;
};

// KernelAgeRowAt (craters.h:531-559): binary search of the tabulated diffusion ages.
__device__ __forceinline__ AgeRow AgeRowAt(const tg_crater_profile& p, const float* data,
                                           float tau) {
    AgeRow age;
    age.row = 0u;
    age.weight = 0.0f;
    const uint32_t last = p.age_count - 1u;
    const float first = data[p.ages_offset + 1u];
    if (tau >= data[p.ages_offset + last]) {
        age.row = last;
    } else if (tau < first) {
        age.weight = tau / first;
    } else {
        uint32_t low = 1u;
        uint32_t high = last;
        while (high - low > 1u) {
            const uint32_t middle = (low + high) / 2u;
            if (data[p.ages_offset + middle] <= tau) {
                low = middle;
            } else {
                high = middle;
            }
        }
        age.row = low;
        const float below = data[p.ages_offset + low];
        const float above = data[p.ages_offset + high];
        age.weight = logf(tau / below) / logf(above / below);
    }
    return age;
}

// KernelDiffusedValue (craters.h:565-577).
__device__ __forceinline__ float DiffusedValue(const tg_crater_profile& p, const float* data,
                                               uint32_t rows, float x, float tau) {
    const TableCell cell = TableAt(p, x);
    const AgeRow age = AgeRowAt(p, data, tau);
    const uint32_t width = p.table_cells + 1u;
    const uint32_t at = rows + age.row * width + cell.index;
    float value = DataLerp(data, at, cell.t);
    if (age.weight > 0.0f) {
        value += age.weight * (DataLerp(data, at + width, cell.t) - value);
    }
    return value;
}

// Per-crater values every profile evaluation of one (cell, crater) pair shares.
struct CraterView {
    const tg_crater_record& c;
    ProfileState current;
    ProfileState fresh;
};

// KernelProfileHeight (craters.h:608-618) with DiffusedProfileHeight (craters.h:580-594) and
// TabulatedProfileHeight (craters.h:598-606).
__device__ __forceinline__ float ProfileHeight(const tg_crater_shape& k, const float* data,
                                               float distance, float diameter,
                                               const CraterView& v) {
    const tg_crater_profile& p = k.profile;
    if (p.simple == TG_PROFILE_TABULATED) {
        const float x = distance / diameter;
        if (!(x < p.table_extent)) {
            return 0.0f;
        }
        const TableCell cell = TableAt(p, x);
        return diameter * DataLerp(data, v.c.profile_offset + cell.index, cell.t);
    }
    if (p.diffused == 1u) {
        if (!(v.c.diffusion > 0.0f)) {
            return FreshProfileHeight(p, distance, diameter, v.fresh);
        }
        const float x = distance / diameter;
        if (x >= p.table_extent) {
            return 0.0f;
        }
        return diameter * DiffusedValue(p, data, v.c.profile_offset, x, v.c.diffusion);
    }
    return FreshProfileHeight(p, distance, diameter, v.current);
}

// KernelRoundedHeight (craters.h:624-653): the rim rounded by a Gaussian on a grid anchored at it.
__device__ __forceinline__ float RoundedHeight(const tg_crater_shape& k, const float* data,
                                               float distance, float diameter,
                                               const CraterView& v) {
    const float height = ProfileHeight(k, data, distance, diameter, v);
    if (k.rim_sigma == 0.0f) {
        return height;
    }
    const float radius = 0.5f * diameter;
    const float radii = distance / radius;
    const float weight =
        radii <= 1.0f
            ? k.rim_strength * Falloff(2.0f - radii, 1.0f, 2.0f - k.rim_inner, k.rim_shape)
            : k.rim_strength * Falloff(radii, 1.0f, k.rim_outer, k.rim_shape);
    if (weight == 0.0f) {
        return height;
    }
    const float sigma = k.rim_sigma * radius;
    const float step = 0.5f * sigma;
    const float position = (distance - radius) / step;
    const float first = floorf(position) - 7.0f;
    float smoothed = 0.0f;
    float total = 0.0f;
    for (int i = 0; i < 16; ++i) {
        const float node = first + float(i);
        const float z = 0.5f * (node - position);
        const float w = expf(-0.5f * z * z) * Falloff(fabsf(z), 3.0f, 4.0f, 2u);
        smoothed += w * ProfileHeight(k, data, fabsf(radius + step * node), diameter, v);
        total += w;
    }
    return height + weight * (smoothed / total - height);
}

// KernelFractalParams (craters.h:657-669) evaluated as Fractal2 (fractal.h:219-239); the octave
// seeds are hashed per call since the key varies by crater.
__device__ __forceinline__ float KernelFractal2(float x, float y, uint32_t seed, float wavelength,
                                                int octaves, float gain, uint32_t type) {
    float frequency = 1.0f / wavelength;
    float turnCos = 1.0f;
    float turnSin = 0.0f;
    float weight = 1.0f;
    float sum = 0.0f;
    float total = 0.0f;
    float feedback = 1.0f;
    for (int octave = 0; octave < octaves; ++octave) {
        const U3 h = Pcg3({seed, uint32_t(octave), 0u});
        const float offsetX = UnitFloat(h.y);
        const float offsetY = UnitFloat(h.z);
        const float rx = x * turnCos - y * turnSin;
        const float ry = x * turnSin + y * turnCos;
        const uint32_t seedLcg = h.x * kLcgMul + kLcgAdd;
        const float n = Perlin2(rx * frequency + offsetX, ry * frequency + offsetY, h.x, seedLcg);
        if (type == TG_FRACTAL_RIDGED) {
            float signal = 1.0f - fabsf(n);
            signal *= signal * feedback;
            feedback = Clamp01(2.0f * signal);
            sum += weight * signal;
        } else {
            sum += weight * n;
        }
        total += weight;
        weight *= gain;
        frequency *= 2.0f;
        const float nextCos = turnCos * -0.7373688780783197f - turnSin * 0.6754902942615238f;
        turnSin = turnSin * -0.7373688780783197f + turnCos * 0.6754902942615238f;
        turnCos = nextCos;
    }
    if (type == TG_FRACTAL_RIDGED) {
        // ridge scale max(1, 0) = 1 for the kernels' offset of 1 (fractal.h:203-205).
        return 1.0f * (2.0f * sum / (1.0f * 1.0f * total) - 1.0f);
    }
    return 1.0f * sum / total;
}

// KernelNoiseHeight (craters.h:672-694).
__device__ __forceinline__ float NoiseHeight(const tg_crater_shape& k, const tg_crater_record& c,
                                             float dx, float dy, float radii) {
    if (k.noise_amplitude == 0.0f && k.noise_rim_amplitude == 0.0f) {
        return 0.0f;
    }
    const float radius = 0.5f * c.diameter;
    const float footprint =
        radii <= 1.0f ? 1.0f : Falloff(radii, 1.0f, k.profile_reach_radii, 2u);
    const float rim = radii <= 1.0f ? Falloff(2.0f - radii, 1.0f, 2.0f - k.noise_rim_inner, 2u)
                                    : Falloff(radii, 1.0f, k.noise_rim_outer, 2u);
    float height = 0.0f;
    if (k.noise_amplitude > 0.0f && footprint > 0.0f) {
        height += k.noise_amplitude * footprint *
                  KernelFractal2(dx, dy, c.noise_key, k.noise_wavelength * radius,
                                 k.noise_octaves, k.noise_gain, TG_FRACTAL_FBM);
    }
    if (k.noise_rim_amplitude > 0.0f && rim > 0.0f) {
        height += k.noise_rim_amplitude * rim *
                  KernelFractal2(dx, dy, c.noise_key + 1u, k.noise_rim_wavelength * radius,
                                 k.noise_octaves, k.noise_gain, TG_FRACTAL_FBM);
    }
    return c.diameter * height;
}

// Fractal3Float (fractal.h:243-277) for Perlin fBm: the 3D octaves the texture streaks read.
__device__ __forceinline__ float PerlinGrad3(uint32_t hash, float x, float y, float z) {
    const uint32_t h = hash & 15u;
    const float u = h < 8u ? x : y;
    const float v = h < 4u ? y : (h == 12u || h == 14u ? x : z);
    return ((h & 1u) == 0u ? u : -u) + ((h & 2u) == 0u ? v : -v);
}

__device__ __forceinline__ float Perlin3(float x, float y, float z, uint32_t seed) {
    const float cx = floorf(x), cy = floorf(y), cz = floorf(z);
    const int32_t ix = int32_t(cx), iy = int32_t(cy), iz = int32_t(cz);
    const float fx = x - cx, fy = y - cy, fz = z - cz;
    const uint32_t seedHash = Pcg(seed);
    float c[8];
    for (uint32_t corner = 0; corner < 8u; ++corner) {
        const uint32_t ox = corner & 1u;
        const uint32_t oy = (corner >> 1u) & 1u;
        const uint32_t oz = corner >> 2u;
        const uint32_t hash =
            Pcg3({uint32_t(ix) + ox, uint32_t(iy) + oy, uint32_t(iz) + oz + seedHash}).x;
        c[corner] = PerlinGrad3(hash, fx - float(ox), fy - float(oy), fz - float(oz));
    }
    const float u = Fade(fx), v = Fade(fy), w = Fade(fz);
    const float nearV = Mix(v, Mix(u, c[0], c[1]), Mix(u, c[2], c[3]));
    const float farV = Mix(v, Mix(u, c[4], c[5]), Mix(u, c[6], c[7]));
    return StdClamp(Mix(w, nearV, farV), -1.0f, 1.0f);
}

__device__ __forceinline__ float Fractal3Float(float x, float y, float z, uint32_t seed,
                                               int octaves) {
    const float kTurn[3][3] = {
        {-0.6132711010727254f, -0.293401570821183f, 0.7333580809050306f},
        {0.7897926788435601f, -0.24097777005594262f, 0.5640542870894418f},
        {0.011228581128535142f, 0.9251190369776896f, 0.37951111497202883f}};
    float turn[3][3] = {{1.0f, 0.0f, 0.0f}, {0.0f, 1.0f, 0.0f}, {0.0f, 0.0f, 1.0f}};
    float frequency = 1.0f;  // 1 / wavelength 1
    float weight = 1.0f, sum = 0.0f, total = 0.0f;
    for (int octave = 0; octave < octaves; ++octave) {
        const U3 h = Pcg3({seed, uint32_t(octave), 0u});
        const U3 g = Pcg3({seed, uint32_t(octave), 1u});
        const float rx = turn[0][0] * x + turn[0][1] * y + turn[0][2] * z;
        const float ry = turn[1][0] * x + turn[1][1] * y + turn[1][2] * z;
        const float rz = turn[2][0] * x + turn[2][1] * y + turn[2][2] * z;
        const float n = Perlin3(rx * frequency + UnitFloat(h.y), ry * frequency + UnitFloat(h.z),
                                rz * frequency + UnitFloat(g.x), h.x);
        sum += weight * n;
        total += weight;
        weight *= 0.5f;
        frequency *= 2.0f;
        float next[3][3];
        for (int r = 0; r < 3; ++r) {
            for (int col = 0; col < 3; ++col) {
                next[r][col] = kTurn[r][0] * turn[0][col] + kTurn[r][1] * turn[1][col] +
                               kTurn[r][2] * turn[2][col];
            }
        }
        for (int r = 0; r < 3; ++r) {
            for (int col = 0; col < 3; ++col) {
                turn[r][col] = next[r][col];
            }
        }
    }
    return 1.0f * sum / total;
}

// KernelRadialStreaks (craters.h:697-703).
__device__ __forceinline__ float RadialStreaks(float radii, float angle, float across, float along,
                                               int octaves, uint32_t seed) {
    const float ring = 1.0f / across;
    return Fractal3Float(ring * cosf(angle), ring * sinf(angle), radii / along, seed, octaves);
}

// KernelTextureHeight (craters.h:706-733).
__device__ __forceinline__ float TextureHeight(const tg_crater_shape& k, const tg_crater_record& c,
                                               float dx, float dy, float radii) {
    if (k.wall_streaks == 0.0f && k.ejecta == 0.0f) {
        return 0.0f;
    }
    const float freshness = c.fresh_depth > 0.0f ? c.current_depth / c.fresh_depth : 1.0f;
    const float fade = Smoothstep((freshness - k.min_freshness) / (1.0f - k.min_freshness));
    if (!(fade > 0.0f)) {
        return 0.0f;
    }
    const float angle = Angle(dy, dx);
    float height = 0.0f;
    if (k.wall_streaks > 0.0f && radii > 0.25f && radii < 1.0f) {
        const float band =
            Smoothstep((radii - 0.25f) / 0.35f) * Smoothstep((1.0f - radii) / 0.1f);
        height += k.wall_streaks * c.diameter * band *
                  RadialStreaks(radii, angle, k.wall_streak_width, 0.5f, 2, c.texture_key);
    }
    if (k.ejecta > 0.0f && radii > 1.0f && radii < k.ejecta_reach_radii) {
        const float radius = 0.5f * c.diameter;
        const float thickness = EjectaScale(radius) * EjectaShare(radii, k.ejecta_reach_radii);
        const float rise = Smoothstep((radii - 1.0f) / 0.1f);
        height += k.ejecta * thickness * rise *
                  RadialStreaks(radii, angle, k.ejecta_lineation, 2.0f * k.ejecta_lineation, 3,
                                c.texture_key + 1u);
    }
    return fade * height;
}

// KernelEllipseScale (craters.h:737-743).
__device__ __forceinline__ float EllipseScale(float theta, float ellipticity, float heading) {
    const float along = sqrtf(ellipticity);
    const float across = 1.0f / along;
    const float u = cosf(theta - heading) / along;
    const float v = sinf(theta - heading) / across;
    return 1.0f / sqrtf(u * u + v * v);
}

// KernelCatmullRom (craters.h:746-749).
__device__ __forceinline__ float CatmullRom(float p0, float p1, float p2, float p3, float t) {
    return 0.5f * (2.0f * p1 + (p2 - p0) * t + (2.0f * p0 - 5.0f * p1 + 4.0f * p2 - p3) * t * t +
                   (3.0f * (p1 - p2) + p3 - p0) * t * t * t);
}

// KernelMarksWindow (craters.h:753-759).
__device__ __forceinline__ float MarksWindow(float radii, float inner, float outer, float fade) {
    if (!(radii > inner - fade) || !(radii < outer + fade)) {
        return 0.0f;
    }
    return Smoothstep((radii - inner + fade) / fade) * Smoothstep((outer + fade - radii) / fade);
}

constexpr float kTableFineFade = 6.5f;

// KernelOutlineScaleAt (craters.h:772-808).
__device__ __forceinline__ float OutlineScaleAt(const float* data, const tg_crater_record& c,
                                                float dx, float dy, float share) {
    const float theta = Angle(dy, dx);
    const float scale =
        c.ellipticity != 1.0f ? EllipseScale(theta, c.ellipticity, c.heading) : 1.0f;
    float wobble = 1.0f;
    if (c.table_count > 0u && share > 0.0f) {
        const PeriodicCell cell = PeriodicAt(theta / kTwoPi, c.table_count);
        const float rim = (1.0f - cell.t) * data[c.table_offset + cell.index] +
                          cell.t * data[c.table_offset + (cell.index + 1u) % c.table_count];
        if (share >= 1.0f) {
            wobble += share * rim;
        } else {
            float broad = 0.0f;
            for (uint32_t n = 0u; n < c.harmonic_count; ++n) {
                const float order = float(c.first_harmonic + n);
                const float wave = data[c.harmonic_offset + 2u * n] *
                                   cosf(order * theta + data[c.harmonic_offset + 2u * n + 1u]);
                broad += wave;
                wobble += powf(share, 0.5f * order) * wave;
            }
            wobble += powf(share, kTableFineFade) * (rim - broad);
        }
    } else if (share > 0.0f) {
        for (uint32_t n = 0u; n < c.harmonic_count; ++n) {
            const float order = float(c.first_harmonic + n);
            const float kept = share < 1.0f ? powf(share, 0.5f * order) : 1.0f;
            wobble += kept * data[c.harmonic_offset + 2u * n] *
                      cosf(order * theta + data[c.harmonic_offset + 2u * n + 1u]);
        }
    }
    return scale * wobble;
}

// KernelWornOutlineShare (craters.h:826-828).
__device__ __forceinline__ float WornOutlineShare(const tg_crater_record& c) {
    return c.fresh_rim > 0.0f ? StdClamp(c.current_rim / c.fresh_rim, 0.0f, 1.0f) : 1.0f;
}

// KernelMarkShift (craters.h:832-848).
__device__ __forceinline__ float MarkShift(const float* data, const tg_crater_record& c, float dx,
                                           float dy, float radii) {
    const float window = MarksWindow(radii, data[c.mark_offset], data[c.mark_offset + 1u],
                                     data[c.mark_offset + 2u]);
    if (window == 0.0f) {
        return 0.0f;
    }
    const uint32_t size = c.mark_count;
    const PeriodicCell cell = PeriodicAt(Angle(dy, dx) / kTwoPi, size);
    const uint32_t index = cell.index;
    const uint32_t knots = c.mark_offset + 3u;
    return window * CatmullRom(data[knots + (index + size - 1u) % size], data[knots + index],
                               data[knots + (index + 1u) % size],
                               data[knots + (index + 2u) % size], cell.t);
}

constexpr uint32_t kTerraceHeader = 20u;
constexpr uint32_t kTerraceScarpValues = 8u;
constexpr uint32_t kTerraceStepValues = 3u;
constexpr float kTerraceBlockBlend = 0.3f;

// KernelTerraceBlockValue (craters.h:875-880).
__device__ __forceinline__ float TerraceBlockValue(uint32_t key, int cell, int count, float gaps) {
    const int wrapped = ((cell % count) + count) % count;
    const float hash = UnitFloat(Pcg(key + uint32_t(wrapped) * 0x9E3779B9u));
    return gaps < 0.0f ? hash : (hash >= gaps ? 1.0f : 0.0f);
}

// KernelTerraceBlocks (craters.h:884-902).
__device__ __forceinline__ float TerraceBlocks(const float* data, uint32_t s, float theta,
                                               uint32_t salt, float gaps) {
    const float turns = data[s + 2u] * theta / kTwoPi + data[s + 3u];
    const int count = int(data[s + 2u]);
    const uint32_t key = uint32_t(data[s + 7u]) ^ (salt * 0x85EBCA6Bu);
    const float cellStart = floorf(turns);
    const int cell = int(cellStart);
    const float u = turns - cellStart;
    const float value = TerraceBlockValue(key, cell, count, gaps);
    if (u < kTerraceBlockBlend) {
        const float b = Smoothstep(0.5f + 0.5f * u / kTerraceBlockBlend);
        return TerraceBlockValue(key, cell - 1, count, gaps) * (1.0f - b) + value * b;
    }
    if (u > 1.0f - kTerraceBlockBlend) {
        const float b = Smoothstep(0.5f + 0.5f * (1.0f - u) / kTerraceBlockBlend);
        return TerraceBlockValue(key, cell + 1, count, gaps) * (1.0f - b) + value * b;
    }
    return value;
}

// KernelHarmonics (craters.h:906-915).
__device__ __forceinline__ float Harmonics(const float* data, uint32_t offset, uint32_t first,
                                           uint32_t count, float theta, float turn) {
    float sum = 0.0f;
    for (uint32_t n = 0u; n < count; ++n) {
        const float order = float(first + n);
        sum += data[offset + 2u * n] * cosf(order * (theta + turn) + data[offset + 2u * n + 1u]);
    }
    return sum;
}

// KernelTerraceArc (craters.h:920-927).
__device__ __forceinline__ float TerraceArc(const float* data, uint32_t o, float theta,
                                            float step) {
    const float cut = data[o + 4u];
    if (!(cut > -1.5f)) {
        return 1.0f;
    }
    const float arcs = Harmonics(data, o + 6u, 2u, 5u, theta, 2.39996323f * step);
    return Smoothstep((arcs - cut) / data[o + 5u]);
}

// KernelTerraceBase (craters.h:931-939).
__device__ __forceinline__ float TerraceBase(const float* data, const tg_crater_record& c,
                                             uint32_t k) {
    if (k == 0u) {
        return 0.0f;
    }
    if (k >= c.terrace_count) {
        return 1.0f;
    }
    return data[c.terrace_offset + kTerraceHeader + kTerraceScarpValues * (k - 1u)];
}

// KernelTerraceScarp (craters.h:946-959).
__device__ __forceinline__ float TerraceScarp(const float* data, const tg_crater_record& c,
                                              uint32_t k, float theta) {
    if (k == 0u || k >= c.terrace_count) {
        return TerraceBase(data, c, k);
    }
    const uint32_t s = c.terrace_offset + kTerraceHeader + kTerraceScarpValues * (k - 1u);
    const float turns = data[s + 2u] * theta / kTwoPi + data[s + 3u];
    const float f = 2.0f * (turns - floorf(turns)) - 1.0f;
    const float scallop = data[s + 1u] * (1.0f - f * f - 2.0f / 3.0f);
    const float wobble = data[s + 4u] * cosf(3.0f * theta + data[s + 5u]);
    const float block = 2.0f * TerraceBlocks(data, s, theta, 0u, -1.0f) - 1.0f;
    return data[s] + scallop + wobble + data[s + 6u] * block;
}

// KernelTerraceRimScale (craters.h:964-970).
__device__ __forceinline__ float TerraceRimScale(const float* data, const tg_crater_record& c,
                                                 float theta) {
    const uint32_t o = c.terrace_offset;
    const float turns = data[o + 17u] * theta / kTwoPi + data[o + 18u];
    const float f = 2.0f * (turns - floorf(turns)) - 1.0f;
    return 1.0f + data[o + 16u] * (1.0f - f * f - 2.0f / 3.0f);
}

// KernelTerracedRadii (craters.h:986-1030).
__device__ __forceinline__ float TerracedRadii(const float* data, const tg_crater_record& c,
                                               float radii, float theta) {
    const uint32_t o = c.terrace_offset;
    const float inner = data[o];
    if (!(radii > inner) || !(radii < 1.0f)) {
        return radii;
    }
    const float t = (radii - inner) / (1.0f - inner);
    uint32_t step = 0u;
    float low = 0.0f;
    float high = TerraceScarp(data, c, 1u, theta);
    while (step + 1u < c.terrace_count && t >= high) {
        ++step;
        low = high;
        high = TerraceScarp(data, c, step + 1u, theta);
    }
    const float within = StdClamp((t - low) / (high - low), 0.0f, 1.0f);
    float strength = data[o + 2u] * TerraceArc(data, o, theta, float(step));
    if (c.terrace_count > 1u) {
        const bool top = step + 1u >= c.terrace_count;
        const uint32_t k = top ? step : step + 1u;
        const uint32_t s = o + kTerraceHeader + kTerraceScarpValues * (k - 1u);
        strength *= TerraceBlocks(data, s, theta, top ? 2u : 1u, data[o + 19u]);
    }
    const float rounding = data[o + 3u];
    const float slope = data[o + 1u];
    const uint32_t s = o + kTerraceHeader + kTerraceScarpValues * (c.terrace_count - 1u) +
                       kTerraceStepValues * step;
    const float level = data[s];
    const float scarp = data[s + 1u];
    strength *= data[s + 2u];
    const float lower = scarp * level;
    const float upper = scarp * (1.0f - level);
    float stair = level;
    if (within < lower) {
        stair = level * (1.0f - TerraceEase(1.0f - within / lower, rounding));
    } else if (within > 1.0f - upper) {
        stair = level + (1.0f - level) * TerraceEase((within - (1.0f - upper)) / upper, rounding);
    }
    const float rise = slope * within + (1.0f - slope) * stair;
    const float shaped = within + strength * (rise - within);
    const float from = TerraceBase(data, c, step);
    const float to = TerraceBase(data, c, step + 1u);
    return inner + (from + shaped * (to - from)) * (1.0f - inner);
}

// KernelOutlineScaleAt for an outline of harmonics and an ellipse only (no rim table), without
// inverse or forward trigonometry per harmonic: theta's cosine and sine are dx / d and dy / d, the
// harmonics' cos(k theta + phi_k) follow from rotating (cos k theta, sin k theta) by theta and the
// crater's precomputed (A_k cos phi_k, A_k sin phi_k) in `harm` (same slots as its data), and the
// kept shares share^(k / 2) are powers of sqrt(share) by repeated multiplication. Equal to the
// reference's form to within FP32 rounding of a dozen products. `ra` holds the crater's
// cos(heading), sin(heading), sqrt(ellipticity) and its reciprocal; `harm` is padded by
// kHarmonicBlock floats past its end.
constexpr uint32_t kHarmonicBlock = 16u;

// The fast outline over `coef`, the crater's (A cos phi, A sin phi) pairs from harmonic
// `first`, `count` of them, padded by kHarmonicBlock floats.
__device__ __forceinline__ float FastOutline(const float* coef, float4 ra, float ellipticity,
                                             uint32_t first, uint32_t count, float dx, float dy,
                                             float d, float share) {
    float ct = 1.0f, st = 0.0f;
    if (d > 0.0f) {
        const float inv = 1.0f / d;
        ct = dx * inv;
        st = dy * inv;
    }
    float scale = 1.0f;
    if (ellipticity != 1.0f) {
        const float cd = ct * ra.x + st * ra.y;  // cos(theta - heading)
        const float sd = st * ra.x - ct * ra.y;  // sin(theta - heading)
        const float u = cd / ra.z;
        const float v = sd / ra.w;
        scale = 1.0f / sqrtf(u * u + v * v);
    }
    float wobble = 1.0f;
    if (share > 0.0f && count > 0u) {
        const bool partial = share < 1.0f;
        const float q = partial ? sqrtf(share) : 1.0f;
        float ck = 1.0f, sk = 0.0f, kept = 1.0f;
        for (uint32_t k = 0; k < first; ++k) {
            const float next = ck * ct - sk * st;
            sk = sk * ct + ck * st;
            ck = next;
            kept *= q;
        }
        // The coefficients are read kHarmonicBlock floats at a time (one wide scalar load per
        // block rather than one per harmonic); the table is padded by a block, so the last block
        // stays in bounds. The sums are the same, in the same order.
        for (uint32_t base = 0u; base < count; base += kHarmonicBlock / 2u) {
            float block[kHarmonicBlock];
#pragma unroll
            for (uint32_t i = 0u; i < kHarmonicBlock; ++i) {
                block[i] = coef[2u * base + i];
            }
#pragma unroll
            for (uint32_t j = 0u; j < kHarmonicBlock / 2u; ++j) {
                if (base + j >= count) {
                    break;
                }
                const float wave = block[2u * j] * ck - block[2u * j + 1u] * sk;
                wobble += (partial ? kept : 1.0f) * wave;
                const float next = ck * ct - sk * st;
                sk = sk * ct + ck * st;
                ck = next;
                kept *= q;
            }
        }
    }
    return scale * wobble;
}

__device__ __forceinline__ float FastOutlineScale(const float* harm, float4 ra,
                                                  const tg_crater_record& c, float dx, float dy,
                                                  float d, float share) {
    return FastOutline(harm + c.harmonic_offset, ra, c.ellipticity, c.first_harmonic,
                       c.harmonic_count, dx, dy, d, share);
}

// KernelProfileDistance (craters.h:1034-1054).
__device__ __forceinline__ float ProfileDistance(const float* data, const tg_crater_record& c,
                                                 float dx, float dy, const float* harm = nullptr,
                                                 float4 ra = float4{1.0f, 0.0f, 1.0f, 1.0f}) {
    float distance = sqrtf(dx * dx + dy * dy);
    if (c.ellipticity != 1.0f || c.harmonic_count > 0u || c.table_count > 0u ||
        c.mark_count > 0u) {
        const float share = Smoothstep(distance / (0.5f * c.diameter)) * WornOutlineShare(c);
        if (harm != nullptr && c.table_count == 0u) {
            distance /= FastOutlineScale(harm, ra, c, dx, dy, distance, share);
        } else {
            distance /= OutlineScaleAt(data, c, dx, dy, share);
        }
    }
    if (c.mark_count > 0u) {
        const float radius = 0.5f * c.diameter;
        distance += radius * MarkShift(data, c, dx, dy, distance / radius);
    }
    if (c.terrace_count > 0u) {
        const float radius = 0.5f * c.diameter;
        const float theta = Angle(dy, dx);
        const float scale = TerraceRimScale(data, c, theta);
        distance /= 1.0f + (scale - 1.0f) * Smoothstep(distance / radius);
        distance = radius * TerracedRadii(data, c, distance / radius, theta);
    }
    return distance;
}

// KernelSlopeShiftWeight (craters.h:1058-1061).
__device__ __forceinline__ float SlopeShiftWeight(const tg_crater_record& c, float dx, float dy) {
    const float radii = sqrtf(dx * dx + dy * dy) / (0.5f * c.diameter);
    return 1.0f - Smoothstep((radii - 0.3f) / 0.7f);
}

// KernelSlopeRaisedShare (craters.h:1065-1073).
__device__ __forceinline__ float SlopeRaisedShare(const tg_crater_record& c, float dx, float dy) {
    const float distance = sqrtf(dx * dx + dy * dy);
    if (!(distance > 0.0f)) {
        return 1.0f;
    }
    const float removed = sqrtf(c.upslope_x * c.upslope_x + c.upslope_y * c.upslope_y);
    const float cosine =
        StdMax(0.0f, (c.upslope_x * dx + c.upslope_y * dy) / (removed * distance));
    return 1.0f - removed * cosine * cosine;
}

// KernelShiftedDistance (craters.h:1077-1081).
__device__ __forceinline__ float ShiftedDistance(const float* data, const tg_crater_record& c,
                                                 float dx, float dy, const float* harm = nullptr,
                                                 float4 ra = float4{1.0f, 0.0f, 1.0f, 1.0f}) {
    const float shift =
        c.shift_x != 0.0f || c.shift_y != 0.0f ? SlopeShiftWeight(c, dx, dy) : 0.0f;
    return ProfileDistance(data, c, dx - shift * c.shift_x, dy - shift * c.shift_y, harm, ra);
}

// KernelUpslopeHeight (craters.h:1085-1091).
__device__ __forceinline__ float UpslopeHeight(const tg_crater_record& c, float height, float dx,
                                               float dy) {
    if (height > 0.0f && (c.upslope_x != 0.0f || c.upslope_y != 0.0f)) {
        return height * SlopeRaisedShare(c, dx, dy);
    }
    return height;
}

constexpr uint32_t kPeakHeader = 6u;
constexpr uint32_t kPeakMassifValues = 11u;
constexpr float kPeakReach = 2.6f;

// KernelPeakMassif (craters.h:1104-1123).
__device__ __forceinline__ float PeakMassif(const float* data, uint32_t m, float px, float py) {
    const float ex = px - data[m];
    const float ey = py - data[m + 1u];
    const float radius = data[m + 2u];
    const float stretch = data[m + 3u];
    const float cosine = data[m + 4u];
    const float sine = data[m + 5u];
    const float qx = (ex * cosine + ey * sine) / (radius * stretch);
    const float qy = (ey * cosine - ex * sine) * stretch / radius;
    const float squared = qx * qx + qy * qy;
    if (!(squared < 1.0f)) {
        return 0.0f;
    }
    const float pull = 1.0f - squared;
    const float sx = qx - data[m + 8u] * pull;
    const float sy = qy - data[m + 9u] * pull;
    const float r = StdMin(sqrtf(sx * sx + sy * sy), 1.0f);
    const float flank = 1.0f - powf(r, data[m + 7u]);
    return data[m + 6u] * powf(flank, data[m + 10u]);
}

// KernelPeakMassifs (craters.h:1128-1155).
__device__ __forceinline__ float PeakMassifs(const float* data, const tg_crater_record& c,
                                             PeakSize size, float dx, float dy) {
    const uint32_t o = c.peak_offset;
    float px = dx / size.base;
    float py = dy / size.base;
    const float roughness = data[o];
    const float wavelength = data[o + 1u];
    const float warp = data[o + 2u];
    const uint32_t key = uint32_t(data[o + 3u]);
    if (warp > 0.0f) {
        const float wx = KernelFractal2(px, py, key + 1u, 2.0f * wavelength, 3, 0.5f,
                                        TG_FRACTAL_FBM);
        const float wy = KernelFractal2(px + 17.31f, py - 9.73f, key + 1u, 2.0f * wavelength, 3,
                                        0.5f, TG_FRACTAL_FBM);
        px += warp * wx;
        py += warp * wy;
    }
    float height = 0.0f;
    for (uint32_t n = 0u; n < c.peak_count; ++n) {
        height = StdMax(height, PeakMassif(data, o + kPeakHeader + kPeakMassifValues * n, px, py));
    }
    if (roughness > 0.0f && height > 0.0f) {
        height *= 1.0f + roughness * KernelFractal2(px, py, key, wavelength, 5, 0.5f,
                                                    TG_FRACTAL_RIDGED);
    }
    return size.height * height;
}

// KernelPeakRelief (craters.h:1161-1182).
__device__ __forceinline__ float PeakRelief(const tg_crater_shape& k, const float* data,
                                            const tg_crater_record& c, float distance, float dx,
                                            float dy) {
    if (c.peak_count == 0u || k.profile.complex_ != 1u || k.profile.diffused == 1u ||
        k.profile.simple == TG_PROFILE_TABULATED || c.diameter < k.profile.transition_diameter) {
        return 0.0f;
    }
    const PeakSize size = ComplexPeak(k.profile, c.diameter);
    PeakSize scaled = size;
    scaled.height *= data[c.peak_offset + 4u];
    scaled.base *= data[c.peak_offset + 5u];
    const float shift =
        c.shift_x != 0.0f || c.shift_y != 0.0f ? SlopeShiftWeight(c, dx, dy) : 0.0f;
    const float sx = dx - shift * c.shift_x;
    const float sy = dy - shift * c.shift_y;
    const float reach = kPeakReach * StdMax(size.base, scaled.base);
    if (!(sx * sx + sy * sy < reach * reach)) {
        return 0.0f;
    }
    return KeptDepth(c.fresh_depth, c.current_depth) *
           (PeakMassifs(data, c, scaled, sx, sy) - PeakDome(size, distance));
}

// ApplyCraterKernel (craters.h:1186-1212): `height` after crater `c` at (dx, dy) from its center.
__device__ __forceinline__ float ApplyCrater(const CraterShape& shape, const float* data,
                                             const tg_crater_record& c, float height, float dx,
                                             float dy, const float* harm = nullptr,
                                             float4 ra = float4{1.0f, 0.0f, 1.0f, 1.0f},
                                             float resolution = -1.0f) {
    const tg_crater_shape& k = shape.k;
    const float diameter = c.diameter;
    const float radius = 0.5f * diameter;
    const float distance = ShiftedDistance(data, c, dx, dy, harm, ra);
    if (distance >= k.reach_radii * radius) {
        return height;
    }
    const float radii = distance / radius;
    float replaced = 0.0f;
    if (c.has_datum == 1u) {
        float weight = c.replace == 1u
                           ? Falloff(radii, k.replace_start, k.replace_end, k.blend_shape)
                           : k.interior_replacement *
                                 Falloff(radii, k.blend_start, k.blend_end, k.blend_shape);
        if (k.mantling > 0.0f && radii > 1.0f) {
            weight = StdMax(weight, k.mantling * EjectaShare(radii, k.ejecta_reach_radii));
        }
        replaced = weight * (c.datum_a + c.datum_b * dx + c.datum_c * dy - height);
    }
    CraterView view{c,
                    {c.diameter, c.fresh_depth, c.fresh_rim, c.current_depth, c.current_rim},
                    {c.diameter, c.fresh_depth, c.fresh_rim, c.fresh_depth, c.fresh_rim}};
    const float bare = UpslopeHeight(c, RoundedHeight(k, data, distance, diameter, view), dx, dy) +
                       PeakRelief(k, data, c, distance, dx, dy);
    const float shaped =
        bare + NoiseHeight(k, c, dx, dy, radii) + TextureHeight(k, c, dx, dy, radii);
    // The record's ResolutionWeight, or the caller's for the grid it shapes (a catalogue record
    // serves grids of every spacing).
    const float weight = resolution >= 0.0f ? resolution : c.resolution;
    return height + weight * shaped + weight * replaced;
}

// ---- The hoisted path for the Martin-Parkes-Dunstan profile ------------------------------------
//
// When only the simple MPD profile and the outline shape a crater (no diffusion, complex form, rim
// rounding, crater noise or texture in the kernel block; no datum, marks, terraces, peaks, floor
// shift, upslope or outline table in the record), most of ApplyCrater's work depends on the crater
// alone. MpdTerms holds those values, computed once per crater by the same FP32 operations in the
// same order as MpdHeight (craters.h:388-416), MartinParkesDunstanExterior (craters.h:374-386) and
// KernelProfileDistance (craters.h:1034-1054) perform them, so ApplyCraterMpd gives the same bits
// as ApplyCrater for such a crater.

// MpdTerms::fast: 0 for the full evaluation, else the hoisted path in FP32 (kFastSingle) or, for a
// crater of kPreciseDiameter or more, in FP64 (kFastPrecise, PreciseTerms below).
constexpr uint32_t kFastSingle = 1u;
constexpr uint32_t kFastPrecise = 2u;
// FP32 carries a relative 6e-8: for a crater of diameter D the distances, depths and rim heights
// of one pair round to about 1.5e-7 D of height (2e-3 m at 20 km, 1.5e-4 m at 1 km, measured
// against the FP64 oracle with the reference's arithmetic). Pairs with craters of a kilometre and
// more are evaluated in FP64, which keeps a cell within the 1e-3 m bound wherever km-scale
// craters overlap.
constexpr float kPreciseDiameter = 1000.0f;

struct MpdTerms {
    float radius;     // 0.5 D
    float reach;      // reach radii x radius
    float worn;       // KernelWornOutlineShare
    float freshRim;   // fresh rim x D
    float rim;        // current rim x D
    float k;          // freshRim - rim + depth
    float k2;         // 2 k
    float alpha;      // (rim - freshRim) / k
    float interior;   // k (alpha + 1) / alpha
    float beta;       // min(1, 1.5 (freshRim - rim) / k)
    float ejecta;     // the ejecta scale
    float joinRise;   // the exterior law at beta, less the rim
    float betaSlope;  // beta x the joint's slope
    float offset;     // rim - ejecta
    float gradient;   // -ejecta / 8 - offset
    uint32_t fast;    // 1 when the hoisted path applies to this crater
};
static_assert(sizeof(MpdTerms) == 64, "MpdTerms is four float4");

// The kernel block's part of the condition.
__host__ __device__ inline bool MpdShapeOnly(const tg_crater_shape& k) {
    return k.profile.simple == TG_PROFILE_MPD && k.profile.diffused == 0u &&
           k.profile.complex_ == 0u && k.rim_sigma == 0.0f && k.noise_amplitude == 0.0f &&
           k.noise_rim_amplitude == 0.0f && k.wall_streaks == 0.0f && k.ejecta == 0.0f;
}

// The record's part of the condition: nothing the hoisted path leaves out.
__host__ __device__ inline bool MpdRecordFast(const tg_crater_record& c) {
    return c.has_datum != 1u && c.mark_count == 0u && c.terrace_count == 0u &&
           c.peak_count == 0u && c.shift_x == 0.0f && c.shift_y == 0.0f && c.upslope_x == 0.0f &&
           c.upslope_y == 0.0f && c.table_count == 0u;
}

__device__ inline MpdTerms MakeMpdTerms(const tg_crater_shape& k, const tg_crater_record& c) {
    MpdTerms t{};
    const float diameter = c.diameter;
    const float depth = c.current_depth * diameter;
    t.rim = c.current_rim * diameter;
    t.freshRim = c.fresh_rim * diameter;
    t.radius = 0.5f * diameter;
    t.reach = k.reach_radii * t.radius;
    t.worn = WornOutlineShare(c);
    t.k = t.freshRim - t.rim + depth;
    t.k2 = 2.0f * t.k;
    t.alpha = (t.rim - t.freshRim) / t.k;
    t.interior = t.k * (t.alpha + 1.0f) / t.alpha;
    t.ejecta = t.freshRim > 0.0f ? EjectaScale(t.radius) * t.rim / t.freshRim
                                 : EjectaScale(t.radius);
    t.beta = StdMin(1.0f, 1.5f * (t.freshRim - t.rim) / t.k);
    const ProfilePoint join = MpdExterior(t.beta, t.rim, t.ejecta);
    const float drop = t.rim - join.height;
    const float slope = drop > 0.0f ? StdMax(join.slope, -3.0f * drop / t.beta) : join.slope;
    t.joinRise = join.height - t.rim;
    t.betaSlope = t.beta * slope;
    t.offset = t.rim - t.ejecta;
    t.gradient = -t.ejecta / 8.0f - t.offset;
    t.fast = MpdShapeOnly(k) && MpdRecordFast(c)
                 ? (c.diameter >= kPreciseDiameter ? kFastPrecise : kFastSingle)
                 : 0u;
    return t;
}

// MpdHeight at `distance` from the hoisted terms.
__device__ __forceinline__ float MpdHeightFrom(const MpdTerms& t, float distance) {
    const float x = distance / t.radius - 1.0f;
    if (x >= 1.0f) {
        return 0.0f;
    }
    if (x <= t.alpha) {
        return (t.k * x + t.k2) * x + t.freshRim;
    }
    if (x <= 0.0f) {
        return t.interior * x * x + t.rim;
    }
    if (x <= t.beta) {
        const float u = x / t.beta;
        return t.rim + t.joinRise * u * u * (3.0f - 2.0f * u) + t.betaSlope * u * u * (u - 1.0f);
    }
    const float r = 1.0f / (x + 1.0f);
    const float law = t.ejecta * r * r * r + t.offset + t.gradient * x;
    const float taper = 1.0f - x * x;
    return law * taper;
}

// ApplyCrater for a crater whose terms say `fast`, with the outline's harmonic terms in `harm`
// and `ra` as FastOutlineScale takes them, and the resolution weight for the grid.
// The hoisted MPD crater from its terms, outline (ellipticity, harmonics from `first`, `count`
// of them in `coef`, `ra`) and resolution weight.
__device__ __forceinline__ float ApplyMpd(const MpdTerms& t, float ellipticity, uint32_t first,
                                          uint32_t count, const float* coef, float4 ra,
                                          float height, float dx, float dy, float weight) {
    float distance = sqrtf(dx * dx + dy * dy);
    if (ellipticity != 1.0f || count > 0u) {
        const float share = Smoothstep(distance / t.radius) * t.worn;
        distance /= FastOutline(coef, ra, ellipticity, first, count, dx, dy, distance, share);
    }
    if (distance >= t.reach) {
        return height;
    }
    // The same sums as ApplyCrater's, whose other terms are zero here.
    const float bare = MpdHeightFrom(t, distance) + 0.0f;
    const float shaped = bare + 0.0f + 0.0f;
    return height + weight * shaped + weight * 0.0f;
}

// A crater as the batched kernel's fast path reads it, in one 128-byte line: the split centre and
// squared cull radius its walks test, the outline's scalars, the FP64 diameter of the resolution
// fade and the hoisted MPD terms. The first kHotHarmonics outline harmonics sit in a slot of a
// parallel table at the same index, so a visit's loads do not wait on one another.
struct HotCrater {
    float x, y, x_low, y_low;
    float cull;
    float ellipticity;
    uint32_t first_harmonic;
    uint32_t harmonic_count;
    double diameter;
    uint32_t harmonic_offset;  // into the padded harmonic table, for outlines longer than a slot
    uint32_t pad;
    float4 ra;                 // cos and sin of the heading, sqrt(ellipticity) and its reciprocal
    MpdTerms terms;
};
static_assert(sizeof(HotCrater) == 128, "HotCrater is one 128-byte line");

// Harmonics a crater's slot of the hot harmonic table holds: 32 floats, 128 bytes.
constexpr uint32_t kHotHarmonics = 16u;

__device__ __forceinline__ float ApplyCraterMpd(const tg_crater_record& c, const MpdTerms& t,
                                                float height, float dx, float dy,
                                                const float* harm, float4 ra, float weight) {
    return ApplyMpd(t, c.ellipticity, c.first_harmonic, c.harmonic_count,
                    harm + c.harmonic_offset, ra, height, dx, dy, weight);
}

// ---- The hoisted MPD path in FP64, for craters of kPreciseDiameter and more ---------------------
//
// The same profile law, outline and blend as ApplyMpd, with every per-pair quantity in FP64: the
// cell's offset from the split cell point and split crater centre (on a body, in the crater's
// frame with FP64 axes), the distance, the outline scale and the height. The record fields the
// terms start from are the FP32 ones the FP64 oracle widens, so only rounding differs from it.
// A cell's precise contributions are summed in FP64 (PreciseSum) and reach its FP32 height once,
// since a km-scale height rounds to 1e-4 m or more on every FP32 store.

struct PreciseTerms {
    double invRadius, reach, worn, freshRim, rim, k, alpha, interior, beta, invBeta, ejecta,
        joinRise, betaSlope, offset, gradient;
    double cosHeading, sinHeading, along, alongInv;  // the outline's ellipse
    double centreX, centreY, centreZ;  // the split centre, summed (z 0 on a plane)
    double eastX, eastY, northX, northY, northZ;     // on a body: the frame's axes (east z = 0)
    double halfInvSphere;              // on a body: 0.5 / the sphere's radius
    double pad[4];
};
static_assert(sizeof(PreciseTerms) == 256, "PreciseTerms is two 128-byte lines");

// 1 / v to within an ulp or so: the hardware estimate and two Newton steps.
__device__ __forceinline__ double Reciprocal64(double v) {
    double r = __builtin_amdgcn_rcp(v);
    r = fma(r, fma(-v, r, 1.0), r);
    return fma(r, fma(-v, r, 1.0), r);
}

// 1 / sqrt(v) for a positive normal v: the hardware estimate and two Newton steps.
__device__ __forceinline__ double InvSqrt64(double v) {
    double y = __builtin_amdgcn_rsq(v);
    const double half = 0.5 * v;
    y = y * fma(-half * y, y, 1.5);
    return y * fma(-half * y, y, 1.5);
}

// sqrt(v) for v >= 0 to within an ulp or so: v / sqrt(v) from InvSqrt64, then one correction.
__device__ __forceinline__ double Sqrt64(double v) {
    if (!(v > 0.0)) {
        return 0.0;
    }
    const double y = InvSqrt64(v);
    const double s = v * y;
    return fma(fma(-s, s, v), 0.5 * y, s);
}

__device__ __forceinline__ double Smoothstep64(double t) {
    const double u = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
    return u * u * (3.0 - 2.0 * u);
}

// MartinParkesDunstanExterior (craters.h:374-386) in FP64.
__device__ __forceinline__ void MpdExterior64(double x, double rim, double ejectaScale,
                                              double& height, double& slope) {
    const double offset = rim - ejectaScale;
    const double gradient = -ejectaScale / 8.0 - offset;
    const double r = 1.0 / (x + 1.0);
    const double law = ejectaScale * r * r * r + offset + gradient * x;
    const double lawSlope = -3.0 * ejectaScale * r * r * r * r + gradient;
    const double taper = 1.0 - x * x;
    height = law * taper;
    slope = lawSlope * taper - 2.0 * x * law;
}

// The terms of crater c (MakeMpdTerms's, in FP64 from the same record fields), its ellipse and,
// with a frame, FP64 east and north axes from the frame's split centre.
__device__ inline PreciseTerms MakePreciseTerms(const tg_crater_shape& k, const tg_crater_record& c,
                                                const tg_crater_frame* frame) {
    PreciseTerms t{};
    const double diameter = c.diameter;
    const double radius = 0.5 * diameter;
    const double depth = double(c.current_depth) * diameter;
    t.rim = double(c.current_rim) * diameter;
    t.freshRim = double(c.fresh_rim) * diameter;
    t.invRadius = 1.0 / radius;
    t.reach = double(k.reach_radii) * radius;
    t.worn = c.fresh_rim > 0.0f ? fmin(fmax(double(c.current_rim) / double(c.fresh_rim), 0.0), 1.0)
                                : 1.0;
    t.k = t.freshRim - t.rim + depth;
    t.alpha = (t.rim - t.freshRim) / t.k;
    t.interior = t.k * (t.alpha + 1.0) / t.alpha;
    const double scale = 0.14 * pow(radius, 0.74);
    t.ejecta = t.freshRim > 0.0 ? scale * t.rim / t.freshRim : scale;
    t.beta = fmin(1.0, 1.5 * (t.freshRim - t.rim) / t.k);
    t.invBeta = 1.0 / t.beta;
    double joinHeight, joinSlope;
    MpdExterior64(t.beta, t.rim, t.ejecta, joinHeight, joinSlope);
    const double drop = t.rim - joinHeight;
    const double slope = drop > 0.0 ? fmax(joinSlope, -3.0 * drop / t.beta) : joinSlope;
    t.joinRise = joinHeight - t.rim;
    t.betaSlope = t.beta * slope;
    t.offset = t.rim - t.ejecta;
    t.gradient = -t.ejecta / 8.0 - t.offset;
    t.cosHeading = cos(double(c.heading));
    t.sinHeading = sin(double(c.heading));
    t.along = sqrt(double(c.ellipticity));
    t.alongInv = 1.0 / t.along;
    t.centreX = double(c.x) + double(c.x_low);
    t.centreY = double(c.y) + double(c.y_low);
    if (frame != nullptr) {
        const double cx = double(frame->x) + double(frame->x_low);
        const double cy = double(frame->y) + double(frame->y_low);
        const double cz = double(frame->z) + double(frame->z_low);
        t.centreX = cx;
        t.centreY = cy;
        t.centreZ = cz;
        t.halfInvSphere = frame->sphere_radius > 0.0f ? 0.5 / double(frame->sphere_radius) : 0.0;
        const double rho = sqrt(cx * cx + cy * cy);
        const double full = sqrt(rho * rho + cz * cz);
        if (rho > 1e-6 * full) {
            t.eastX = -cy / rho;
            t.eastY = cx / rho;
            t.northX = -cz * cx / (full * rho);
            t.northY = -cz * cy / (full * rho);
            t.northZ = rho / full;
        } else {  // at a pole the centre fixes no longitude: the frame's own axes
            t.eastX = frame->east_x;
            t.eastY = frame->east_y;
            t.northX = frame->north_x;
            t.northY = frame->north_y;
            t.northZ = frame->north_z;
        }
    }
    return t;
}

// MpdHeightFrom in FP64.
__device__ __forceinline__ double MpdHeight64(const PreciseTerms& t, double distance) {
    const double x = distance * t.invRadius - 1.0;
    if (x >= 1.0) {
        return 0.0;
    }
    if (x <= t.alpha) {
        return fma(fma(t.k, x, 2.0 * t.k), x, t.freshRim);
    }
    if (x <= 0.0) {
        return fma(t.interior * x, x, t.rim);
    }
    if (x <= t.beta) {
        const double u = x * t.invBeta;
        const double u2 = u * u;
        return fma(t.betaSlope * u2, u - 1.0, fma(t.joinRise * u2, fma(-2.0, u, 3.0), t.rim));
    }
    const double r = Reciprocal64(x + 1.0);
    const double law = fma(t.ejecta * r, r * r, fma(t.gradient, x, t.offset));
    return law * fma(-x, x, 1.0);
}

// The offset of a cell from a crater centre: its distance and direction cosines (1, 0 at the
// centre).
struct Offset64 {
    double distance, ct, st;
};

// FastOutline in FP64 over the same (A cos phi, A sin phi) coefficients. cos k theta and
// sin k theta follow the three-term recurrence x_(k+1) = 2 cos theta x_k - x_(k-1), which for
// the dozen or so harmonics of an outline stays within a few ulps in FP64.
__device__ __forceinline__ double FastOutline64(const float* coef, const PreciseTerms& t,
                                                float ellipticity, uint32_t first, uint32_t count,
                                                const Offset64& o, double share) {
    const double ct = o.ct, st = o.st;
    double scale = 1.0;
    if (ellipticity != 1.0f) {
        const double cd = fma(ct, t.cosHeading, st * t.sinHeading);
        const double sd = fma(st, t.cosHeading, -(ct * t.sinHeading));
        const double u = cd * t.alongInv;
        const double v = sd * t.along;
        scale = InvSqrt64(fma(u, u, v * v));
    }
    double wobble = 1.0;
    if (share > 0.0 && count > 0u) {
        const bool partial = share < 1.0;
        const double q = partial ? Sqrt64(share) : 1.0;
        const double twoCos = 2.0 * ct;
        // (cos, sin) of (k - 1) theta and k theta, from k = 0.
        double cPrev = ct, sPrev = -st, ck = 1.0, sk = 0.0, kept = 1.0;
        for (uint32_t k = 0; k < first; ++k) {
            const double cNext = fma(twoCos, ck, -cPrev);
            const double sNext = fma(twoCos, sk, -sPrev);
            cPrev = ck;
            sPrev = sk;
            ck = cNext;
            sk = sNext;
            kept *= q;
        }
        for (uint32_t j = 0; j < count; ++j) {
            const double wave = fma(double(coef[2u * j]), ck, -(double(coef[2u * j + 1u]) * sk));
            wobble = fma(partial ? kept : 1.0, wave, wobble);
            const double cNext = fma(twoCos, ck, -cPrev);
            const double sNext = fma(twoCos, sk, -sPrev);
            cPrev = ck;
            sPrev = sk;
            ck = cNext;
            sk = sNext;
            kept *= q;
        }
    }
    return scale * wobble;
}

// ApplyMpd's contribution in FP64 at offset o, before the resolution weight: 0 beyond the
// outline-scaled reach.
__device__ __forceinline__ double MpdPrecise(const PreciseTerms& t, float ellipticity,
                                             uint32_t first, uint32_t count, const float* coef,
                                             const Offset64& o) {
    double distance = o.distance;
    if (ellipticity != 1.0f || count > 0u) {
        const double share = Smoothstep64(distance * t.invRadius) * t.worn;
        distance *= Reciprocal64(FastOutline64(coef, t, ellipticity, first, count, o, share));
    }
    if (distance >= t.reach) {
        return 0.0;
    }
    return MpdHeight64(t, distance);
}

// The plane offset of the split cell point (x + xLow, y + yLow) from the crater's centre.
__device__ __forceinline__ Offset64 PlaneOffset64(const PreciseTerms& t, float x, float xLow,
                                                  float y, float yLow) {
    const double dx = (double(x) + double(xLow)) - t.centreX;
    const double dy = (double(y) + double(yLow)) - t.centreY;
    Offset64 o{Sqrt64(dx * dx + dy * dy), 1.0, 0.0};
    if (o.distance > 0.0) {
        const double inv = Reciprocal64(o.distance);
        o.ct = dx * inv;
        o.st = dy * inv;
    }
    return o;
}

// KernelFrameOffset (craters.h:168-186) in FP64: the split cell point's offset in the crater's
// frame, its length the arc along the sphere rather than the chord.
__device__ __forceinline__ Offset64 BodyOffset64(const PreciseTerms& t, const tg_split_point& p) {
    const double qx = (double(p.x) + double(p.x_low)) - t.centreX;
    const double qy = (double(p.y) + double(p.y_low)) - t.centreY;
    const double qz = (double(p.z) + double(p.z_low)) - t.centreZ;
    const double ox = qx * t.eastX + qy * t.eastY;
    const double oy = qx * t.northX + qy * t.northY + qz * t.northZ;
    const double chord = Sqrt64(qx * qx + qy * qy + qz * qz);
    const double s = fmin(chord * t.halfInvSphere, 1.0);
    // asin(s) / s by its series below 0.25. Above (an offset of more than half the radius), the
    // angle is halved until its sine is below 0.25 (sin(a/2) = sin a / sqrt(2 (1 + cos a))),
    // and asin(s) = 2^k asin(t_k): a short loop rather than atan2, whose polynomial's constants
    // the compiler keeps in registers across every kernel that inlines this.
    const auto series = [](double u) {
        const double u2 = u * u;
        return 1.0 + u2 * (1.0 / 6.0 +
                           u2 * (3.0 / 40.0 + u2 * (5.0 / 112.0 + u2 * (35.0 / 1152.0 +
                                                                       u2 * (63.0 / 2816.0)))));
    };
    double ratio;
    if (s < 0.25) {
        ratio = series(s);
    } else {
        double u = s, scale = 1.0;
#pragma unroll 1
        while (u >= 0.25) {
            u *= InvSqrt64(2.0 * (1.0 + Sqrt64(1.0 - u * u)));
            scale *= 2.0;
        }
        ratio = scale * u * series(u) / s;
    }
    Offset64 o{chord * ratio, 0.0, 1.0};
    const double across2 = ox * ox + oy * oy;
    if (across2 > 0.0) {
        // The direction cosines take 1 / |(ox, oy)| alone, so its inverse square root serves.
        const double inv = InvSqrt64(across2);
        o.ct = ox * inv;
        o.st = oy * inv;
    }
    return o;
}

// A cell's height as its craters are applied: the FP32 height of the reference's arithmetic and
// the FP64 sum of the precise craters not yet added to it. Both the batched and the strip kernels
// fold the sum in before a crater whose evaluation reads the height, and at the end.
struct PreciseSum {
    float h;
    double pending = 0.0;

    __device__ __forceinline__ float Folded() {
        if (pending != 0.0) {
            h = float(double(h) + pending);
            pending = 0.0;
        }
        return h;
    }
};

}  // namespace tg
