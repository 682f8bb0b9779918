// SPDX-License-Identifier: BSD-3-Clause
//
// ORACLE — test infrastructure only (see noise_oracle.cpp). The crater kernel
// (include/terraingen/layers/kernels/craters.h:150-1212 of the reference) restated for the CPU,
// with the precision of its per-pair arithmetic as the macro ORC_R: this header is included once
// with float (the reference-equivalent path, pinned bit for bit against oracle/_ref, the
// reference's own header compiled beside it) and once with double (the FP64 oracle: records, data
// and constants as the reference stores them in FP32, every operation on them in FP64).
// Included only by crater_oracle.cpp.

namespace ORC_NS {

using R = ORC_R;
using u32 = std::uint32_t;
using uint32_t = std::uint32_t;

inline R StdMax(R a, R b) {
    return a < b ? b : a;
}
inline R StdMin(R a, R b) {
    return b < a ? b : a;
}
inline R StdClamp(R v, R lo, R hi) {
    return v < lo ? lo : (hi < v ? hi : v);
}
inline R Clamp01(R v) {
    return StdClamp(v, R(0), R(1));
}

// The reference's records with every FP32 field widened to R.
struct Prof {
    u32 simple, rim, complex_, diffused;
    R transition_diameter, floor_coefficient, floor_exponent, peak_height_coefficient,
        peak_height_exponent, peak_width_coefficient, peak_width_exponent;
    u32 peak_from_volume;
    R peak_fit_max_diameter, ejecta_exponent, complex_reach_radii, table_extent;
    u32 table_cells, ages_offset, age_count;
};

struct Shp {
    Prof profile;
    R reach_radii, profile_reach_radii, blend_start, blend_end;
    u32 blend_shape;
    R interior_replacement, replace_start, replace_end, mantling, ejecta_reach_radii, rim_sigma,
        rim_inner, rim_outer, rim_strength;
    u32 rim_shape;
    R noise_amplitude, noise_wavelength;
    int noise_octaves;
    R noise_gain, noise_rim_amplitude, noise_rim_wavelength, noise_rim_inner, noise_rim_outer,
        wall_streaks, wall_streak_width, ejecta, ejecta_lineation, min_freshness;
};

struct Rec {
    R x, y, diameter, resolution, fresh_depth, fresh_rim, current_depth, current_rim, diffusion,
        ellipticity, heading;
    u32 first_harmonic, harmonic_count, harmonic_offset, table_count, table_offset, profile_offset,
        has_datum;
    R datum_a, datum_b, datum_c;
    u32 noise_key, texture_key, mark_count, mark_offset;
    R shift_x, shift_y, upslope_x, upslope_y;
    u32 terrace_count, terrace_offset, peak_count, peak_offset, replace;
    R x_low, y_low;
};

struct Frm {
    R x, y, z, east_x, east_y, east_z, north_x, north_y, north_z, sphere_radius, x_low, y_low,
        z_low;
};

struct Pt {
    R x, x_low, y, y_low, z, z_low;
};

inline Prof Widen(const tg_crater_profile& p) {
    return {p.simple, p.rim, p.complex_, p.diffused, p.transition_diameter, p.floor_coefficient,
            p.floor_exponent, p.peak_height_coefficient, p.peak_height_exponent,
            p.peak_width_coefficient, p.peak_width_exponent, p.peak_from_volume,
            p.peak_fit_max_diameter, p.ejecta_exponent, p.complex_reach_radii, p.table_extent,
            p.table_cells, p.ages_offset, p.age_count};
}

inline Shp Widen(const tg_crater_shape& k) {
    return {Widen(k.profile), k.reach_radii, k.profile_reach_radii, k.blend_start, k.blend_end,
            k.blend_shape, k.interior_replacement, k.replace_start, k.replace_end, k.mantling,
            k.ejecta_reach_radii, k.rim_sigma, k.rim_inner, k.rim_outer, k.rim_strength,
            k.rim_shape, k.noise_amplitude, k.noise_wavelength, k.noise_octaves, k.noise_gain,
            k.noise_rim_amplitude, k.noise_rim_wavelength, k.noise_rim_inner, k.noise_rim_outer,
            k.wall_streaks, k.wall_streak_width, k.ejecta, k.ejecta_lineation, k.min_freshness};
}

inline Rec Widen(const tg_crater_record& c) {
    return {c.x, c.y, c.diameter, c.resolution, c.fresh_depth, c.fresh_rim, c.current_depth,
            c.current_rim, c.diffusion, c.ellipticity, c.heading, c.first_harmonic,
            c.harmonic_count, c.harmonic_offset, c.table_count, c.table_offset, c.profile_offset,
            c.has_datum, c.datum_a, c.datum_b, c.datum_c, c.noise_key, c.texture_key,
            c.mark_count, c.mark_offset, c.shift_x, c.shift_y, c.upslope_x, c.upslope_y,
            c.terrace_count, c.terrace_offset, c.peak_count, c.peak_offset, c.replace, c.x_low,
            c.y_low};
}

inline Frm Widen(const tg_crater_frame& f) {
    return {f.x, f.y, f.z, f.east_x, f.east_y, f.east_z, f.north_x, f.north_y, f.north_z,
            f.sphere_radius, f.x_low, f.y_low, f.z_low};
}

inline R UnitFloat(u32 h) {
    return R(orc::UnitFloat(h));
}
using orc::Pcg;

// Fractal2 with KernelFractalParams (craters.h:657-669): Perlin, lacunarity 2, offset 1,
// amplitude 1; its FP32 parameters as the reference holds them.
inline R KernelFractal2(R x, R y, u32 seed, R wavelength, int octaves, R gain, u32 type) {
    const orc::Params p{seed, float(wavelength), 1.0f, octaves, 2.0f, float(gain),
                        TG_BASIS_PERLIN, type, 1.0f};
    return orc::Fractal2<R>(x, y, p);
}

// PerlinGradient3 and PerlinNoise3 (perlin.h:36-41,107-122), lattice.h:96-99 keys.
inline R Grad3(u32 hash, R x, R y, R z) {
    const u32 h = hash & 15u;
    const R u = h < 8u ? x : y;
    const R v = h < 4u ? y : (h == 12u || h == 14u ? x : z);
    return ((h & 1u) == 0u ? u : -u) + ((h & 2u) == 0u ? v : -v);
}

inline R Perlin3(R x, R y, R z, u32 seed) {
    const R cx = std::floor(x), cy = std::floor(y), cz = std::floor(z);
    const auto ix = std::int32_t(std::int64_t(cx));
    const auto iy = std::int32_t(std::int64_t(cy));
    const auto iz = std::int32_t(std::int64_t(cz));
    const R fx = x - cx, fy = y - cy, fz = z - cz;
    const u32 key = Pcg(seed);
    R c[8];
    for (u32 corner = 0; corner < 8u; ++corner) {
        const u32 ox = corner & 1u, oy = (corner >> 1u) & 1u, oz = corner >> 2u;
        const u32 h = orc::Pcg3({u32(ix) + ox, u32(iy) + oy, u32(iz) + oz + key}).x;
        c[corner] = Grad3(h, fx - R(ox), fy - R(oy), fz - R(oz));
    }
    const R u = orc::Fade(fx), v = orc::Fade(fy), w = orc::Fade(fz);
    const R nearV = orc::Mix(v, orc::Mix(u, c[0], c[1]), orc::Mix(u, c[2], c[3]));
    const R farV = orc::Mix(v, orc::Mix(u, c[4], c[5]), orc::Mix(u, c[6], c[7]));
    return StdClamp(orc::Mix(w, nearV, farV), R(-1), R(1));
}

// Fractal3Float (fractal.h:243-277) for unit-wavelength Perlin fBm of gain 0.5; the octave turns
// kept in FP32 as the reference keeps them.
inline R Fractal3Float(R x, R y, R z, u32 seed, int octaves) {
    const float kTurn[3][3] = {
        {-0.6132711010727254f, -0.293401570821183f, 0.7333580809050306f},
        {0.7897926788435601f, -0.24097777005594262f, 0.5640542870894418f},
        {0.011228581128535142f, 0.9251190369776896f, 0.37951111497202883f}};
    float turn[3][3] = {{1.0f, 0.0f, 0.0f}, {0.0f, 1.0f, 0.0f}, {0.0f, 0.0f, 1.0f}};
    float frequency = 1.0f;
    float weight = 1.0f;
    R sum = 0, total = 0;
    for (int o = 0; o < octaves; ++o) {
        const orc::U3 h = orc::Pcg3({seed, u32(o), 0u});
        const orc::U3 g = orc::Pcg3({seed, u32(o), 1u});
        const R rx = R(turn[0][0]) * x + R(turn[0][1]) * y + R(turn[0][2]) * z;
        const R ry = R(turn[1][0]) * x + R(turn[1][1]) * y + R(turn[1][2]) * z;
        const R rz = R(turn[2][0]) * x + R(turn[2][1]) * y + R(turn[2][2]) * z;
        const R n = Perlin3(rx * R(frequency) + UnitFloat(h.y), ry * R(frequency) + UnitFloat(h.z),
                            rz * R(frequency) + UnitFloat(g.x), h.x);
        sum += R(weight) * n;
        total += R(weight);
        weight *= 0.5f;
        frequency *= 2.0f;
        float next[3][3];
        for (int r = 0; r < 3; ++r) {
            for (int col = 0; col < 3; ++col) {
                next[r][col] = kTurn[r][0] * turn[0][col] + kTurn[r][1] * turn[1][col] +
                               kTurn[r][2] * turn[2][col];
            }
        }
        std::memcpy(turn, next, sizeof turn);
    }
    return R(1) * sum / total;
}

constexpr float kPi = 3.14159265358979f;
constexpr float kTwoPi = 6.28318530717959f;

struct Offset {
    R dx, dy;
};

// A thin view of the shape block so device code reads it by name.
struct CraterShape {
    const Shp& k;
    explicit CraterShape(const Shp& block) : k(block) {}
};

// SplitDifference (craters.h:152-154).
inline R SplitDifference(R a, R aLow, R b, R bLow) {
    return (a - b) + (aLow - bLow);
}

// KernelArcsineRatio (craters.h:156-166): asin(s)/s by its series below 0.25.
inline R ArcsineRatio(R s) {
    if (s < 0.25f) {
        const R s2 = s * s;
        return 1.0f + s2 * (1.0f / 6.0f +
                            s2 * (3.0f / 40.0f + s2 * (5.0f / 112.0f + s2 * (35.0f / 1152.0f))));
    }
    return std::atan2(s, std::sqrt(1.0f - s * s)) / s;
}

// KernelFrameOffset (craters.h:168-186).
inline Offset FrameOffset(const Frm& f, R dx, R dy,
                                              R dz) {
    Offset o;
    o.dx = dx * f.east_x + dy * f.east_y + dz * f.east_z;
    o.dy = dx * f.north_x + dy * f.north_y + dz * f.north_z;
    if (f.sphere_radius > 0.0f) {
        const R chord = std::sqrt(dx * dx + dy * dy + dz * dz);
        const R arc = chord * ArcsineRatio(StdMin(0.5f * chord / f.sphere_radius, 1.0f));
        const R across = std::sqrt(o.dx * o.dx + o.dy * o.dy);
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
inline Offset BodySplitOffset(const Frm& f,
                                                  const Pt& p) {
    return FrameOffset(f, SplitDifference(p.x, p.x_low, f.x, f.x_low),
                       SplitDifference(p.y, p.y_low, f.y, f.y_low),
                       SplitDifference(p.z, p.z_low, f.z, f.z_low));
}

// KernelFalloffShape / KernelFalloff (craters.h:241-270).
inline R FalloffShape(R t, uint32_t shape) {
    if (shape == 0u) {
        return 1.0f - t;
    }
    if (shape == 1u) {
        return 1.0f - t * t * (3.0f - 2.0f * t);
    }
    if (shape == 2u) {
        return 1.0f - t * t * t * (t * (t * 6.0f - 15.0f) + 10.0f);
    }
    const R tail = std::exp(-4.5f);
    const R z = 3.0f * t;
    return (std::exp(-0.5f * z * z) - tail) / (1.0f - tail);
}

inline R Falloff(R x, R start, R end, uint32_t shape) {
    if (x <= start) {
        return 1.0f;
    }
    if (x >= end) {
        return 0.0f;
    }
    return FalloffShape((x - start) / (end - start), shape);
}

// KernelSmoothstep (craters.h:272-276).
inline R Smoothstep(R t) {
    const R u = StdClamp(t, 0.0f, 1.0f);
    return u * u * (3.0f - 2.0f * u);
}

// KernelAngle (craters.h:278-285).
inline R Angle(R dy, R dx) {
    if (dx == 0.0f && dy == 0.0f) {
        return 0.0f;
    }
    return std::atan2(dy, dx);
}

struct PeriodicCell {
    uint32_t index;
    R t;
};

// KernelPeriodicAt (craters.h:297-304).
inline PeriodicCell PeriodicAt(R turns, uint32_t size) {
    const R position = (turns - std::floor(turns)) * R(size);
    PeriodicCell cell;
    const uint32_t whole = uint32_t(position);
    cell.index = whole < size - 1u ? whole : size - 1u;
    cell.t = position - R(cell.index);
    return cell;
}

// KernelEjectaScale / KernelEjectaShare (craters.h:306-324).
inline R EjectaScale(R radius) {
    return 0.14f * std::pow(radius, 0.74f);
}

inline R EjectaShare(R radii, R reachRadii) {
    if (radii <= 1.0f) {
        return 1.0f;
    }
    if (radii >= reachRadii) {
        return 0.0f;
    }
    const R atReach = 1.0f / (reachRadii * reachRadii * reachRadii);
    return (1.0f / (radii * radii * radii) - atReach) / (1.0f - atReach);
}

// FassettThomsonExterior / FassettThomsonShape (craters.h:326-362).
inline R FassettThomsonExterior(R x, R start) {
    const R polynomial =
        0.188253307f + x * (-0.187050452f + x * (0.01844746f + x * 0.01505647f));
    const R u = (x - start) / (1.5f - start);
    return polynomial * (1.0f - u * u);
}

inline R FassettThomsonShape(R x, uint32_t rim, R freshRim,
                                                     R currentDepth) {
    if (x > 1.5f) {
        return 0.0f;
    }
    const R interior =
        -0.228809953f + x * (0.227533882f + x * (0.083116795f - x * 0.039499407f));
    R h;
    if (rim == 2u) {
        if (x <= 0.98f) {
            h = interior;
        } else if (x <= 1.02f) {
            h = freshRim;
        } else {
            h = FassettThomsonExterior(x, 1.02f);
        }
    } else {
        const R edge = rim == 0u ? 0.98f : 1.0f;
        if (x <= 0.2f) {
            h = -0.181f;
        } else if (x <= edge) {
            h = interior;
        } else {
            h = FassettThomsonExterior(x, edge);
        }
    }
    return StdMax(h, freshRim - currentDepth);
}

struct ProfilePoint {
    R height, slope;
};

// MartinParkesDunstanExterior (craters.h:374-386).
inline ProfilePoint MpdExterior(R x, R rim, R ejectaScale) {
    const R offset = rim - ejectaScale;
    const R gradient = -ejectaScale / 8.0f - offset;
    const R r = 1.0f / (x + 1.0f);
    const R law = ejectaScale * r * r * r + offset + gradient * x;
    const R lawSlope = -3.0f * ejectaScale * r * r * r * r + gradient;
    const R taper = 1.0f - x * x;
    ProfilePoint p;
    p.height = law * taper;
    p.slope = lawSlope * taper - 2.0f * x * law;
    return p;
}

// The record fields a profile reads; a diffused crater's fresh state swaps the current ones.
struct ProfileState {
    R diameter, freshDepth, freshRim, currentDepth, currentRim;
};

// MartinParkesDunstanHeight (craters.h:388-416).
inline R MpdHeight(R distance, R diameter, const ProfileState& c) {
    const R depth = c.currentDepth * diameter;
    const R rim = c.currentRim * diameter;
    const R freshRim = c.freshRim * diameter;
    const R radius = 0.5f * diameter;
    const R x = distance / radius - 1.0f;
    if (x >= 1.0f) {
        return 0.0f;
    }
    const R k = freshRim - rim + depth;
    const R alpha = (rim - freshRim) / k;
    if (x <= alpha) {
        return (k * x + 2.0f * k) * x + freshRim;
    }
    if (x <= 0.0f) {
        return k * (alpha + 1.0f) / alpha * x * x + rim;
    }
    const R ejectaScale =
        freshRim > 0.0f ? EjectaScale(radius) * rim / freshRim : EjectaScale(radius);
    const R beta = StdMin(1.0f, 1.5f * (freshRim - rim) / k);
    if (x <= beta) {
        const ProfilePoint join = MpdExterior(beta, rim, ejectaScale);
        const R drop = rim - join.height;
        const R slope = drop > 0.0f ? StdMax(join.slope, -3.0f * drop / beta) : join.slope;
        const R t = x / beta;
        return rim + (join.height - rim) * t * t * (3.0f - 2.0f * t) +
               beta * slope * t * t * (t - 1.0f);
    }
    return MpdExterior(x, rim, ejectaScale).height;
}

// SimpleProfileHeight (craters.h:418-427).
inline R SimpleProfileHeight(const Prof& p, R distance,
                                                     R diameter, const ProfileState& c) {
    if (p.simple == TG_PROFILE_MPD) {
        return MpdHeight(distance, diameter, c);
    }
    return FassettThomsonShape(distance / (0.5f * diameter), p.rim, c.freshRim, c.currentDepth) *
           diameter;
}

// KilometerLaw (craters.h:429-433).
inline R KilometerLaw(R coefficient, R exponent, R diameter,
                                              bool volume) {
    const R kilometers = coefficient * std::pow(diameter / 1000.0f, exponent);
    return volume ? kilometers * 1e9f : kilometers * 1000.0f;
}

struct PeakSize {
    R height, base;
};

// KernelComplexPeak (craters.h:441-456).
inline PeakSize ComplexPeak(const Prof& p, R diameter) {
    const R fitted =
        p.peak_fit_max_diameter > 0.0f ? StdMin(diameter, p.peak_fit_max_diameter) : diameter;
    PeakSize size;
    size.height = KilometerLaw(p.peak_height_coefficient, p.peak_height_exponent, fitted, false);
    if (p.peak_from_volume == 1u) {
        const R volume =
            KilometerLaw(p.peak_width_coefficient, p.peak_width_exponent, fitted, true);
        size.base = std::sqrt(3.0f * volume / (kPi * size.height));
    } else {
        size.base =
            0.5f * KilometerLaw(p.peak_width_coefficient, p.peak_width_exponent, fitted, false);
    }
    size.base *= diameter / fitted;
    return size;
}

// KernelPeakDome (craters.h:458-465).
inline R PeakDome(PeakSize size, R distance) {
    if (!(distance < size.base)) {
        return 0.0f;
    }
    const R u = 1.0f - (distance / size.base) * (distance / size.base);
    return size.height * u * u;
}

// KernelKeptDepth (craters.h:467-470).
inline R KeptDepth(R freshDepth, R currentDepth) {
    return freshDepth > 0.0f ? currentDepth / freshDepth : 1.0f;
}

// ComplexProfileHeight (craters.h:472-494).
inline R ComplexProfileHeight(const Prof& p, R distance,
                                                      R diameter, const ProfileState& c) {
    const R radius = 0.5f * diameter;
    const R x = distance / radius;
    const R depth = c.currentDepth * diameter;
    const R rim = c.currentRim * diameter;
    if (x >= 1.0f) {
        if (x >= p.complex_reach_radii) {
            return 0.0f;
        }
        const R edge = std::pow(p.complex_reach_radii, -p.ejecta_exponent);
        return rim * (std::pow(x, -p.ejecta_exponent) - edge) / (1.0f - edge);
    }
    const R floorShare =
        StdMin(0.9f, KilometerLaw(p.floor_coefficient, p.floor_exponent, diameter, false) / diameter);
    R height = rim - depth;
    if (x > floorShare) {
        const R t = (x - floorShare) / (1.0f - floorShare);
        height += depth * t * t;
    }
    return height +
           KeptDepth(c.freshDepth, c.currentDepth) * PeakDome(ComplexPeak(p, diameter), distance);
}

// FreshProfileHeight (craters.h:496-503).
inline R FreshProfileHeight(const Prof& p, R distance,
                                                    R diameter, const ProfileState& c) {
    if (p.complex_ == 1u && diameter >= p.transition_diameter) {
        return ComplexProfileHeight(p, distance, diameter, c);
    }
    return SimpleProfileHeight(p, distance, diameter, c);
}

struct TableCell {
    uint32_t index;
    R t;
};

// KernelTableAt (craters.h:511-518).
inline TableCell TableAt(const Prof& p, R x) {
    const R position = x / p.table_extent * R(p.table_cells);
    TableCell cell;
    const uint32_t whole = uint32_t(position);
    cell.index = whole < p.table_cells - 1u ? whole : p.table_cells - 1u;
    cell.t = position - R(cell.index);
    return cell;
}

// KernelLerp (craters.h:520-523).
inline R DataLerp(const float* data, uint32_t at, R t) {
    return (1.0f - t) * data[at] + t * data[at + 1u];
}

struct AgeRow {
    uint32_t row;
    R weight;
};

// KernelAgeRowAt (craters.h:531-559): binary search of the tabulated diffusion ages.
inline AgeRow AgeRowAt(const Prof& p, const float* data,
                                           R tau) {
    AgeRow age;
    age.row = 0u;
    age.weight = 0.0f;
    const uint32_t last = p.age_count - 1u;
    const R first = data[p.ages_offset + 1u];
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
        const R below = data[p.ages_offset + low];
        const R above = data[p.ages_offset + high];
        age.weight = std::log(tau / below) / std::log(above / below);
    }
    return age;
}

// KernelDiffusedValue (craters.h:565-577).
inline R DiffusedValue(const Prof& p, const float* data,
                                               uint32_t rows, R x, R tau) {
    const TableCell cell = TableAt(p, x);
    const AgeRow age = AgeRowAt(p, data, tau);
    const uint32_t width = p.table_cells + 1u;
    const uint32_t at = rows + age.row * width + cell.index;
    R value = DataLerp(data, at, cell.t);
    if (age.weight > 0.0f) {
        value += age.weight * (DataLerp(data, at + width, cell.t) - value);
    }
    return value;
}

// Per-crater values every profile evaluation of one (cell, crater) pair shares.
struct CraterView {
    const Rec& c;
    ProfileState current;
    ProfileState fresh;
};

// KernelProfileHeight (craters.h:608-618) with DiffusedProfileHeight (craters.h:580-594) and
// TabulatedProfileHeight (craters.h:598-606).
inline R ProfileHeight(const Shp& k, const float* data,
                                               R distance, R diameter,
                                               const CraterView& v) {
    const Prof& p = k.profile;
    if (p.simple == TG_PROFILE_TABULATED) {
        const R x = distance / diameter;
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
        const R x = distance / diameter;
        if (x >= p.table_extent) {
            return 0.0f;
        }
        return diameter * DiffusedValue(p, data, v.c.profile_offset, x, v.c.diffusion);
    }
    return FreshProfileHeight(p, distance, diameter, v.current);
}

// KernelRoundedHeight (craters.h:624-653): the rim rounded by a Gaussian on a grid anchored at it.
inline R RoundedHeight(const Shp& k, const float* data,
                                               R distance, R diameter,
                                               const CraterView& v) {
    const R height = ProfileHeight(k, data, distance, diameter, v);
    if (k.rim_sigma == 0.0f) {
        return height;
    }
    const R radius = 0.5f * diameter;
    const R radii = distance / radius;
    const R weight =
        radii <= 1.0f
            ? k.rim_strength * Falloff(2.0f - radii, 1.0f, 2.0f - k.rim_inner, k.rim_shape)
            : k.rim_strength * Falloff(radii, 1.0f, k.rim_outer, k.rim_shape);
    if (weight == 0.0f) {
        return height;
    }
    const R sigma = k.rim_sigma * radius;
    const R step = 0.5f * sigma;
    const R position = (distance - radius) / step;
    const R first = std::floor(position) - 7.0f;
    R smoothed = 0.0f;
    R total = 0.0f;
    for (int i = 0; i < 16; ++i) {
        const R node = first + R(i);
        const R z = 0.5f * (node - position);
        const R w = std::exp(-0.5f * z * z) * Falloff(std::abs(z), 3.0f, 4.0f, 2u);
        smoothed += w * ProfileHeight(k, data, std::abs(radius + step * node), diameter, v);
        total += w;
    }
    return height + weight * (smoothed / total - height);
}

// KernelNoiseHeight (craters.h:672-694).
inline R NoiseHeight(const Shp& k, const Rec& c,
                                             R dx, R dy, R radii) {
    if (k.noise_amplitude == 0.0f && k.noise_rim_amplitude == 0.0f) {
        return 0.0f;
    }
    const R radius = 0.5f * c.diameter;
    const R footprint =
        radii <= 1.0f ? 1.0f : Falloff(radii, 1.0f, k.profile_reach_radii, 2u);
    const R rim = radii <= 1.0f ? Falloff(2.0f - radii, 1.0f, 2.0f - k.noise_rim_inner, 2u)
                                    : Falloff(radii, 1.0f, k.noise_rim_outer, 2u);
    R height = 0.0f;
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

// KernelRadialStreaks (craters.h:697-703).
inline R RadialStreaks(R radii, R angle, R across, R along,
                                               int octaves, uint32_t seed) {
    const R ring = 1.0f / across;
    return Fractal3Float(ring * std::cos(angle), ring * std::sin(angle), radii / along, seed, octaves);
}

// KernelTextureHeight (craters.h:706-733).
inline R TextureHeight(const Shp& k, const Rec& c,
                                               R dx, R dy, R radii) {
    if (k.wall_streaks == 0.0f && k.ejecta == 0.0f) {
        return 0.0f;
    }
    const R freshness = c.fresh_depth > 0.0f ? c.current_depth / c.fresh_depth : 1.0f;
    const R fade = Smoothstep((freshness - k.min_freshness) / (1.0f - k.min_freshness));
    if (!(fade > 0.0f)) {
        return 0.0f;
    }
    const R angle = Angle(dy, dx);
    R height = 0.0f;
    if (k.wall_streaks > 0.0f && radii > 0.25f && radii < 1.0f) {
        const R band =
            Smoothstep((radii - 0.25f) / 0.35f) * Smoothstep((1.0f - radii) / 0.1f);
        height += k.wall_streaks * c.diameter * band *
                  RadialStreaks(radii, angle, k.wall_streak_width, 0.5f, 2, c.texture_key);
    }
    if (k.ejecta > 0.0f && radii > 1.0f && radii < k.ejecta_reach_radii) {
        const R radius = 0.5f * c.diameter;
        const R thickness = EjectaScale(radius) * EjectaShare(radii, k.ejecta_reach_radii);
        const R rise = Smoothstep((radii - 1.0f) / 0.1f);
        height += k.ejecta * thickness * rise *
                  RadialStreaks(radii, angle, k.ejecta_lineation, 2.0f * k.ejecta_lineation, 3,
                                c.texture_key + 1u);
    }
    return fade * height;
}

// KernelEllipseScale (craters.h:737-743).
inline R EllipseScale(R theta, R ellipticity, R heading) {
    const R along = std::sqrt(ellipticity);
    const R across = 1.0f / along;
    const R u = std::cos(theta - heading) / along;
    const R v = std::sin(theta - heading) / across;
    return 1.0f / std::sqrt(u * u + v * v);
}

// KernelCatmullRom (craters.h:746-749).
inline R CatmullRom(R p0, R p1, R p2, R p3, R t) {
    return 0.5f * (2.0f * p1 + (p2 - p0) * t + (2.0f * p0 - 5.0f * p1 + 4.0f * p2 - p3) * t * t +
                   (3.0f * (p1 - p2) + p3 - p0) * t * t * t);
}

// KernelMarksWindow (craters.h:753-759).
inline R MarksWindow(R radii, R inner, R outer, R fade) {
    if (!(radii > inner - fade) || !(radii < outer + fade)) {
        return 0.0f;
    }
    return Smoothstep((radii - inner + fade) / fade) * Smoothstep((outer + fade - radii) / fade);
}

constexpr float kTableFineFade = 6.5f;

// KernelOutlineScaleAt (craters.h:772-808).
inline R OutlineScaleAt(const float* data, const Rec& c,
                                                R dx, R dy, R share) {
    const R theta = Angle(dy, dx);
    const R scale =
        c.ellipticity != 1.0f ? EllipseScale(theta, c.ellipticity, c.heading) : 1.0f;
    R wobble = 1.0f;
    if (c.table_count > 0u && share > 0.0f) {
        const PeriodicCell cell = PeriodicAt(theta / kTwoPi, c.table_count);
        const R rim = (1.0f - cell.t) * data[c.table_offset + cell.index] +
                          cell.t * data[c.table_offset + (cell.index + 1u) % c.table_count];
        if (share >= 1.0f) {
            wobble += share * rim;
        } else {
            R broad = 0.0f;
            for (uint32_t n = 0u; n < c.harmonic_count; ++n) {
                const R order = R(c.first_harmonic + n);
                const R wave = data[c.harmonic_offset + 2u * n] *
                                   std::cos(order * theta + data[c.harmonic_offset + 2u * n + 1u]);
                broad += wave;
                wobble += std::pow(share, 0.5f * order) * wave;
            }
            wobble += std::pow(share, kTableFineFade) * (rim - broad);
        }
    } else if (share > 0.0f) {
        for (uint32_t n = 0u; n < c.harmonic_count; ++n) {
            const R order = R(c.first_harmonic + n);
            const R kept = share < 1.0f ? std::pow(share, 0.5f * order) : 1.0f;
            wobble += kept * data[c.harmonic_offset + 2u * n] *
                      std::cos(order * theta + data[c.harmonic_offset + 2u * n + 1u]);
        }
    }
    return scale * wobble;
}

// KernelWornOutlineShare (craters.h:826-828).
inline R WornOutlineShare(const Rec& c) {
    return c.fresh_rim > 0.0f ? StdClamp(c.current_rim / c.fresh_rim, 0.0f, 1.0f) : 1.0f;
}

// KernelMarkShift (craters.h:832-848).
inline R MarkShift(const float* data, const Rec& c, R dx,
                                           R dy, R radii) {
    const R window = MarksWindow(radii, data[c.mark_offset], data[c.mark_offset + 1u],
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
inline R TerraceBlockValue(uint32_t key, int cell, int count, R gaps) {
    const int wrapped = ((cell % count) + count) % count;
    const R hash = UnitFloat(Pcg(key + uint32_t(wrapped) * 0x9E3779B9u));
    return gaps < 0.0f ? hash : (hash >= gaps ? 1.0f : 0.0f);
}

// KernelTerraceBlocks (craters.h:884-902).
inline R TerraceBlocks(const float* data, uint32_t s, R theta,
                                               uint32_t salt, R gaps) {
    const R turns = data[s + 2u] * theta / kTwoPi + data[s + 3u];
    const int count = int(data[s + 2u]);
    const uint32_t key = uint32_t(data[s + 7u]) ^ (salt * 0x85EBCA6Bu);
    const R cellStart = std::floor(turns);
    const int cell = int(cellStart);
    const R u = turns - cellStart;
    const R value = TerraceBlockValue(key, cell, count, gaps);
    if (u < kTerraceBlockBlend) {
        const R b = Smoothstep(0.5f + 0.5f * u / kTerraceBlockBlend);
        return TerraceBlockValue(key, cell - 1, count, gaps) * (1.0f - b) + value * b;
    }
    if (u > 1.0f - kTerraceBlockBlend) {
        const R b = Smoothstep(0.5f + 0.5f * (1.0f - u) / kTerraceBlockBlend);
        return TerraceBlockValue(key, cell + 1, count, gaps) * (1.0f - b) + value * b;
    }
    return value;
}

// KernelHarmonics (craters.h:906-915).
inline R Harmonics(const float* data, uint32_t offset, uint32_t first,
                                           uint32_t count, R theta, R turn) {
    R sum = 0.0f;
    for (uint32_t n = 0u; n < count; ++n) {
        const R order = R(first + n);
        sum += data[offset + 2u * n] * std::cos(order * (theta + turn) + data[offset + 2u * n + 1u]);
    }
    return sum;
}

// KernelTerraceArc (craters.h:920-927).
inline R TerraceArc(const float* data, uint32_t o, R theta,
                                            R step) {
    const R cut = data[o + 4u];
    if (!(cut > -1.5f)) {
        return 1.0f;
    }
    const R arcs = Harmonics(data, o + 6u, 2u, 5u, theta, 2.39996323f * step);
    return Smoothstep((arcs - cut) / data[o + 5u]);
}

// KernelTerraceBase (craters.h:931-939).
inline R TerraceBase(const float* data, const Rec& c,
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
inline R TerraceScarp(const float* data, const Rec& c,
                                              uint32_t k, R theta) {
    if (k == 0u || k >= c.terrace_count) {
        return TerraceBase(data, c, k);
    }
    const uint32_t s = c.terrace_offset + kTerraceHeader + kTerraceScarpValues * (k - 1u);
    const R turns = data[s + 2u] * theta / kTwoPi + data[s + 3u];
    const R f = 2.0f * (turns - std::floor(turns)) - 1.0f;
    const R scallop = data[s + 1u] * (1.0f - f * f - 2.0f / 3.0f);
    const R wobble = data[s + 4u] * std::cos(3.0f * theta + data[s + 5u]);
    const R block = 2.0f * TerraceBlocks(data, s, theta, 0u, -1.0f) - 1.0f;
    return data[s] + scallop + wobble + data[s + 6u] * block;
}

// KernelTerraceRimScale (craters.h:964-970).
inline R TerraceRimScale(const float* data, const Rec& c,
                                                 R theta) {
    const uint32_t o = c.terrace_offset;
    const R turns = data[o + 17u] * theta / kTwoPi + data[o + 18u];
    const R f = 2.0f * (turns - std::floor(turns)) - 1.0f;
    return 1.0f + data[o + 16u] * (1.0f - f * f - 2.0f / 3.0f);
}

// KernelTerraceEase (craters.h:975-981).
inline R TerraceEase(R v, R rounding) {
    const R steepest = 1.0f / (1.0f - 0.5f * rounding);
    if (v < rounding) {
        return steepest * v * v / (2.0f * rounding);
    }
    return steepest * (v - 0.5f * rounding);
}

// KernelTerracedRadii (craters.h:986-1030).
inline R TerracedRadii(const float* data, const Rec& c,
                                               R radii, R theta) {
    const uint32_t o = c.terrace_offset;
    const R inner = data[o];
    if (!(radii > inner) || !(radii < 1.0f)) {
        return radii;
    }
    const R t = (radii - inner) / (1.0f - inner);
    uint32_t step = 0u;
    R low = 0.0f;
    R high = TerraceScarp(data, c, 1u, theta);
    while (step + 1u < c.terrace_count && t >= high) {
        ++step;
        low = high;
        high = TerraceScarp(data, c, step + 1u, theta);
    }
    const R within = StdClamp((t - low) / (high - low), 0.0f, 1.0f);
    R strength = data[o + 2u] * TerraceArc(data, o, theta, R(step));
    if (c.terrace_count > 1u) {
        const bool top = step + 1u >= c.terrace_count;
        const uint32_t k = top ? step : step + 1u;
        const uint32_t s = o + kTerraceHeader + kTerraceScarpValues * (k - 1u);
        strength *= TerraceBlocks(data, s, theta, top ? 2u : 1u, data[o + 19u]);
    }
    const R rounding = data[o + 3u];
    const R slope = data[o + 1u];
    const uint32_t s = o + kTerraceHeader + kTerraceScarpValues * (c.terrace_count - 1u) +
                       kTerraceStepValues * step;
    const R level = data[s];
    const R scarp = data[s + 1u];
    strength *= data[s + 2u];
    const R lower = scarp * level;
    const R upper = scarp * (1.0f - level);
    R stair = level;
    if (within < lower) {
        stair = level * (1.0f - TerraceEase(1.0f - within / lower, rounding));
    } else if (within > 1.0f - upper) {
        stair = level + (1.0f - level) * TerraceEase((within - (1.0f - upper)) / upper, rounding);
    }
    const R rise = slope * within + (1.0f - slope) * stair;
    const R shaped = within + strength * (rise - within);
    const R from = TerraceBase(data, c, step);
    const R to = TerraceBase(data, c, step + 1u);
    return inner + (from + shaped * (to - from)) * (1.0f - inner);
}

// KernelProfileDistance (craters.h:1034-1054).
inline R ProfileDistance(const float* data, const Rec& c,
                                                 R dx, R dy) {
    R distance = std::sqrt(dx * dx + dy * dy);
    if (c.ellipticity != 1.0f || c.harmonic_count > 0u || c.table_count > 0u ||
        c.mark_count > 0u) {
        const R share = Smoothstep(distance / (0.5f * c.diameter)) * WornOutlineShare(c);
        distance /= OutlineScaleAt(data, c, dx, dy, share);
    }
    if (c.mark_count > 0u) {
        const R radius = 0.5f * c.diameter;
        distance += radius * MarkShift(data, c, dx, dy, distance / radius);
    }
    if (c.terrace_count > 0u) {
        const R radius = 0.5f * c.diameter;
        const R theta = Angle(dy, dx);
        const R scale = TerraceRimScale(data, c, theta);
        distance /= 1.0f + (scale - 1.0f) * Smoothstep(distance / radius);
        distance = radius * TerracedRadii(data, c, distance / radius, theta);
    }
    return distance;
}

// KernelSlopeShiftWeight (craters.h:1058-1061).
inline R SlopeShiftWeight(const Rec& c, R dx, R dy) {
    const R radii = std::sqrt(dx * dx + dy * dy) / (0.5f * c.diameter);
    return 1.0f - Smoothstep((radii - 0.3f) / 0.7f);
}

// KernelSlopeRaisedShare (craters.h:1065-1073).
inline R SlopeRaisedShare(const Rec& c, R dx, R dy) {
    const R distance = std::sqrt(dx * dx + dy * dy);
    if (!(distance > 0.0f)) {
        return 1.0f;
    }
    const R removed = std::sqrt(c.upslope_x * c.upslope_x + c.upslope_y * c.upslope_y);
    const R cosine =
        StdMax(0.0f, (c.upslope_x * dx + c.upslope_y * dy) / (removed * distance));
    return 1.0f - removed * cosine * cosine;
}

// KernelShiftedDistance (craters.h:1077-1081).
inline R ShiftedDistance(const float* data, const Rec& c,
                                                 R dx, R dy) {
    const R shift =
        c.shift_x != 0.0f || c.shift_y != 0.0f ? SlopeShiftWeight(c, dx, dy) : 0.0f;
    return ProfileDistance(data, c, dx - shift * c.shift_x, dy - shift * c.shift_y);
}

// KernelUpslopeHeight (craters.h:1085-1091).
inline R UpslopeHeight(const Rec& c, R height, R dx,
                                               R dy) {
    if (height > 0.0f && (c.upslope_x != 0.0f || c.upslope_y != 0.0f)) {
        return height * SlopeRaisedShare(c, dx, dy);
    }
    return height;
}

constexpr uint32_t kPeakHeader = 6u;
constexpr uint32_t kPeakMassifValues = 11u;
constexpr float kPeakReach = 2.6f;

// KernelPeakMassif (craters.h:1104-1123).
inline R PeakMassif(const float* data, uint32_t m, R px, R py) {
    const R ex = px - data[m];
    const R ey = py - data[m + 1u];
    const R radius = data[m + 2u];
    const R stretch = data[m + 3u];
    const R cosine = data[m + 4u];
    const R sine = data[m + 5u];
    const R qx = (ex * cosine + ey * sine) / (radius * stretch);
    const R qy = (ey * cosine - ex * sine) * stretch / radius;
    const R squared = qx * qx + qy * qy;
    if (!(squared < 1.0f)) {
        return 0.0f;
    }
    const R pull = 1.0f - squared;
    const R sx = qx - data[m + 8u] * pull;
    const R sy = qy - data[m + 9u] * pull;
    const R r = StdMin(std::sqrt(sx * sx + sy * sy), 1.0f);
    const R flank = 1.0f - std::pow(r, data[m + 7u]);
    return data[m + 6u] * std::pow(flank, data[m + 10u]);
}

// KernelPeakMassifs (craters.h:1128-1155).
inline R PeakMassifs(const float* data, const Rec& c,
                                             PeakSize size, R dx, R dy) {
    const uint32_t o = c.peak_offset;
    R px = dx / size.base;
    R py = dy / size.base;
    const R roughness = data[o];
    const R wavelength = data[o + 1u];
    const R warp = data[o + 2u];
    const uint32_t key = uint32_t(data[o + 3u]);
    if (warp > 0.0f) {
        const R wx = KernelFractal2(px, py, key + 1u, 2.0f * wavelength, 3, 0.5f,
                                        TG_FRACTAL_FBM);
        const R wy = KernelFractal2(px + 17.31f, py - 9.73f, key + 1u, 2.0f * wavelength, 3,
                                        0.5f, TG_FRACTAL_FBM);
        px += warp * wx;
        py += warp * wy;
    }
    R height = 0.0f;
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
inline R PeakRelief(const Shp& k, const float* data,
                                            const Rec& c, R distance, R dx,
                                            R dy) {
    if (c.peak_count == 0u || k.profile.complex_ != 1u || k.profile.diffused == 1u ||
        k.profile.simple == TG_PROFILE_TABULATED || c.diameter < k.profile.transition_diameter) {
        return 0.0f;
    }
    const PeakSize size = ComplexPeak(k.profile, c.diameter);
    PeakSize scaled = size;
    scaled.height *= data[c.peak_offset + 4u];
    scaled.base *= data[c.peak_offset + 5u];
    const R shift =
        c.shift_x != 0.0f || c.shift_y != 0.0f ? SlopeShiftWeight(c, dx, dy) : 0.0f;
    const R sx = dx - shift * c.shift_x;
    const R sy = dy - shift * c.shift_y;
    const R reach = kPeakReach * StdMax(size.base, scaled.base);
    if (!(sx * sx + sy * sy < reach * reach)) {
        return 0.0f;
    }
    return KeptDepth(c.fresh_depth, c.current_depth) *
           (PeakMassifs(data, c, scaled, sx, sy) - PeakDome(size, distance));
}

// ApplyCraterKernel (craters.h:1186-1212): `height` after crater `c` at (dx, dy) from its center.
inline R ApplyCrater(const CraterShape& shape, const float* data,
                                             const Rec& c, R height, R dx,
                                             R dy) {
    const Shp& k = shape.k;
    const R diameter = c.diameter;
    const R radius = 0.5f * diameter;
    const R distance = ShiftedDistance(data, c, dx, dy);
    if (distance >= k.reach_radii * radius) {
        return height;
    }
    const R radii = distance / radius;
    R replaced = 0.0f;
    if (c.has_datum == 1u) {
        R weight = c.replace == 1u
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
    const R bare = UpslopeHeight(c, RoundedHeight(k, data, distance, diameter, view), dx, dy) +
                       PeakRelief(k, data, c, distance, dx, dy);
    const R shaped =
        bare + NoiseHeight(k, c, dx, dy, radii) + TextureHeight(k, c, dx, dy, radii);
    return height + c.resolution * shaped + c.resolution * replaced;
}

// CraterRelief (craters.h:1225-1238): the relief crater `c` adds at (dx, dy) without its datum, as
// datums are fitted to it: resolution times the bare profile and the central peak, within the
// profile's reach; nothing beyond `bound`.
inline R CraterRelief(const Shp& k, const float* data, const Rec& c, R bound, R dx, R dy) {
    const R squared = dx * dx + dy * dy;
    if (!(squared < bound * bound)) {
        return 0.0f;
    }
    const R distance = ShiftedDistance(data, c, dx, dy);
    if (!(distance < k.profile_reach_radii * 0.5f * c.diameter)) {
        return 0.0f;
    }
    CraterView view{c,
                    {c.diameter, c.fresh_depth, c.fresh_rim, c.current_depth, c.current_rim},
                    {c.diameter, c.fresh_depth, c.fresh_rim, c.fresh_depth, c.fresh_rim}};
    return c.resolution * (UpslopeHeight(c, ProfileHeight(k, data, distance, c.diameter, view),
                                         dx, dy) +
                           PeakRelief(k, data, c, distance, dx, dy));
}

}  // namespace ORC_NS
