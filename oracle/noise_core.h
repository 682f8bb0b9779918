// SPDX-License-Identifier: BSD-3-Clause
//
// ORACLE — test infrastructure only (see noise_oracle.cpp). The noise core shared by the noise
// and crater oracles: PCG hashes, Perlin/simplex/value/Worley bases in 2D and 3D and the fractal
// octave loops (Fractal2 on a plane, Fractal3 on a body), templated on the precision of the
// per-cell arithmetic.
#pragma once

#include "tg_layouts.h"

#include <algorithm>
#include <cstring>
#include <cmath>
#include <cstdint>
#include <vector>

namespace orc {

using u32 = std::uint32_t;

// include/terraingen/noise/hash.h:33-37
inline u32 Pcg(u32 v) {
    const u32 state = v * 747796405u + 2891336453u;
    const u32 word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

struct U3 {
    u32 x, y, z;
};

// include/terraingen/noise/hash.h:40-57
inline U3 Pcg3(U3 v) {
    v.x = v.x * 1664525u + 1013904223u;
    v.y = v.y * 1664525u + 1013904223u;
    v.z = v.z * 1664525u + 1013904223u;
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

// include/terraingen/noise/hash.h:60-62
inline float UnitFloat(u32 h) {
    return float(h >> 8u) * 0x1p-24f;
}

template <class R> struct Lattice2 {
    std::int32_t ix, iy;
    R fx, fy;
};

// include/terraingen/noise/lattice.h:31-36 (floor, then the fraction in the same precision).
template <class R> Lattice2<R> Split2(R x, R y) {
    const R cx = std::floor(x);
    const R cy = std::floor(y);
    return {std::int32_t(std::int64_t(cx)), std::int32_t(std::int64_t(cy)), x - cx, y - cy};
}

// include/terraingen/noise/lattice.h:76-83
template <class R> inline R Fade(R t) {
    return t * t * t * (t * (t * R(6) - R(15)) + R(10));
}
template <class R> inline R Mix(R t, R a, R b) {
    return a + t * (b - a);
}

// include/terraingen/noise/perlin.h:17-26; the constant is the reference's float constant.
template <class R> inline R Grad2(u32 hash, R x, R y) {
    const R invSqrt2 = R(0.707106781186547524f);
    const u32 h = hash & 7u;
    const R axis = (h & 2u) != 0u ? y : x;
    const R sx = h == 5u ? -x : x;
    const R sy = h == 6u ? -y : y;
    const R sum = sx + sy;
    const R diag = (h == 7u ? -sum : sum) * invSqrt2;
    return h < 4u ? ((h & 1u) != 0u ? -axis : axis) : diag;
}

// include/terraingen/noise/perlin.h:31-33, lattice.h:89-92
inline u32 CornerHash(u32 ix, u32 iy, u32 seed) {
    return Pcg3({ix, iy, seed}).x;
}

// include/terraingen/noise/perlin.h:72-90
template <class R> inline R Perlin2(std::int32_t ix, std::int32_t iy, R fx, R fy, u32 seed) {
    const u32 x0 = u32(ix), y0 = u32(iy), x1 = x0 + 1u, y1 = y0 + 1u;
    const R n00 = Grad2<R>(CornerHash(x0, y0, seed), fx, fy);
    const R n10 = Grad2<R>(CornerHash(x1, y0, seed), fx - R(1), fy);
    const R n01 = Grad2<R>(CornerHash(x0, y1, seed), fx, fy - R(1));
    const R n11 = Grad2<R>(CornerHash(x1, y1, seed), fx - R(1), fy - R(1));
    const R u = Fade(fx), v = Fade(fy);
    return R(1.41421356237309505f) * Mix(v, Mix(u, n00, n10), Mix(u, n01, n11));
}

template <class R> inline R Perlin2At(R x, R y, u32 seed) {
    const Lattice2<R> p = Split2(x, y);
    return Perlin2(p.ix, p.iy, p.fx, p.fy, seed);
}

// include/terraingen/noise/simplex.h:18-24 (corner) and :34-59 (noise).
template <class R> inline R SimplexCorner(u32 ci, u32 cj, R dx, R dy, u32 seed) {
    const R t = R(0.5) - dx * dx - dy * dy;
    const R t2 = t * t;
    const R c = t2 * t2 * Grad2<R>(CornerHash(ci, cj, seed), dx, dy);
    return !(t <= R(0)) ? c : R(0);
}

template <class R> inline R Simplex2(R x, R y, u32 seed) {
    const R skewK = R(0.366025403784438647f);
    const R unskewK = R(0.211324865405187118f);
    const R skew = (x + y) * skewK;
    const R cx = std::floor(x + skew);
    const R cy = std::floor(y + skew);
    const R unskew = (cx + cy) * unskewK;
    const R x0 = x - (cx - unskew);
    const R y0 = y - (cy - unskew);
    const u32 stepX = x0 > y0 ? 1u : 0u;
    const u32 stepY = 1u - stepX;
    const R x1 = x0 - R(stepX) + unskewK;
    const R y1 = y0 - R(stepY) + unskewK;
    const R x2 = x0 - R(1) + R(2) * unskewK;
    const R y2 = y0 - R(1) + R(2) * unskewK;
    const u32 i = u32(std::int32_t(std::int64_t(cx)));
    const u32 j = u32(std::int32_t(std::int64_t(cy)));
    return R(99) * (SimplexCorner(i, j, x0, y0, seed) +
                    SimplexCorner(i + stepX, j + stepY, x1, y1, seed) +
                    SimplexCorner(i + 1u, j + 1u, x2, y2, seed));
}

// include/terraingen/noise/value.h:19-21 and :34-45
template <class R> inline R LatticeValue(U3 key) {
    return R(2) * R(UnitFloat(Pcg3(key).y)) - R(1);
}

template <class R> inline R Value2(R x, R y, u32 seed) {
    const Lattice2<R> p = Split2(x, y);
    const u32 x0 = u32(p.ix), y0 = u32(p.iy);
    const R v00 = LatticeValue<R>({x0, y0, seed});
    const R v10 = LatticeValue<R>({x0 + 1u, y0, seed});
    const R v01 = LatticeValue<R>({x0, y0 + 1u, seed});
    const R v11 = LatticeValue<R>({x0 + 1u, y0 + 1u, seed});
    const R u = Fade(p.fx), v = Fade(p.fy);
    return Mix(v, Mix(u, v00, v10), Mix(u, v01, v11));
}

template <class R> struct Worley {
    R f1, f2, value;
};

// include/terraingen/noise/worley.h:38-63: 5x5 neighbourhood, F1, F2 and the nearest cell's value.
template <class R> inline Worley<R> Worley2(R x, R y, u32 seed) {
    const Lattice2<R> p = Split2(x, y);
    R d1 = R(1e30f), d2 = R(1e30f);
    u32 nearest = 0u;
    for (std::int32_t dy = -2; dy <= 2; ++dy) {
        for (std::int32_t dx = -2; dx <= 2; ++dx) {
            const U3 h = Pcg3({u32(p.ix + dx), u32(p.iy + dy), seed});
            const R px = R(dx) + R(UnitFloat(h.y)) - p.fx;
            const R py = R(dy) + R(UnitFloat(h.z)) - p.fy;
            const R d = px * px + py * py;
            if (d < d1) {
                d2 = d1;
                d1 = d;
                nearest = h.x;
            } else if (d < d2) {
                d2 = d;
            }
        }
    }
    return {std::sqrt(d1), std::sqrt(d2), R(2) * R(UnitFloat(nearest)) - R(1)};
}

// include/terraingen/noise/fractal.h:89-111
template <class R> inline R Basis2(u32 basis, R x, R y, u32 seed) {
    const R invSqrt2 = R(0.707106781186547524f);
    const R invSqrt5 = R(0.447213595499957939f);
    switch (basis) {
    case TG_BASIS_SIMPLEX:
        return Simplex2(x, y, seed);
    case TG_BASIS_VALUE:
        return Value2(x, y, seed);
    case TG_BASIS_WORLEY_F1:
        return R(2) * std::min(Worley2(x, y, seed).f1 * invSqrt2, R(1)) - R(1);
    case TG_BASIS_WORLEY_F2:
        return R(2) * std::min(Worley2(x, y, seed).f2 * invSqrt5, R(1)) - R(1);
    case TG_BASIS_WORLEY_F2_MINUS_F1: {
        const Worley<R> w = Worley2(x, y, seed);
        return R(2) * std::min((w.f2 - w.f1) * invSqrt5, R(1)) - R(1);
    }
    case TG_BASIS_VORONOI:
        return Worley2(x, y, seed).value;
    default:
        return Perlin2At(x, y, seed);
    }
}

// The per-octave parameters of a fractal, derived in FP32 exactly as the reference's octave loop
// does (fractal.h:219-239): the seed and lattice offsets from Pcg3(OctaveKey(seed, octave, 0)),
// the frequency by repeated multiplication, the golden-angle turn by its FP32 recurrence
// (fractal.h:142-144).
struct Octave {
    u32 seed;
    float offsetX, offsetY, frequency, turnCos, turnSin;
};

struct Params {
    u32 seed;
    float wavelength, amplitude;
    int octaves;
    float lacunarity, gain;
    u32 basis, type;
    float offset;
};

inline std::vector<Octave> Octaves(const Params& p) {
    std::vector<Octave> out;
    const float kCos = -0.7373688780783197f;
    const float kSin = 0.6754902942615238f;
    float frequency = 1.0f / p.wavelength;
    float c = 1.0f, s = 0.0f;
    for (int o = 0; o < p.octaves; ++o) {
        const U3 h = Pcg3({p.seed, u32(o), 0u});
        out.push_back({h.x, UnitFloat(h.y), UnitFloat(h.z), frequency, c, s});
        frequency *= p.lacunarity;
        const float nc = c * kCos - s * kSin;
        s = s * kCos + c * kSin;
        c = nc;
    }
    return out;
}

template <class R> struct Sums {
    R weight = R(1), sum = R(0), total = R(0), feedback = R(1);
};

// include/terraingen/noise/fractal.h:164-197; the weights follow the gain in precision R.
template <class R> inline void AddOctave(Sums<R>& s, const Params& p, int octave, R n) {
    const R offset = R(p.offset);
    switch (p.type) {
    case TG_FRACTAL_BILLOW:
        s.sum += s.weight * (R(2) * std::abs(n) - R(1));
        break;
    case TG_FRACTAL_RIDGED: {
        R signal = offset - std::abs(n);
        signal *= signal * s.feedback;
        s.feedback = std::clamp(R(2) * signal, R(0), R(1));
        s.sum += s.weight * signal;
        break;
    }
    case TG_FRACTAL_HYBRID: {
        const R signal = (n + offset) * s.weight;
        if (octave == 0) {
            s.sum = signal;
            s.feedback = signal;
        } else {
            s.feedback = std::clamp(s.feedback, R(0), R(1));
            s.sum += s.feedback * signal;
            s.feedback *= signal;
        }
        break;
    }
    default:
        s.sum += s.weight * n;
        break;
    }
    s.total += s.weight;
    s.weight *= R(p.gain);
}

// include/terraingen/noise/fractal.h:201-211
template <class R> inline R Finish(const Sums<R>& s, const Params& p) {
    const R amp = R(p.amplitude), offset = R(p.offset);
    switch (p.type) {
    case TG_FRACTAL_RIDGED: {
        const R ridge = std::max(offset, R(1) - offset);
        return amp * (R(2) * s.sum / (ridge * ridge * s.total) - R(1));
    }
    case TG_FRACTAL_HYBRID:
        return amp * s.sum / ((R(1) + offset) * s.total);
    default:
        return amp * s.sum / s.total;
    }
}

// include/terraingen/noise/fractal.h:219-239
template <class R> inline R Fractal2(R x, R y, const Params& p) {
    Sums<R> sums;
    const std::vector<Octave> octaves = Octaves(p);
    for (int o = 0; o < p.octaves; ++o) {
        const Octave& q = octaves[std::size_t(o)];
        const R c = R(q.turnCos), s = R(q.turnSin), f = R(q.frequency);
        const R rx = x * c - y * s;
        const R ry = x * s + y * c;
        const R n = Basis2<R>(p.basis, rx * f + R(q.offsetX), ry * f + R(q.offsetY), q.seed);
        AddOctave(sums, p, o, n);
    }
    return Finish(sums, p);
}

// ---- 3D noise on a body ------------------------------------------------------------------------

// PerlinGradient3 (include/terraingen/noise/perlin.h:38-43).
template <class R> inline R Grad3(u32 hash, R x, R y, R z) {
    const u32 h = hash & 15u;
    const R u = h < 8u ? x : y;
    const R v = h < 4u ? y : (h == 12u || h == 14u ? x : z);
    return ((h & 1u) == 0u ? u : -u) + ((h & 2u) == 0u ? v : -v);
}

// PerlinNoise3 over a split lattice point (perlin.h:56-63,113-128): each corner's hash is the x
// word of pcg3d(x, y, z + PcgHash(seed)) (PerlinLatticeKey3, lattice.h:99-102).
template <class R>
inline R Perlin3(std::int32_t ix, std::int32_t iy, std::int32_t iz, R fx, R fy, R fz, u32 seed) {
    const u32 zKey = Pcg(seed);
    R c[8];
    for (u32 corner = 0; corner < 8u; ++corner) {
        const u32 dx = corner & 1u, dy = (corner >> 1u) & 1u, dz = corner >> 2u;
        const u32 hash = Pcg3({u32(ix) + dx, u32(iy) + dy, u32(iz) + dz + zKey}).x;
        c[corner] = Grad3<R>(hash, fx - R(dx), fy - R(dy), fz - R(dz));
    }
    const R u = Fade(fx), v = Fade(fy), w = Fade(fz);
    const R nearSide = Mix(v, Mix(u, c[0], c[1]), Mix(u, c[2], c[3]));
    const R farSide = Mix(v, Mix(u, c[4], c[5]), Mix(u, c[6], c[7]));
    return std::clamp(Mix(w, nearSide, farSide), R(-1), R(1));
}

// SimplexCorner3 (include/terraingen/noise/simplex.h:27-35); the corner hash is Perlin's
// (PerlinCornerHash3, perlin.h:49-51).
template <class R>
inline R SimplexCorner3(u32 ci, u32 cj, u32 ck, R dx, R dy, R dz, u32 zKey) {
    const R t = R(0.5) - dx * dx - dy * dy - dz * dz;
    if (t <= R(0)) {
        return R(0);
    }
    const R t2 = t * t;
    return t2 * t2 * Grad3<R>(Pcg3({ci, cj, ck + zKey}).x, dx, dy, dz);
}

// SimplexNoise3 over a split lattice point (simplex.h:74-114): the skew by (x + y + z) / 3 split
// into the cell sum's integer third and a fraction; the unskew constant is the reference's float
// 1 / 6.
template <class R>
inline R Simplex3(std::int32_t ix, std::int32_t iy, std::int32_t iz, R fx, R fy, R fz, u32 seed) {
    const u32 zKey = Pcg(seed);
    const std::int32_t cellSum = std::int32_t(u32(ix) + u32(iy) + u32(iz));
    std::int32_t third = cellSum / 3;
    if (cellSum - 3 * third < 0) {
        third -= 1;
    }
    const R remainder = R(cellSum - 3 * third);
    const R skew = (remainder + fx + fy + fz) / R(3);
    const R stepX = std::floor(fx + skew);
    const R stepY = std::floor(fy + skew);
    const R stepZ = std::floor(fz + skew);
    const R unskew = (remainder + stepX + stepY + stepZ) / R(6);
    const R x0 = fx - stepX + unskew;
    const R y0 = fy - stepY + unskew;
    const R z0 = fz - stepZ + unskew;
    const u32 i1 = x0 >= y0 && x0 >= z0 ? 1u : 0u;
    const u32 j1 = y0 > x0 && y0 >= z0 ? 1u : 0u;
    const u32 k1 = z0 > x0 && z0 > y0 ? 1u : 0u;
    const u32 i2 = x0 >= y0 || x0 >= z0 ? 1u : 0u;
    const u32 j2 = y0 > x0 || y0 >= z0 ? 1u : 0u;
    const u32 k2 = z0 > x0 || z0 > y0 ? 1u : 0u;
    const R g = R(1.0f / 6.0f);
    const R g2 = R(2.0f * (1.0f / 6.0f));
    const R g3 = R(3.0f * (1.0f / 6.0f));
    const u32 i = u32(ix) + u32(third) + u32(std::int32_t(stepX));
    const u32 j = u32(iy) + u32(third) + u32(std::int32_t(stepY));
    const u32 k = u32(iz) + u32(third) + u32(std::int32_t(stepZ));
    const R sum = SimplexCorner3<R>(i, j, k, x0, y0, z0, zKey) +
                  SimplexCorner3<R>(i + i1, j + j1, k + k1, x0 - R(i1) + g, y0 - R(j1) + g,
                                    z0 - R(k1) + g, zKey) +
                  SimplexCorner3<R>(i + i2, j + j2, k + k2, x0 - R(i2) + g2, y0 - R(j2) + g2,
                                    z0 - R(k2) + g2, zKey) +
                  SimplexCorner3<R>(i + 1u, j + 1u, k + 1u, x0 - R(1) + g3, y0 - R(1) + g3,
                                    z0 - R(1) + g3, zKey);
    return std::clamp(R(76) * sum, R(-1), R(1));
}

// ValueNoise3 (value.h:22-26,57-73): corner values from pcg3d(x, y, z ^ PcgHash(seed)).y
// (ValueLatticeKey3, lattice.h:108-111), blended with the quintic fade.
template <class R>
inline R Value3(std::int32_t ix, std::int32_t iy, std::int32_t iz, R fx, R fy, R fz, u32 seed) {
    const u32 key = Pcg(seed);
    R c[8];
    for (u32 corner = 0; corner < 8u; ++corner) {
        c[corner] = LatticeValue<R>({u32(ix) + (corner & 1u), u32(iy) + ((corner >> 1u) & 1u),
                                     (u32(iz) + (corner >> 2u)) ^ key});
    }
    const R u = Fade(fx), v = Fade(fy), w = Fade(fz);
    const R nearSide = Mix(v, Mix(u, c[0], c[1]), Mix(u, c[2], c[3]));
    const R farSide = Mix(v, Mix(u, c[4], c[5]), Mix(u, c[6], c[7]));
    return Mix(w, nearSide, farSide);
}

// CellGapSquared (worley.h:24-32): the squared gap along one axis from the point to a cell d
// cells away.
template <class R> inline R CellGap2(std::int32_t d, R f) {
    R gap = R(0);
    if (d > 0) {
        gap = R(d) - f;
    } else if (d < 0) {
        gap = f - R(d + 1);
    }
    return gap * gap;
}

// Worley3 (worley.h:75-114): the 3 x 3 x 3 cells around the point, then the outer shell of the
// 5 x 5 x 5 cells not yet farther than F2; feature points from pcg3d(x ^ PcgHash(seed), y, z)
// (CellularLatticeKey3, lattice.h:116-119), the nearest cell's value from PcgHash of its words.
template <class R>
inline Worley<R> Worley3(std::int32_t ix, std::int32_t iy, std::int32_t iz, R fx, R fy, R fz,
                         u32 seed) {
    const u32 seedKey = Pcg(seed);
    R d1 = R(1e30f), d2 = R(1e30f);
    u32 nearest = 0u;
    for (std::int32_t shell = 0; shell < 2; ++shell) {
        for (std::int32_t dz = -2; dz <= 2; ++dz) {
            for (std::int32_t dy = -2; dy <= 2; ++dy) {
                for (std::int32_t dx = -2; dx <= 2; ++dx) {
                    const bool inner = std::abs(dx) <= 1 && std::abs(dy) <= 1 && std::abs(dz) <= 1;
                    if (inner != (shell == 0)) {
                        continue;
                    }
                    if (shell == 1 &&
                        CellGap2(dx, fx) + CellGap2(dy, fy) + CellGap2(dz, fz) >= d2) {
                        continue;
                    }
                    const U3 h =
                        Pcg3({(u32(ix) + u32(dx)) ^ seedKey, u32(iy) + u32(dy), u32(iz) + u32(dz)});
                    const R px = R(dx) + R(UnitFloat(h.x)) - fx;
                    const R py = R(dy) + R(UnitFloat(h.y)) - fy;
                    const R pz = R(dz) + R(UnitFloat(h.z)) - fz;
                    const R d = px * px + py * py + pz * pz;
                    if (d < d1) {
                        d2 = d1;
                        d1 = d;
                        nearest = Pcg(h.x ^ h.y ^ h.z);
                    } else if (d < d2) {
                        d2 = d;
                    }
                }
            }
        }
    }
    return {std::sqrt(d1), std::sqrt(d2), R(2) * R(UnitFloat(nearest)) - R(1)};
}

// BasisNoise3 (fractal.h:115-136): Worley distances over their 3D bounds sqrt(3) and sqrt(6).
template <class R>
inline R Basis3(u32 basis, std::int32_t ix, std::int32_t iy, std::int32_t iz, R fx, R fy, R fz,
                u32 seed) {
    const R invSqrt3 = R(0.577350269189625765f);
    const R invSqrt6 = R(0.408248290463863016f);
    switch (basis) {
    case TG_BASIS_SIMPLEX:
        return Simplex3<R>(ix, iy, iz, fx, fy, fz, seed);
    case TG_BASIS_VALUE:
        return Value3<R>(ix, iy, iz, fx, fy, fz, seed);
    case TG_BASIS_WORLEY_F1:
        return R(2) * std::min(Worley3<R>(ix, iy, iz, fx, fy, fz, seed).f1 * invSqrt3, R(1)) -
               R(1);
    case TG_BASIS_WORLEY_F2:
        return R(2) * std::min(Worley3<R>(ix, iy, iz, fx, fy, fz, seed).f2 * invSqrt6, R(1)) -
               R(1);
    case TG_BASIS_WORLEY_F2_MINUS_F1: {
        const Worley<R> w = Worley3<R>(ix, iy, iz, fx, fy, fz, seed);
        return R(2) * std::min((w.f2 - w.f1) * invSqrt6, R(1)) - R(1);
    }
    case TG_BASIS_VORONOI:
        return Worley3<R>(ix, iy, iz, fx, fy, fz, seed).value;
    default:
        return Perlin3<R>(ix, iy, iz, fx, fy, fz, seed);
    }
}

// Fractal3 (include/terraingen/noise/fractal.h:284-317) over every basis: FP64 turns about the
// axis (1, 2, 3), FP64 frequency by repeated multiplication, FP64 lattice coordinates split before
// narrowing (SplitLatticePrecise3, lattice.h:59-69), octave sums in R (FP32 as the reference;
// FP64 for the FP64 oracle, where the fractions are not narrowed either).
template <class R = float> inline R Fractal3(double x, double y, double z, const Params& p) {
    static constexpr double kTurn[3][3] = {
        {-0.6132711010727254, -0.293401570821183, 0.7333580809050306},
        {0.7897926788435601, -0.24097777005594262, 0.5640542870894418},
        {0.011228581128535142, 0.9251190369776896, 0.37951111497202883}};
    double frequency = 1.0 / static_cast<double>(p.wavelength);
    double turn[3][3] = {{1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}};
    Sums<R> sums;
    for (int o = 0; o < p.octaves; ++o) {
        const U3 h = Pcg3({p.seed, u32(o), 0u});
        const U3 g = Pcg3({p.seed, u32(o), 1u});
        const double rx = turn[0][0] * x + turn[0][1] * y + turn[0][2] * z;
        const double ry = turn[1][0] * x + turn[1][1] * y + turn[1][2] * z;
        const double rz = turn[2][0] * x + turn[2][1] * y + turn[2][2] * z;
        const double px = rx * frequency + static_cast<double>(UnitFloat(h.y));
        const double py = ry * frequency + static_cast<double>(UnitFloat(h.z));
        const double pz = rz * frequency + static_cast<double>(UnitFloat(g.x));
        const double cx = std::floor(px), cy = std::floor(py), cz = std::floor(pz);
        const R n = Basis3<R>(p.basis, std::int32_t(static_cast<std::int64_t>(cx)),
                              std::int32_t(static_cast<std::int64_t>(cy)),
                              std::int32_t(static_cast<std::int64_t>(cz)), R(px - cx), R(py - cy),
                              R(pz - cz), h.x);
        AddOctave(sums, p, o, n);
        frequency *= static_cast<double>(p.lacunarity);
        double next[3][3] = {};
        for (int row = 0; row < 3; ++row) {
            for (int column = 0; column < 3; ++column) {
                next[row][column] = kTurn[row][0] * turn[0][column] +
                                    kTurn[row][1] * turn[1][column] +
                                    kTurn[row][2] * turn[2][column];
            }
        }
        std::memcpy(turn, next, sizeof(turn));
    }
    return Finish(sums, p);
}

}  // namespace orc
