// SPDX-License-Identifier: BSD-3-Clause
//
// ORACLE — test infrastructure only. Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load this library, and only as the checker or the CPU baseline; the hip
// backend never calls it.
//
// A CPU restatement of the reference's noise (PCG hashes, Perlin/simplex/value/Worley bases,
// fBm/ridged/billow/hybrid fractals, stacks with domain warp and blend; in 2D on a plane and in 3D
// at body-fixed positions) in two precisions:
//   * `float`  — the reference-equivalent path: every operation in the reference's order, built
//                with -ffp-contract=off as the reference is (CMakeLists.txt:32-38), so its bits are
//                the reference CPU backend's bits;
//   * `double` — the FP64 oracle: the per-octave parameters (seeds, lattice offsets, the FP32
//                golden-angle recurrence and frequencies) exactly as the reference derives them,
//                and all per-cell arithmetic in double, so the difference to it is the per-cell
//                FP32 rounding error.
// Pinned by: tests/golden/noise/kernels_test.cpp:21-103 (golden values), oracle/_ref (the
// reference's own headers compiled beside it) and brute-force invariants in tests/.

#include "noise_core.h"
#include "tg_layouts.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

using namespace orc;

namespace {

Params ParamsOf(const tg_fractal& f) {
    return {f.seed, f.wavelength, f.amplitude, f.octaves, f.lacunarity,
            f.gain, f.basis,      f.type,      f.offset};
}

// include/terraingen/noise/fractals.h:98-104 (warp fields) and :110-135 (warp and unit value).
template <class R> R FractalValue(R x, R y, const tg_fractal& layer) {
    Params unit = ParamsOf(layer);
    unit.amplitude = 1.0f;
    if (layer.warp_amplitude > 0.0f) {
        Params warp{Pcg(layer.seed ^ 0x5bd1e995u), layer.warp_wavelength, layer.warp_amplitude, 4,
                    2.0f, 0.5f, TG_BASIS_PERLIN, TG_FRACTAL_FBM, 1.0f};
        const R dx = Fractal2(x, y, warp);
        warp.seed = Pcg(warp.seed);
        const R dy = Fractal2(x, y, warp);
        return Fractal2(x + dx, y + dy, unit);
    }
    return Fractal2(x, y, unit);
}

// include/terraingen/noise/fractals.h:151-196
template <class R> R Stack(R x, R y, const tg_fractal* fractals, std::size_t count) {
    R height = R(0);
    bool started = false;
    for (std::size_t i = 0; i < count; ++i) {
        const tg_fractal& layer = fractals[i];
        if (!layer.enabled) {
            continue;
        }
        const R value = FractalValue(x, y, layer);
        const R scaled = R(layer.amplitude) * value;
        if (!started) {
            height = scaled;
            started = true;
            continue;
        }
        switch (layer.blend) {
        case TG_BLEND_MAX:
            height = std::max(height, scaled);
            break;
        case TG_BLEND_MIN:
            height = std::min(height, scaled);
            break;
        case TG_BLEND_MASK:
            height *= R(0.5) * (value + R(1));
            break;
        default:
            height += scaled;
            break;
        }
    }
    return height;
}

// EvaluateFractal3 (include/terraingen/noise/fractals.h:212-243): the fractal's unit value at
// body-fixed (x, y, z) meters, warped by three Fractal3 fields (DomainWarp3) when warpAmplitude
// > 0; the offsets are added in FP64.
template <class R> R FractalValue3(double x, double y, double z, const tg_fractal& layer) {
    Params unit = ParamsOf(layer);
    unit.amplitude = 1.0f;
    if (layer.warp_amplitude > 0.0f) {
        Params warp{Pcg(layer.seed ^ 0x5bd1e995u), layer.warp_wavelength, layer.warp_amplitude, 4,
                    2.0f, 0.5f, TG_BASIS_PERLIN, TG_FRACTAL_FBM, 1.0f};
        const double dx = static_cast<double>(Fractal3<R>(x, y, z, warp));
        warp.seed = Pcg(warp.seed);
        const double dy = static_cast<double>(Fractal3<R>(x, y, z, warp));
        warp.seed = Pcg(warp.seed);
        const double dz = static_cast<double>(Fractal3<R>(x, y, z, warp));
        return Fractal3<R>(x + dx, y + dy, z + dz, unit);
    }
    return Fractal3<R>(x, y, z, unit);
}

// EvaluateFractals3 (include/terraingen/noise/fractals.h:245-269) at body-fixed (x, y, z) meters,
// in precision R.
template <class R>
R Stack3(double x, double y, double z, const tg_fractal* fractals, std::size_t count) {
    R height = R(0);
    bool started = false;
    for (std::size_t i = 0; i < count; ++i) {
        const tg_fractal& layer = fractals[i];
        if (!layer.enabled) {
            continue;
        }
        const R value = FractalValue3<R>(x, y, z, layer);
        const R scaled = R(layer.amplitude) * value;
        if (!started) {
            height = scaled;
            started = true;
            continue;
        }
        switch (layer.blend) {
        case TG_BLEND_MAX:
            height = std::max(height, scaled);
            break;
        case TG_BLEND_MIN:
            height = std::min(height, scaled);
            break;
        case TG_BLEND_MASK:
            height *= R(0.5) * (value + R(1));
            break;
        default:
            height += scaled;
            break;
        }
    }
    return height;
}

bool IsNoData(const tg_kernel_params& p, float h) {
    if (p.no_data == TG_NODATA_NAN) {
        return std::isnan(h);
    }
    return p.no_data != TG_NODATA_NONE && h == p.no_data_value;
}

// Runs `rowTask` over rows on `threads` workers, each taking every `threads`-th row; like the
// reference's ParallelFor (src/util/parallel.cpp:126-151) results do not depend on the count.
template <class F> void ForRows(std::size_t rows, unsigned threads, F rowTask) {
    if (threads == 0) {
        threads = std::max(1u, std::thread::hardware_concurrency());
    }
    threads = unsigned(std::min<std::size_t>(threads, std::max<std::size_t>(rows, 1)));
    if (threads <= 1) {
        for (std::size_t r = 0; r < rows; ++r) {
            rowTask(r);
        }
        return;
    }
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < threads; ++t) {
        pool.emplace_back([&, t] {
            for (std::size_t r = t; r < rows; r += threads) {
                rowTask(r);
            }
        });
    }
    for (std::thread& th : pool) {
        th.join();
    }
}

}  // namespace

extern "C" {

u32 or_pcg_hash(u32 v) {
    return Pcg(v);
}

void or_pcg3d(const u32* in, u32* out) {
    const U3 r = Pcg3({in[0], in[1], in[2]});
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
}

float or_hash_to_unit_float(u32 h) {
    return UnitFloat(h);
}

float or_perlin2(float x, float y, u32 seed) {
    return Perlin2At(x, y, seed);
}

float or_simplex2(float x, float y, u32 seed) {
    return Simplex2(x, y, seed);
}

float or_value2(float x, float y, u32 seed) {
    return Value2(x, y, seed);
}

void or_worley2(float x, float y, u32 seed, float* out3) {
    const Worley<float> w = Worley2(x, y, seed);
    out3[0] = w.f1;
    out3[1] = w.f2;
    out3[2] = w.value;
}

float or_basis2(u32 basis, float x, float y, u32 seed) {
    return Basis2<float>(basis, x, y, seed);
}

// FractalParams given as a tg_fractal (its `enabled`, `blend` and warp fields unused).
float or_fractal2(float x, float y, const tg_fractal* f) {
    return Fractal2<float>(x, y, ParamsOf(*f));
}

double or_fractal2_f64(double x, double y, const tg_fractal* f) {
    return Fractal2<double>(x, y, ParamsOf(*f));
}

float or_eval_fractals(float x, float y, const tg_fractal* fractals, std::size_t count) {
    return Stack<float>(x, y, fractals, count);
}

double or_eval_fractals_f64(double x, double y, const tg_fractal* fractals, std::size_t count) {
    return Stack<double>(x, y, fractals, count);
}

// Fractal3 of one fractal (amplitude included), and the stack, at body-fixed (x, y, z) meters.
float or_fractal3(double x, double y, double z, const tg_fractal* f) {
    return Fractal3<float>(x, y, z, ParamsOf(*f));
}

// BasisNoise3 at a point in lattice units split in FP32 (SplitLattice3, lattice.h:39-46).
float or_basis3(u32 basis, float x, float y, float z, u32 seed) {
    const float cx = std::floor(x), cy = std::floor(y), cz = std::floor(z);
    return Basis3<float>(basis, std::int32_t(cx), std::int32_t(cy), std::int32_t(cz), x - cx,
                         y - cy, z - cz, seed);
}

// Worley3 at a point in lattice units: out = {F1, F2, the nearest cell's value}.
void or_worley3(float x, float y, float z, u32 seed, float* out) {
    const float cx = std::floor(x), cy = std::floor(y), cz = std::floor(z);
    const Worley<float> w = Worley3<float>(std::int32_t(cx), std::int32_t(cy), std::int32_t(cz),
                                           x - cx, y - cy, z - cz, seed);
    out[0] = w.f1;
    out[1] = w.f2;
    out[2] = w.value;
}

float or_eval_fractals3(double x, double y, double z, const tg_fractal* fractals,
                        std::size_t count) {
    return Stack3<float>(x, y, z, fractals, count);
}

double or_eval_fractals3_f64(double x, double y, double z, const tg_fractal* fractals,
                             std::size_t count) {
    return Stack3<double>(x, y, z, fractals, count);
}

// The noise of a terrain on a body, as SyntheticSource and the noise layer add it there
// (src/noise/surface_noise.cpp:9-20): heights[i] += stack at positions[3 i .. 3 i + 2].
void or_body_noise_grid(const tg_noise_layer_kernel* kernel, const double* positions,
                        std::size_t cells, const tg_kernel_params* params, float* heights,
                        unsigned threads) {
    const std::size_t rows = (cells + 1023) / 1024;
    ForRows(rows, threads, [&](std::size_t r) {
        for (std::size_t i = r * 1024; i < std::min(cells, (r + 1) * 1024); ++i) {
            if (!IsNoData(*params, heights[i])) {
                const double* p = positions + 3 * i;
                heights[i] = heights[i] + Stack3<float>(p[0], p[1], p[2], kernel->fractals,
                                                        kernel->count);
            }
        }
    });
}

// The FP64 oracle of or_body_noise_grid: heights[i] += the stack in FP64 (every cell holds data).
void or_body_noise_grid_f64(const tg_noise_layer_kernel* kernel, const double* positions,
                            std::size_t cells, double* heights, unsigned threads) {
    const std::size_t rows = (cells + 1023) / 1024;
    ForRows(rows, threads, [&](std::size_t r) {
        for (std::size_t i = r * 1024; i < std::min(cells, (r + 1) * 1024); ++i) {
            const double* p = positions + 3 * i;
            heights[i] += Stack3<double>(p[0], p[1], p[2], kernel->fractals, kernel->count);
        }
    });
}

// The noise layer over a grid, as the reference's CPU path computes it (noise_layer.cpp:28-58):
// heights[r * width + c] += stack(xs[c], ys[r]) for each cell holding data, rows over `threads`.
void or_noise_grid(const tg_noise_layer_kernel* kernel, const float* xs, const float* ys,
                   std::size_t width, std::size_t height, const tg_kernel_params* params,
                   float* heights, unsigned threads) {
    ForRows(height, threads, [&](std::size_t r) {
        for (std::size_t c = 0; c < width; ++c) {
            float& cell = heights[r * width + c];
            if (!IsNoData(*params, cell)) {
                cell = cell + Stack<float>(xs[c], ys[r], kernel->fractals, kernel->count);
            }
        }
    });
}

// The FP64 oracle of or_noise_grid: cell positions given in double, heights returned in double.
void or_noise_grid_f64(const tg_noise_layer_kernel* kernel, const double* xs, const double* ys,
                       std::size_t width, std::size_t height, double* heights, unsigned threads) {
    ForRows(height, threads, [&](std::size_t r) {
        for (std::size_t c = 0; c < width; ++c) {
            heights[r * width + c] += Stack<double>(xs[c], ys[r], kernel->fractals, kernel->count);
        }
    });
}

}  // extern "C"
