// SPDX-License-Identifier: BSD-3-Clause
//
// ORACLE — test infrastructure only (see noise_oracle.cpp). Built by oracle/Makefile into
// oracle/_ref/libtgref.so against the reference's own headers under /root/reference/include, so
// the tests can compare this repository's restatement with the reference's code itself. Nothing
// of the reference is copied here: this file only calls its header-only noise functions.

#include "terraingen/layers/kernels/noise_layer.h"
#include "terraingen/noise/fractals.h"

#include "tg_layouts.h"

#include <cstring>

using namespace terraingen;

namespace {

Fractal FromLayout(const tg_fractal& f) {
    Fractal out;
    out.enabled = f.enabled != 0;
    out.seed = f.seed;
    out.wavelength = f.wavelength;
    out.amplitude = f.amplitude;
    out.octaves = f.octaves;
    out.lacunarity = f.lacunarity;
    out.gain = f.gain;
    out.basis = static_cast<NoiseBasis>(f.basis);
    out.type = static_cast<FractalType>(f.type);
    out.offset = f.offset;
    out.blend = static_cast<FractalBlend>(f.blend);
    out.warpAmplitude = f.warp_amplitude;
    out.warpWavelength = f.warp_wavelength;
    return out;
}

}  // namespace

extern "C" {

std::uint32_t ref_pcg_hash(std::uint32_t v) {
    return PcgHash(v);
}

void ref_pcg3d(const std::uint32_t* in, std::uint32_t* out) {
    const UVec3 r = Pcg3dHash(UVec3{in[0], in[1], in[2]});
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
}

float ref_basis2(std::uint32_t basis, float x, float y, std::uint32_t seed) {
    return BasisNoise2(static_cast<NoiseBasis>(basis), x, y, seed);
}

float ref_fractal2(float x, float y, const tg_fractal* f) {
    return Fractal2(x, y, FractalParamsFrom(FromLayout(*f)));
}

float ref_basis3(std::uint32_t basis, float x, float y, float z, std::uint32_t seed) {
    return BasisNoise3(static_cast<NoiseBasis>(basis), SplitLattice3(x, y, z), seed);
}

float ref_fractal3(double x, double y, double z, const tg_fractal* f) {
    return Fractal3(x, y, z, FractalParamsFrom(FromLayout(*f)));
}

float ref_eval_fractals3(double x, double y, double z, const tg_fractal* fractals,
                         std::size_t count) {
    std::vector<Fractal> list;
    for (std::size_t i = 0; i < count; ++i) {
        list.push_back(FromLayout(fractals[i]));
    }
    return EvaluateFractals3(x, y, z, list.data(), list.size());
}

float ref_eval_fractals(float x, float y, const tg_fractal* fractals, std::size_t count) {
    std::vector<Fractal> list;
    for (std::size_t i = 0; i < count; ++i) {
        list.push_back(FromLayout(fractals[i]));
    }
    return EvaluateFractals(x, y, list.data(), list.size());
}

// The reference's noise-layer kernel struct built by its own MakeNoiseLayerKernel, as raw bytes,
// so the tests can check this repository's layout against it.
void ref_make_noise_layer_kernel(const tg_fractal* fractals, std::size_t count, void* out420) {
    std::vector<Fractal> list;
    for (std::size_t i = 0; i < count; ++i) {
        list.push_back(FromLayout(fractals[i]));
    }
    NoiseLayerKernel kernel = MakeNoiseLayerKernel(list);
    // The reference keeps `bool enabled` in one byte; its kernel upload relies on the padding being
    // zero, which a value-initialised struct guarantees only per member, so clear it here.
    unsigned char bytes[sizeof(NoiseLayerKernel)];
    std::memset(bytes, 0, sizeof bytes);
    for (std::uint32_t i = 0; i < kMaxFractals; ++i) {
        std::memcpy(bytes + i * sizeof(Fractal), &kernel.stack.fractals[i], sizeof(Fractal));
        bytes[i * sizeof(Fractal) + 1] = 0;
        bytes[i * sizeof(Fractal) + 2] = 0;
        bytes[i * sizeof(Fractal) + 3] = 0;
    }
    std::memcpy(bytes + kMaxFractals * sizeof(Fractal), &kernel.stack.count, sizeof(std::uint32_t));
    std::memcpy(out420, bytes, sizeof bytes);
}

// The reference's ApplyCell over a grid, one cell at a time.
void ref_noise_grid(const void* kernel420, const float* xs, const float* ys, std::size_t width,
                    std::size_t height, float* heights) {
    NoiseLayerKernel kernel;
    std::memcpy(&kernel, kernel420, sizeof kernel);
    for (std::size_t r = 0; r < height; ++r) {
        for (std::size_t c = 0; c < width; ++c) {
            float& cell = heights[r * width + c];
            cell = ApplyCell(kernel, cell, CellPosition{xs[c], ys[r]});
        }
    }
}

}  // extern "C"
