// SPDX-License-Identifier: BSD-3-Clause
//
// Measurement aids, no counterpart in the reference.
// A VALU issue probe: independent dependency chains per lane, no memory traffic in the loop, run
// at full occupancy and timed with HIP events. It measures the vector-ALU issue ceiling the
// roofline of the shaping kernels is read against: FP32 FMAs (the kernels are built without packed
// FP32 forms) and 32-bit integer logic/add, the two classes of their arithmetic.
// A store-shape probe: the batched kernels' output writes alone (tiles packed one after another,
// 16 x 16 bins in contiguous runs per wave) in three shapes, for reading the HBM bytes each shape
// writes with the WRITE_SIZE counter.

#include "internal.h"

#include <hip/hip_runtime.h>

using namespace tgi;

namespace {

constexpr int kChains = 8;

// Eight v_fma_f32 per iteration.
__global__ void __launch_bounds__(256) FmaChains(float* out, uint32_t iters) {
    float a[kChains];
    for (int i = 0; i < kChains; ++i) {
        a[i] = float(threadIdx.x) * 0.001f + float(i);
    }
    for (uint32_t it = 0; it < iters; ++it) {
#pragma unroll
        for (int i = 0; i < kChains; ++i) {
            a[i] = __builtin_fmaf(a[i], 0.999f, 0.5f);
        }
    }
    float s = 0.0f;
    for (int i = 0; i < kChains; ++i) {
        s += a[i];
    }
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// Eight v_xor_b32 and eight v_add_u32 per iteration.
__global__ void __launch_bounds__(256) IntChains(uint32_t* out, uint32_t iters) {
    uint32_t a[kChains];
    for (int i = 0; i < kChains; ++i) {
        a[i] = threadIdx.x * 747796405u + uint32_t(i);
    }
    for (uint32_t it = 0; it < iters; ++it) {
#pragma unroll
        for (int i = 0; i < kChains; ++i) {
            a[i] = (a[i] ^ 0x9E3779B9u) + 0x7F4A7C15u;
        }
    }
    uint32_t s = 0;
    for (int i = 0; i < kChains; ++i) {
        s ^= a[i];
    }
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// One wave per run of `binsPerWave` consecutive bins of `tiles` tiles of width x width floats
// packed one after another; shape 0: a bin's four 8 x 8 squares (the batched kernels' stores);
// 1: four 16 x 4 strips; 2: rows of 64 floats, a wave writing four bins across at a time.
template <int kShape>
__global__ void __launch_bounds__(256) StoreShapes(float* __restrict__ out, uint32_t width,
                                                   uint32_t tiles, uint32_t binsPerWave) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = (blockIdx.x * 256u + threadIdx.x) >> 6u;
    const uint32_t across = (width + 15u) / 16u;
    const uint64_t binsPerTile = uint64_t(across) * across;
    const uint64_t total = binsPerTile * tiles;
    const uint64_t first = uint64_t(wave) * binsPerWave;
    for (uint64_t b = first; b < first + binsPerWave && b < total; ++b) {
        const uint32_t tile = uint32_t(b / binsPerTile);
        const uint32_t local = uint32_t(b % binsPerTile);
        const uint32_t col0 = (local % across) * 16u, row0 = (local / across) * 16u;
        float* base = out + uint64_t(tile) * width * width;
        if constexpr (kShape == 0) {
            for (uint32_t sq = 0; sq < 4u; ++sq) {
                const uint32_t col = col0 + (sq & 1u) * 8u + (lane & 7u);
                const uint32_t row = row0 + (sq >> 1u) * 8u + (lane >> 3u);
                if (col < width && row < width) {
                    base[uint64_t(row) * width + col] = float(col + row);
                }
            }
        } else if constexpr (kShape == 1) {
            for (uint32_t k = 0; k < 4u; ++k) {
                const uint32_t col = col0 + (lane & 15u);
                const uint32_t row = row0 + k * 4u + (lane >> 4u);
                if (col < width && row < width) {
                    base[uint64_t(row) * width + col] = float(col + row);
                }
            }
        } else {
            if ((local % across) % 4u != 0u) {
                continue;  // the first bin of each four across writes their rows
            }
            for (uint32_t r = 0; r < 16u; ++r) {
                const uint32_t col = col0 + lane;
                const uint32_t row = row0 + r;
                if (col < width && row < width) {
                    base[uint64_t(row) * width + col] = float(col + row);
                }
            }
        }
    }
}

}  // namespace

extern "C" int tg_hip_store_probe(tg_dev* dev, int shape, uint32_t width, uint32_t tiles,
                                  float* out, size_t out_capacity, float* ms, char* err,
                                  size_t errLen) {
    if (dev == nullptr || out == nullptr || ms == nullptr || shape < 0 || shape > 2 ||
        width == 0 || tiles == 0 || uint64_t(width) * width * tiles > out_capacity) {
        SetErr(err, errLen, "tg_hip_store_probe: null argument, unknown shape or too small out");
        return TG_ERR_ARGUMENT;
    }
    return Guard(err, errLen, [&] {
        Check(hipSetDevice(dev->ordinal), "hipSetDevice");
        constexpr uint32_t kWaves = 8192;
        const uint32_t across = (width + 15u) / 16u;
        const uint64_t total = uint64_t(across) * across * tiles;
        const uint32_t per = uint32_t((total + kWaves - 1) / kWaves);
        hipEvent_t a = nullptr, b = nullptr;
        hipError_t status = hipEventCreate(&a);
        if (status == hipSuccess) {
            status = hipEventCreate(&b);
        }
        if (status == hipSuccess) {
            status = hipEventRecord(a, nullptr);
        }
        if (status == hipSuccess) {
            auto* fn = shape == 0 ? StoreShapes<0> : shape == 1 ? StoreShapes<1> : StoreShapes<2>;
            hipLaunchKernelGGL(fn, dim3(kWaves / 4u), dim3(256), 0, nullptr, out, width, tiles,
                               per);
            status = hipGetLastError();
        }
        if (status == hipSuccess) {
            status = hipEventRecord(b, nullptr);
        }
        if (status == hipSuccess) {
            status = hipEventSynchronize(b);
        }
        if (status == hipSuccess) {
            status = hipEventElapsedTime(ms, a, b);
        }
        if (a != nullptr) {
            (void)hipEventDestroy(a);
        }
        if (b != nullptr) {
            (void)hipEventDestroy(b);
        }
        Check(status, "running the store probe");
    });
}

extern "C" int tg_hip_valu_probe(tg_dev* dev, int kind, uint32_t iters, uint32_t blocks,
                                 float* ms, double* lane_ops, char* err, size_t errLen) {
    if (dev == nullptr || ms == nullptr || lane_ops == nullptr || (kind != 0 && kind != 1) ||
        blocks == 0 || iters == 0) {
        SetErr(err, errLen, "tg_hip_valu_probe: null argument, unknown kind or empty launch");
        return TG_ERR_ARGUMENT;
    }
    return Guard(err, errLen, [&] {
        Check(hipSetDevice(dev->ordinal), "hipSetDevice");
        DeviceArray<uint32_t> out(size_t(blocks) * 256);
        hipStream_t stream = nullptr;
        Check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
        hipEvent_t a = nullptr, b = nullptr;
        hipError_t status = hipEventCreate(&a);
        if (status == hipSuccess) {
            status = hipEventCreate(&b);
        }
        for (int rep = 0; rep < 2 && status == hipSuccess; ++rep) {  // the first run warms up
            status = hipEventRecord(a, stream);
            if (kind == 0) {
                hipLaunchKernelGGL(FmaChains, dim3(blocks), dim3(256), 0, stream,
                                   reinterpret_cast<float*>(out.ptr), iters);
            } else {
                hipLaunchKernelGGL(IntChains, dim3(blocks), dim3(256), 0, stream, out.ptr,
                                   iters);
            }
            if (status == hipSuccess) {
                status = hipGetLastError();
            }
            if (status == hipSuccess) {
                status = hipEventRecord(b, stream);
            }
            if (status == hipSuccess) {
                status = hipEventSynchronize(b);
            }
        }
        if (status == hipSuccess) {
            status = hipEventElapsedTime(ms, a, b);
        }
        if (a != nullptr) {
            (void)hipEventDestroy(a);
        }
        if (b != nullptr) {
            (void)hipEventDestroy(b);
        }
        (void)hipStreamDestroy(stream);
        Check(status, "running the VALU probe");
        *lane_ops = double(blocks) * 256.0 * double(iters) * (kind == 0 ? 8.0 : 16.0);
    });
}
