// SPDX-License-Identifier: BSD-3-Clause
//
// A VALU issue probe: independent dependency chains per lane, no memory traffic in the loop, run
// at full occupancy and timed with HIP events. It measures the vector-ALU issue ceiling the
// roofline of the shaping kernels is read against: FP32 FMAs (the kernels are built without packed
// FP32 forms) and 32-bit integer logic/add, the two classes of their arithmetic.

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

}  // namespace

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
