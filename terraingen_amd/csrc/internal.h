// SPDX-License-Identifier: BSD-3-Clause
//
// Internals shared by the translation units of libtg_hip.so: the handles behind the C ABI and the
// error plumbing (every HIP status checked, every failure a message, never an exception across the
// ABI).
#pragma once

#include "tg_hip.h"

#include "noise_plan.h"

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdio>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

namespace tg {
struct MpdTerms;
struct PreciseTerms;
}

struct tg_dev {
    int ordinal = 0;
    std::string name;
    std::mutex mutex;  // runs are serialized, as ComputeDevice requires (src/gpu/device.h:48)
    // This device's copy of the library's static scratch (tg_hip_batch_run_static), from
    // hipGetSymbolAddress at tg_hip_open; null, with the reason, if that failed.
    unsigned char* scratch = nullptr;
    std::string scratchError;
};

// A boulder set (boulders.hip): the boulders of one grid spacing on the device, binned into the
// grid's 16 x 16-cell tiles.
struct tg_boulder_set {
    tg_dev* dev = nullptr;
    tg_boulder_kernel kernel{};
    tg_boulder* boulders = nullptr;
    uint32_t* tiles = nullptr;
    uint32_t* entries = nullptr;
    float* lows = nullptr;
    float* cull = nullptr;  // per boulder: squared distance beyond which its relief is 0
    size_t count = 0;
    size_t lowCount = 0;
    uint32_t tilesAcross = 0;
    uint32_t tileRows = 0;
    uint32_t tileSide = 16;
};

namespace tgi {

// The static scratch of tg_hip_batch_run_static, one per device: the queue counters, then the
// staged tiles, bins and entries of one batch, at these byte offsets.
constexpr size_t kStaticCounterBytes = 2048;
constexpr size_t kStaticTilesAt = kStaticCounterBytes;
constexpr size_t kStaticBinsAt = kStaticTilesAt + size_t(TG_STATIC_MAX_TILES) * sizeof(tg_tile_desc);
constexpr size_t kStaticEntriesAt = kStaticBinsAt + size_t(TG_STATIC_MAX_BINS) * sizeof(tg_bin);
constexpr size_t kStaticScratchBytes =
    kStaticEntriesAt + size_t(TG_STATIC_MAX_ENTRIES) * sizeof(uint32_t);
static_assert(kStaticBinsAt % 256 == 0 && kStaticEntriesAt % 256 == 0, "scratch alignment");

}  // namespace tgi

struct tg_crater_set {
    tg_dev* dev = nullptr;
    tg_crater_shape* kernel = nullptr;
    tg_crater_record* records = nullptr;
    float* data = nullptr;
    uint32_t* tiles = nullptr;
    uint32_t* entries = nullptr;
    tg_crater_frame* frames = nullptr;
    tg_split_point* points = nullptr;
    float* lows = nullptr;
    float* cull = nullptr;  // per record: squared round distance (m^2) beyond which it is inert
    float* harm = nullptr;  // A cos(phi), A sin(phi) in the data's harmonic slots
    float4* outline = nullptr;  // per record: cos and sin of the heading, sqrt(ellipticity), 1/that
    tg::MpdTerms* terms = nullptr;  // per record: the hoisted MPD path's constants
    tg::PreciseTerms* precise = nullptr;  // per record: FP64 terms of km-scale craters
    bool allFast = false;           // every record takes the hoisted path
    bool anyPrecise = false;        // some record takes the hoisted path in FP64
    size_t pointCount = 0;
    size_t lowCount = 0;
    uint32_t tilesAcross = 0;
    uint32_t tileRows = 0;
    uint32_t tileSide = 16;
    bool onBody = false;
};

namespace tgi {

// A number no other object of the library has had or will have: captured graphs are keyed on the
// generations of what they read rather than on addresses, which a freed object's successor may
// reuse.
inline uint64_t NextGeneration() {
    static std::atomic<uint64_t> next{1};
    return next.fetch_add(1, std::memory_order_relaxed);
}

}  // namespace tgi

struct tg_noise_plan {
    tg_dev* dev = nullptr;
    tg::NoisePlan* device = nullptr;
    tg::NoisePlan host{};
    uint64_t generation = tgi::NextGeneration();
};

namespace tgi {

inline void SetErr(char* err, size_t len, const std::string& message) {
    if (err != nullptr && len > 0) {
        std::snprintf(err, len, "%s", message.c_str());
    }
}

struct HipError {
    std::string message;
};

inline void Check(hipError_t status, const char* what) {
    if (status != hipSuccess) {
        throw HipError{std::string(what) + ": " + hipGetErrorString(status)};
    }
}

// Device memory freed on scope exit.
template <class T> struct DeviceArray {
    T* ptr = nullptr;
    DeviceArray() = default;
    explicit DeviceArray(size_t count) {
        if (count > 0) {
            Check(hipMalloc(reinterpret_cast<void**>(&ptr), count * sizeof(T)), "hipMalloc");
        }
    }
    DeviceArray(const DeviceArray&) = delete;
    DeviceArray& operator=(const DeviceArray&) = delete;
    ~DeviceArray() {
        if (ptr != nullptr) {
            (void)hipFree(ptr);
        }
    }
    void Upload(const void* src, size_t bytes) {
        if (bytes > 0) {
            Check(hipMemcpy(ptr, src, bytes, hipMemcpyHostToDevice), "hipMemcpy to device");
        }
    }
};

}  // namespace tgi

namespace tgi {

template <class F> int Guard(char* err, size_t errLen, F body) {
    try {
        body();
        return TG_OK;
    } catch (const HipError& e) {
        SetErr(err, errLen, e.message);
        return TG_ERR_DEVICE;
    } catch (const std::bad_alloc&) {
        SetErr(err, errLen, "host allocation failed");
        return TG_ERR_DEVICE;
    }
}

// Host-side derived per-crater terms (tg_hip.hip): the squared cull distance per record, and the
// fast outline path's (A cos phi, A sin phi) per harmonic slot and heading/ellipse terms per record.
std::vector<float> CullRadii(const tg_crater_shape& k, const tg_crater_record* records, size_t n,
                             const float* data, size_t dataCount);
void OutlineTerms(const tg_crater_record* records, size_t n, const float* data, size_t dataCount,
                  std::vector<float>& harm, std::vector<float4>& outline);

// Boulder sets (boulders.hip): validation of a bound batch, and the strip launch.
void ValidateBoulderBatch(const tg_boulder_kernel& k, const tg_boulder* boulders, size_t n,
                          const uint32_t* tiles, size_t tileWords, const uint32_t* entries,
                          size_t entryCount);
void LaunchBoulders(const tg_boulder_set& set, float* heights, const float* xs, const float* ys,
                    const tg_kernel_params& params, hipStream_t stream);

// Per-record MpdTerms of records already on the device, computed there (tg_hip.hip) into a new
// device array of n entries (at least one).
tg::MpdTerms* DeviceMpdTerms(const tg_crater_shape* kernelOnDevice,
                             const tg_crater_record* recordsOnDevice, size_t n);
// FP64 terms of every crater (tg::MakePreciseTerms), with frame axes when frames are given.
tg::PreciseTerms* DevicePreciseTerms(const tg_crater_shape* kernelOnDevice,
                                     const tg_crater_record* recordsOnDevice,
                                     const tg_crater_frame* framesOnDevice, size_t n);

// True when every record takes the hoisted MPD path (MakeMpdTerms's `fast`), decided on the host
// by the same conditions (tg_hip.hip).
bool AllMpdFast(const tg_crater_shape& k, const tg_crater_record* records, size_t n);
// Whether any record takes the hoisted path in FP64 (tg::kFastPrecise).
bool AnyPrecise(const tg_crater_shape& k, const tg_crater_record* records, size_t n);

// Throws HipError unless every index a crater batch holds stays inside its arrays (tg_hip.hip).
void ValidateCraterBatch(const tg_crater_shape& k, const tg_crater_record* records, size_t n,
                         size_t dataCount, const uint32_t* tiles, size_t tileWords,
                         const uint32_t* entries, size_t entryCount, size_t frameCount);

__device__ __forceinline__ bool IsNoData(const tg_kernel_params& p, float height) {
    if (p.no_data == TG_NODATA_NAN) {
        return isnan(height);
    }
    return p.no_data != TG_NODATA_NONE && height == p.no_data_value;
}

}  // namespace tgi
