// SPDX-License-Identifier: BSD-3-Clause
//
// Boulder relief on the device: BoulderLayer's dome relief (src/layers/boulders/
// boulder_relief.cpp:426-440, DrawFeatures, sample_block.h:107-131) as a gather over boulders
// binned into 16 x 16-cell tiles. Every cell keeps its height before the layer as the base and
// takes the highest of its own height and base + weight x relief over the boulders reaching it;
// the maximum does not depend on the order the boulders are visited in, so the result equals the
// reference's feature-by-feature loop.

#include "boulder_device.h"
#include "internal.h"

#include <hip/hip_runtime.h>

#include <cstring>

using namespace tgi;

namespace {

constexpr uint32_t kTile = 16;

__global__ void __launch_bounds__(kTile * kTile)
    BoulderStrip(float* __restrict__ heights, const float* __restrict__ xs,
                 const float* __restrict__ ys, tg_boulder_kernel kernel,
                 const tg_boulder* __restrict__ boulders, const uint32_t* __restrict__ tiles,
                 const uint32_t* __restrict__ entries, const float* __restrict__ lows,
                 const float* __restrict__ cull, tg_kernel_params params) {
    // 8 x 8 cells per wave, as the crater kernel: a boulder's disc leaves few waves part-active.
    const uint32_t t = threadIdx.x;
    const uint32_t wave = t >> 6u;
    const uint32_t lane = t & 63u;
    const uint32_t col = blockIdx.x * kTile + (wave & 1u) * 8u + (lane & 7u);
    const uint32_t localRow = blockIdx.y * kTile + (wave >> 1u) * 8u + (lane >> 3u);
    if (col >= params.width || localRow >= params.rows) {
        return;
    }
    const size_t i = size_t(localRow) * params.width + col;
    const float base = heights[i];
    if (IsNoData(params, base)) {
        return;
    }
    const uint32_t row = params.first_row + localRow;
    const uint32_t side = tiles[1];
    const uint32_t head = 3u + 2u * ((row / side) * tiles[0] + col / side);
    const uint32_t first = tiles[head];
    const uint32_t count = tiles[head + 1];
    const float x = xs[col];
    const float xLow = lows[col];
    const float y = ys[localRow];
    const float yLow = lows[params.width + row];
    float h = base;
    for (uint32_t e = first; e < first + count; ++e) {
        const uint32_t n = entries[e];
        const tg_boulder& b = boulders[n];
        const float dx = (x - b.x) + (xLow - b.x_low);
        const float dy = (y - b.y) + (yLow - b.y_low);
        if (dx * dx + dy * dy >= cull[n]) {
            continue;  // beyond the dome and its band: no relief
        }
        h = tg::ApplyBoulder(b, kernel.width, h, base, dx, dy);
    }
    heights[i] = h;
}

template <class T> T* Put(const T* src, size_t count, std::vector<void*>& owned) {
    const size_t bytes = (count > 0 ? count : 1) * sizeof(T);
    void* p = nullptr;
    Check(hipMalloc(&p, bytes), "hipMalloc");
    owned.push_back(p);
    if (src != nullptr && count > 0) {
        Check(hipMemcpy(p, src, count * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy to device");
    } else {
        Check(hipMemset(p, 0, bytes), "hipMemset");
    }
    return static_cast<T*>(p);
}

}  // namespace

namespace tgi {

void ValidateBoulderBatch(const tg_boulder_kernel& k, const tg_boulder* boulders, size_t n,
                          const uint32_t* tiles, size_t tileWords, const uint32_t* entries,
                          size_t entryCount) {
    if (k.count != n) {
        throw HipError{"the boulder kernel's count differs from the boulders bound"};
    }
    if (!(k.width >= 0.0f) || !std::isfinite(k.width)) {
        throw HipError{"the boulder band width must be finite and not negative"};
    }
    for (size_t i = 0; i < n; ++i) {
        const tg_boulder& b = boulders[i];
        if (!(b.radius > 0.0f) || !std::isfinite(b.radius) || !std::isfinite(b.exposed) ||
            !(b.reach >= 0.0f) || !std::isfinite(b.weight) || b.facet_count > 2) {
            throw HipError{"boulder " + std::to_string(i) + " is malformed"};
        }
    }
    if (tileWords < 3 || tiles[0] == 0 || tiles[1] == 0 || tiles[2] != 0) {
        throw HipError{"boulder tiles: need [tiles across, side, 0 (plane), (first, count)...]"};
    }
    if ((tileWords - 3) % 2 != 0 || ((tileWords - 3) / 2) % tiles[0] != 0) {
        throw HipError{"boulder tiles: the heads do not fill whole rows of tiles"};
    }
    for (size_t h = 3; h + 1 < tileWords; h += 2) {
        if (uint64_t(tiles[h]) + tiles[h + 1] > entryCount) {
            throw HipError{"boulder tiles: a tile's list runs past the entries"};
        }
    }
    for (size_t e = 0; e < entryCount; ++e) {
        if (entries[e] >= n) {
            throw HipError{"boulder entries: an entry names a boulder that is not bound"};
        }
    }
}

void LaunchBoulders(const tg_boulder_set& set, float* heights, const float* xs, const float* ys,
                    const tg_kernel_params& params, hipStream_t stream) {
    if (params.width == 0 || params.rows == 0) {
        return;
    }
    if (set.tileSide != kTile) {
        throw HipError{"the boulder kernel bins 16 x 16-cell tiles"};
    }
    if (uint64_t(set.tilesAcross) * kTile < params.width ||
        uint64_t(set.tileRows) * kTile < uint64_t(params.first_row) + params.rows ||
        set.lowCount < size_t(params.width) + params.first_row + params.rows) {
        throw HipError{"the boulder tiles or lows do not cover the grid"};
    }
    const dim3 block(kTile * kTile, 1, 1);
    const dim3 grid((params.width + kTile - 1) / kTile, (params.rows + kTile - 1) / kTile, 1);
    hipLaunchKernelGGL(BoulderStrip, grid, block, 0, stream, heights, xs, ys, set.kernel,
                       set.boulders, set.tiles, set.entries, set.lows, set.cull, params);
    Check(hipGetLastError(), "launching the boulder kernel");
}

}  // namespace tgi

extern "C" {

int tg_hip_boulder_set_create(tg_dev* dev, const tg_boulder_kernel* kernel,
                              const tg_boulder* boulders, size_t count, const uint32_t* tiles,
                              size_t tileWords, const uint32_t* entries, size_t entryCount,
                              const float* lows, size_t lowCount, tg_boulder_set** out,
                              char* err, size_t errLen) {
    if (dev == nullptr || kernel == nullptr || out == nullptr || tiles == nullptr ||
        (count > 0 && boulders == nullptr) || (entryCount > 0 && entries == nullptr) ||
        (lowCount > 0 && lows == nullptr)) {
        SetErr(err, errLen, "tg_hip_boulder_set_create: null argument");
        return TG_ERR_ARGUMENT;
    }
    *out = nullptr;
    std::vector<void*> owned;
    const int code = Guard(err, errLen, [&] {
        ValidateBoulderBatch(*kernel, boulders, count, tiles, tileWords, entries, entryCount);
        Check(hipSetDevice(dev->ordinal), "hipSetDevice");
        auto set = std::make_unique<tg_boulder_set>();
        set->dev = dev;
        set->kernel = *kernel;
        set->boulders = Put(boulders, count, owned);
        set->tiles = Put(tiles, tileWords, owned);
        set->entries = Put(entries, entryCount, owned);
        set->lows = Put(lows, lowCount, owned);
        // Relief is 0 once (d - width / 2) / radius >= 1, i.e. beyond the reach; a relative
        // slack far above FP32 rounding keeps every cell the dome touches.
        std::vector<float> cull(count > 0 ? count : 1, 0.0f);
        for (size_t i = 0; i < count; ++i) {
            const double r = double(boulders[i].reach) * (1.0 + 1e-4);
            cull[i] = float(r * r);
        }
        set->cull = Put(cull.data(), cull.size(), owned);
        set->count = count;
        set->lowCount = lowCount;
        set->tileSide = tiles[1];
        set->tilesAcross = tiles[0];
        set->tileRows = uint32_t((tileWords - 3) / 2 / tiles[0]);
        owned.clear();
        *out = set.release();
    });
    for (void* p : owned) {
        (void)hipFree(p);
    }
    return code;
}

int tg_hip_boulder_set_free(tg_boulder_set* set) {
    if (set == nullptr) {
        return TG_OK;
    }
    (void)hipSetDevice(set->dev->ordinal);
    for (void* p : {static_cast<void*>(set->boulders), static_cast<void*>(set->tiles),
                    static_cast<void*>(set->entries), static_cast<void*>(set->lows),
                    static_cast<void*>(set->cull)}) {
        (void)hipFree(p);
    }
    delete set;
    return TG_OK;
}

int tg_hip_boulder_launch(tg_dev* dev, const tg_boulder_set* set, float* heights, const float* xs,
                          const float* ys, const tg_kernel_params* params, void* stream,
                          char* err, size_t errLen) {
    if (dev == nullptr || set == nullptr || params == nullptr ||
        (size_t(params->width) * params->rows > 0 &&
         (heights == nullptr || xs == nullptr || ys == nullptr))) {
        SetErr(err, errLen, "tg_hip_boulder_launch: null argument");
        return TG_ERR_ARGUMENT;
    }
    return Guard(err, errLen, [&] {
        Check(hipSetDevice(dev->ordinal), "hipSetDevice");
        LaunchBoulders(*set, heights, xs, ys, *params, static_cast<hipStream_t>(stream));
    });
}

}  // extern "C"
