// SPDX-License-Identifier: BSD-3-Clause
//
// Boulder dome relief on the device, in FP32: BoulderLayer's default look (BoulderRelief::Dome,
// no fillet) as src/layers/boulders/boulder_relief.cpp evaluates it in FP64 on the CPU.
#pragma once

#include "tg_layouts.h"

#include <hip/hip_runtime.h>

namespace tg {

// The area under the unit half-circle from 0 to u, for u in [-1, 1]: the dome's band integral.
__device__ __forceinline__ float DomeArea(float u) {
    return 0.5f * (u * sqrtf(1.0f - u * u) + asinf(u));
}

// DomeHeight (boulder_relief.cpp:44-57) cut by the boulder's facets (Boulder::CutByFacets,
// boulder.h:99-104) at east/north (dx, dy) from its centre: the rock's height above its base,
// never below 0.
__device__ __forceinline__ float BoulderRelief(const tg_boulder& b, float width, float dx,
                                               float dy) {
    const float distance = sqrtf(dx * dx + dy * dy);
    float top;
    if (!(width > 0.0f)) {
        const float r2 = (distance * distance) / (b.radius * b.radius);
        top = r2 < 1.0f ? b.exposed * sqrtf(1.0f - r2) : 0.0f;
    } else {
        const float low = fminf(fmaxf((distance - 0.5f * width) / b.radius, -1.0f), 1.0f);
        const float high = fminf(fmaxf((distance + 0.5f * width) / b.radius, -1.0f), 1.0f);
        if (!(high > low)) {
            return 0.0f;
        }
        top = b.exposed * (DomeArea(high) - DomeArea(low)) * b.radius / width;
    }
    for (uint32_t k = 0; k < b.facet_count; ++k) {
        const float* f = b.facet + 4u * k;
        const float along = dx * f[0] + dy * f[1];
        top = fminf(top, (f[2] - along) * f[3]);
    }
    return fmaxf(top, 0.0f);
}

// BoulderLayer::AddRelief without a fillet (boulder_relief.cpp:244-262): the cell's height `h` and
// its height `base` before the layer's boulders, the relief faded by `weight`.
__device__ __forceinline__ float ApplyBoulder(const tg_boulder& b, float width, float weight,
                                              float h, float base, float dx, float dy) {
    const float relief = BoulderRelief(b, width, dx, dy);
    return relief > 0.0f ? fmaxf(h, base + weight * relief) : h;
}

// The same at the weight the record was prepared with.
__device__ __forceinline__ float ApplyBoulder(const tg_boulder& b, float width, float h,
                                              float base, float dx, float dy) {
    return ApplyBoulder(b, width, b.weight, h, base, dx, dy);
}

}  // namespace tg
