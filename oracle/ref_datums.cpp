// SPDX-License-Identifier: BSD-3-Clause
//
// ORACLE — test infrastructure only (see noise_oracle.cpp). The reference's datum relief, built
// into oracle/_ref/libtgref.so from its sources: KernelDatumRelief and KernelBodyDatumRelief
// (include/terraingen/layers/kernels/craters.h:1266-1326) at every lattice point of a strip, as
// the reference's CPU path runs the first (AddDatumRelief, src/layers/craters/shaping/
// datum_fit.cpp:316-327) and its crater_body_datums.slang runs the second, on this repository's
// layouts, which are the reference's (tg_layouts.h). Nothing of the reference is copied here: this
// file calls its functions.

#include "terraingen/layers/kernels/craters.h"

#include "tg_layouts.h"

#include <cmath>
#include <cstdint>
#include <cstring>

using namespace terraingen;

extern "C" {

void ref_datum_grid(const void* kernel, const void* records, const float* data,
                    const float* bandFloats, const std::uint32_t* bandCells,
                    const std::uint32_t* cells, const std::uint32_t* entries,
                    const std::uint32_t* targets, const tg_kernel_params* params,
                    float* heights) {
    static_assert(sizeof(CraterDatumKernel) == sizeof(tg_crater_datum_kernel), "layout");
    CraterDatumKernel k;
    std::memcpy(&k, kernel, sizeof k);
    const auto* recs = static_cast<const CraterRecord*>(records);
    for (std::uint32_t r = 0; r < params->rows; ++r) {
        for (std::uint32_t col = 0; col < params->width; ++col) {
            float& h = heights[std::size_t(r) * params->width + col];
            h = h + KernelDatumRelief(k, CraterData{data}, recs, bandFloats, bandCells, cells,
                                      entries, targets[params->first_row + r], col);
        }
    }
}

void ref_body_datum_grid(const void* kernel, const void* records, const float* data,
                         const void* frames, const float* bounds, const std::uint32_t* lists,
                         const std::uint32_t* entries, const void* positions,
                         const tg_kernel_params* params, float* heights) {
    static_assert(sizeof(CraterBodyDatumKernel) == sizeof(tg_crater_body_datum_kernel), "layout");
    static_assert(sizeof(SplitPoint) == sizeof(tg_split_point), "layout");
    CraterBodyDatumKernel k;
    std::memcpy(&k, kernel, sizeof k);
    const auto* recs = static_cast<const CraterRecord*>(records);
    const auto* frs = static_cast<const CraterFrame*>(frames);
    const auto* pts = static_cast<const SplitPoint*>(positions);
    for (std::uint32_t r = 0; r < params->rows; ++r) {
        const std::uint32_t row = params->first_row + r;
        for (std::uint32_t col = 0; col < params->width; ++col) {
            const std::size_t at = std::size_t(r) * params->width + col;
            float& h = heights[at];
            const bool noData = params->no_data == TG_NODATA_NAN
                                    ? std::isnan(h)
                                    : params->no_data != TG_NODATA_NONE && h == params->no_data_value;
            if (noData) {
                continue;
            }
            h = h + KernelBodyDatumRelief(k.shape, CraterData{data}, recs, frs, bounds, entries,
                                          lists[2 * row], lists[2 * row + 1], pts[at]);
        }
    }
}

}  // extern "C"
