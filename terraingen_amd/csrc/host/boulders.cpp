// SPDX-License-Identifier: BSD-3-Clause
//
// Boulder relief preparation for one grid spacing: BoulderLayer::PrepareRelief
// (src/layers/boulders/boulder_relief.cpp:373-423) for the default look. A boulder is drawn
// when its resolution weight at the spacing is above 0 (ResolutionWeight with the 2-4 sample
// fade, src/layers/fade.cpp:48-52) and the relief range holds it (BoulderReliefRange::Holds,
// boulder_options.h:382-385: every boulder that is not a model); the dome is averaged over a
// band one sample wide (BoulderRelief::Dome) and its box reaches half a diameter plus half the
// band. The records carry what the device reads in FP32; the centres are split hi + lo.

#include "common.h"

#include <cmath>

namespace tgh {

namespace {

double ResolutionWeight(double diameter, double spacing) {
    const double x = diameter / spacing;
    if (x <= 2.0) {
        return 0.0;
    }
    if (x >= 4.0) {
        return 1.0;
    }
    const double t = std::clamp((x - 2.0) / (4.0 - 2.0), 0.0, 1.0);
    return 1.0 - (1.0 - t * t * (3.0 - 2.0 * t));
}

}  // namespace

}  // namespace tgh

extern "C" {

int tgx_prepare_boulders(const tgx_boulder* boulders, size_t n, double spacing, int dome,
                         tg_boulder_kernel* kernel, tg_boulder** records, size_t* count,
                         double** footprint, char* err, size_t errLen) {
    return tgh::Guard(err, errLen, [&] {
        if ((n > 0 && boulders == nullptr) || kernel == nullptr || records == nullptr ||
            count == nullptr) {
            throw std::invalid_argument("tgx_prepare_boulders: null argument");
        }
        if (!(spacing > 0.0) || !std::isfinite(spacing)) {
            throw std::invalid_argument("tgx_prepare_boulders: spacing must be positive");
        }
        const double width = dome != 0 ? spacing : 0.0;
        std::vector<tg_boulder> out;
        std::vector<double> foot;
        for (size_t i = 0; i < n; ++i) {
            const tgx_boulder& b = boulders[i];
            if (!(b.diameter > 0.0) || !std::isfinite(b.x) || !std::isfinite(b.y) ||
                !std::isfinite(b.exposed) || b.facet_count > 2) {
                throw std::invalid_argument("boulder " + std::to_string(i) + " is malformed");
            }
            const double weight = tgh::ResolutionWeight(b.diameter, spacing);
            if (b.model != 0 || !(weight > 0.0)) {
                continue;
            }
            tg_boulder r{};
            r.x = static_cast<float>(b.x);
            r.y = static_cast<float>(b.y);
            r.x_low = static_cast<float>(b.x - double(r.x));
            r.y_low = static_cast<float>(b.y - double(r.y));
            r.radius = static_cast<float>(0.5 * b.diameter);
            r.exposed = static_cast<float>(b.exposed);
            const double reach = 0.5 * (b.diameter + width);
            r.reach = static_cast<float>(reach);
            r.weight = static_cast<float>(weight);
            for (uint32_t k = 0; k < b.facet_count; ++k) {
                const double azimuth = b.facets[3 * k], angle = b.facets[3 * k + 1];
                r.facet[4 * k] = static_cast<float>(std::cos(azimuth));
                r.facet[4 * k + 1] = static_cast<float>(std::sin(azimuth));
                r.facet[4 * k + 2] = static_cast<float>(b.facets[3 * k + 2]);
                r.facet[4 * k + 3] = static_cast<float>(std::tan(angle));
            }
            r.facet_count = b.facet_count;
            out.push_back(r);
            foot.insert(foot.end(), {b.x, b.y, reach, b.diameter});
        }
        kernel->width = static_cast<float>(width);
        kernel->count = static_cast<uint32_t>(out.size());
        kernel->pad[0] = kernel->pad[1] = 0;
        *count = out.size();
        *records = tgh::Export(out);
        if (footprint != nullptr) {
            *footprint = tgh::Export(foot);
        }
    });
}

}  // extern "C"
