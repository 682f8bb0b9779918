// SPDX-License-Identifier: BSD-3-Clause
//
// ORACLE — test infrastructure only (see noise_oracle.cpp). The crater kernel of the reference
// (include/terraingen/layers/kernels/craters.h:150-1370) on the CPU, in two precisions:
// or_* in FP32, the reference-equivalent path, and or_*_f64, the FP64 oracle (crater_oracle_impl.h).
// The grid functions follow the reference's CPU shaping (src/layers/craters/shaping/apply.cpp:
// 199-222): each cell runs its 16x16 tile's crater list, in list order.

#include "noise_core.h"
#include "tg_layouts.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#define ORC_R float
#define ORC_NS c32
#include "crater_oracle_impl.h"
#undef ORC_NS
#undef ORC_R

#define ORC_R double
#define ORC_NS c64
#include "crater_oracle_impl.h"
#undef ORC_NS
#undef ORC_R

namespace {

template <class F> void Rows(std::size_t rows, unsigned threads, F task) {
    if (threads == 0) {
        threads = std::max(1u, std::thread::hardware_concurrency());
    }
    threads = unsigned(std::min<std::size_t>(threads, std::max<std::size_t>(rows, 1)));
    std::vector<std::thread> pool;
    for (unsigned t = 1; t < threads; ++t) {
        pool.emplace_back([&, t] {
            for (std::size_t r = t; r < rows; r += threads) {
                task(r);
            }
        });
    }
    for (std::size_t r = 0; r < rows; r += threads) {
        task(r);
    }
    for (std::thread& th : pool) {
        th.join();
    }
}

bool IsNoData(const tg_kernel_params& p, float h) {
    if (p.no_data == TG_NODATA_NAN) {
        return std::isnan(h);
    }
    return p.no_data != TG_NODATA_NONE && h == p.no_data_value;
}

// CraterTileHead (craters.h:1341-1344).
std::uint32_t Head(const std::uint32_t* tiles, std::uint32_t col, std::uint32_t row) {
    const std::uint32_t side = tiles[1];
    return 3u + 2u * ((row / side) * tiles[0] + col / side);
}

}  // namespace

extern "C" {

float or_apply_crater(const tg_crater_shape* k, const float* data, const tg_crater_record* c,
                      float height, float dx, float dy) {
    const c32::Shp shape = c32::Widen(*k);
    return c32::ApplyCrater(c32::CraterShape(shape), data, c32::Widen(*c), height, dx, dy);
}

double or_apply_crater_f64(const tg_crater_shape* k, const float* data, const tg_crater_record* c,
                           double height, double dx, double dy) {
    const c64::Shp shape = c64::Widen(*k);
    return c64::ApplyCrater(c64::CraterShape(shape), data, c64::Widen(*c), height, dx, dy);
}

// The crater kernel over a plane grid strip with the binding of craters.slang: xs and ys from the
// origin, `lows` what they leave (width columns then every grid row), tiles and entries as
// FlattenKernelLists lays them out.
void or_shape_grid(const tg_crater_shape* k, const tg_crater_record* records, const float* data,
                   const std::uint32_t* tiles, const std::uint32_t* entries, const float* xs,
                   const float* ys, const float* lows, std::size_t width, std::size_t rows,
                   const tg_kernel_params* params, float* heights, unsigned threads) {
    const c32::Shp shape = c32::Widen(*k);
    const c32::CraterShape view(shape);
    Rows(rows, threads, [&](std::size_t r) {
        const std::uint32_t row = params->first_row + std::uint32_t(r);
        for (std::size_t col = 0; col < width; ++col) {
            float& cell = heights[r * width + col];
            if (IsNoData(*params, cell)) {
                continue;
            }
            const std::uint32_t head = Head(tiles, std::uint32_t(col), row);
            const std::uint32_t first = tiles[head], count = tiles[head + 1];
            float h = cell;
            for (std::uint32_t e = first; e < first + count; ++e) {
                const c32::Rec c = c32::Widen(records[entries[e]]);
                const float dx = c32::SplitDifference(xs[col], lows[col], c.x, c.x_low);
                const float dy = c32::SplitDifference(ys[r], lows[width + row], c.y, c.y_low);
                h = c32::ApplyCrater(view, data, c, h, dx, dy);
            }
            cell = h;
        }
    });
}

// The FP64 oracle of or_shape_grid: cell centres given exactly in FP64 (meters, the same frame as
// the records' split centres), heights accumulated in FP64.
void or_shape_grid_f64(const tg_crater_shape* k, const tg_crater_record* records,
                       const float* data, const std::uint32_t* tiles,
                       const std::uint32_t* entries, const double* cx, const double* cy,
                       std::size_t width, std::size_t rows, std::uint32_t firstRow,
                       double* heights, unsigned threads) {
    const c64::Shp shape = c64::Widen(*k);
    const c64::CraterShape view(shape);
    Rows(rows, threads, [&](std::size_t r) {
        const std::uint32_t row = firstRow + std::uint32_t(r);
        for (std::size_t col = 0; col < width; ++col) {
            const std::uint32_t head = Head(tiles, std::uint32_t(col), row);
            const std::uint32_t first = tiles[head], count = tiles[head + 1];
            double h = heights[r * width + col];
            for (std::uint32_t e = first; e < first + count; ++e) {
                const c64::Rec c = c64::Widen(records[entries[e]]);
                const double dx = cx[col] - (c.x + c.x_low);
                const double dy = cy[r] - (c.y + c.y_low);
                h = c64::ApplyCrater(view, data, c, h, dx, dy);
            }
            heights[r * width + col] = h;
        }
    });
}

// The crater kernel on a body: each cell's split position and each crater's frame
// (ShapeCratersOnBody, craters.h:1359-1370).
void or_shape_grid_body(const tg_crater_shape* k, const tg_crater_record* records,
                        const float* data, const std::uint32_t* tiles,
                        const std::uint32_t* entries, const tg_crater_frame* frames,
                        const tg_split_point* points, std::size_t width, std::size_t rows,
                        const tg_kernel_params* params, float* heights, unsigned threads) {
    const c32::Shp shape = c32::Widen(*k);
    const c32::CraterShape view(shape);
    Rows(rows, threads, [&](std::size_t r) {
        const std::uint32_t row = params->first_row + std::uint32_t(r);
        for (std::size_t col = 0; col < width; ++col) {
            const std::size_t i = r * width + col;
            if (IsNoData(*params, heights[i])) {
                continue;
            }
            const std::uint32_t head = Head(tiles, std::uint32_t(col), row);
            const std::uint32_t first = tiles[head], count = tiles[head + 1];
            const tg_split_point& s = points[i];
            const c32::Pt p{s.x, s.x_low, s.y, s.y_low, s.z, s.z_low};
            float h = heights[i];
            for (std::uint32_t e = first; e < first + count; ++e) {
                const std::uint32_t n = entries[e];
                const c32::Offset at = c32::BodySplitOffset(c32::Widen(frames[n]), p);
                h = c32::ApplyCrater(view, data, c32::Widen(records[n]), h, at.dx, at.dy);
            }
            heights[i] = h;
        }
    });
}

// The FP64 oracle of or_shape_grid_body: each cell's body-fixed position (x, y, z per cell) and
// each crater's frame (centre x, y, z, east x, y, z, north x, y, z, sphere radius per crater) given
// exactly in FP64, heights accumulated in FP64.
void or_shape_grid_body_f64(const tg_crater_shape* k, const tg_crater_record* records,
                            const float* data, const std::uint32_t* tiles,
                            const std::uint32_t* entries, const double* frames,
                            const double* positions, std::size_t width, std::size_t rows,
                            std::uint32_t firstRow, double* heights, unsigned threads) {
    const c64::Shp shape = c64::Widen(*k);
    const c64::CraterShape view(shape);
    Rows(rows, threads, [&](std::size_t r) {
        const std::uint32_t row = firstRow + std::uint32_t(r);
        for (std::size_t col = 0; col < width; ++col) {
            const std::size_t i = r * width + col;
            const std::uint32_t head = Head(tiles, std::uint32_t(col), row);
            const std::uint32_t first = tiles[head], count = tiles[head + 1];
            const double* p = positions + 3 * i;
            double h = heights[i];
            for (std::uint32_t e = first; e < first + count; ++e) {
                const std::uint32_t n = entries[e];
                const double* f = frames + 10 * std::size_t(n);
                const c64::Frm frame{f[0], f[1], f[2], f[3], f[4], f[5], f[6], f[7], f[8], f[9],
                                     0.0, 0.0, 0.0};
                const c64::Offset at =
                    c64::FrameOffset(frame, p[0] - f[0], p[1] - f[1], p[2] - f[2]);
                h = c64::ApplyCrater(view, data, c64::Widen(records[n]), h, at.dx, at.dy);
            }
            heights[i] = h;
        }
    });
}

// crater_datums.slang over a strip (datum_fit.cpp:306-328): row r of the lattice holds crater
// targets[first_row + r]'s footprint, column p its lattice point p; each adds KernelDatumRelief
// (craters.h:1266-1305), the relief of the craters before the target found through the size
// bands of the crater index.
void or_datum_grid(const tg_crater_datum_kernel* kd, const tg_crater_record* records,
                   const float* data, const float* bandFloats, const std::uint32_t* bandCells,
                   const std::uint32_t* cells, const std::uint32_t* entries,
                   const std::uint32_t* targets, const tg_kernel_params* params, float* heights,
                   unsigned threads) {
    const c32::Shp shape = c32::Widen(kd->shape);
    Rows(params->rows, threads, [&](std::size_t r) {
        const std::uint32_t k = targets[params->first_row + r];
        const c32::Rec target = c32::Widen(records[k]);
        const std::uint32_t side = kd->lattice_side;
        const float fitRadius = kd->fit_radii * 0.5f * target.diameter;
        for (std::uint32_t col = 0; col < params->width; ++col) {
            const float dx = fitRadius * ((2.0f * float(col % side) + 1.0f) / float(side) - 1.0f);
            const float dy = fitRadius * ((2.0f * float(col / side) + 1.0f) / float(side) - 1.0f);
            float relief = 0.0f;
            if (dx * dx + dy * dy < fitRadius * fitRadius) {
                const float x = target.x, xLow = target.x_low + dx;
                const float y = target.y, yLow = target.y_low + dy;
                const float px = x + xLow, py = y + yLow;
                const float slackX = 4.8e-7f * std::fabs(px), slackY = 4.8e-7f * std::fabs(py);
                // KernelCellsNear (craters.h:1256-1264).
                const auto near = [](float p, float size, float slack, float start,
                                     std::int32_t count, std::int32_t& lo, std::int32_t& hi) {
                    lo = std::max(std::int32_t(std::floor((p - size - slack) / size) - start), 0);
                    hi = std::min(std::int32_t(std::floor((p + size + slack) / size) - start),
                                  count - 1);
                };
                for (std::uint32_t b = 0; b < kd->band_count; ++b) {
                    const float size = bandFloats[3 * b];
                    const std::int32_t across = std::int32_t(bandCells[3 * b]);
                    std::int32_t i0, i1, j0, j1;
                    near(px, size, slackX, bandFloats[3 * b + 1], across, i0, i1);
                    near(py, size, slackY, bandFloats[3 * b + 2],
                         std::int32_t(bandCells[3 * b + 1]), j0, j1);
                    for (std::int32_t j = j0; j <= j1; ++j) {
                        for (std::int32_t i = i0; i <= i1; ++i) {
                            const std::uint32_t cell =
                                bandCells[3 * b + 2] + std::uint32_t(j * across + i);
                            const std::uint32_t first = cells[2 * cell];
                            const std::uint32_t end = first + cells[2 * cell + 1];
                            for (std::uint32_t e = first; e < end && entries[e] < k; ++e) {
                                const c32::Rec c = c32::Widen(records[entries[e]]);
                                relief += c32::CraterRelief(
                                    shape, data, c, size,
                                    c32::SplitDifference(x, xLow, c.x, c.x_low),
                                    c32::SplitDifference(y, yLow, c.y, c.y_low));
                            }
                        }
                    }
                }
            }
            float& h = heights[r * params->width + col];
            h = h + relief;
        }
    });
}

// crater_body_datums.slang over a strip (datum_fit.cpp:544-574): each lattice point with data adds
// KernelBodyDatumRelief (craters.h:1313-1326), the relief of its row's listed craters in order,
// each read in its frame within its bound.
void or_body_datum_grid(const tg_crater_body_datum_kernel* kd, const tg_crater_record* records,
                        const float* data, const tg_crater_frame* frames, const float* bounds,
                        const std::uint32_t* lists, const std::uint32_t* entries,
                        const tg_split_point* positions, const tg_kernel_params* params,
                        float* heights, unsigned threads) {
    const c32::Shp shape = c32::Widen(kd->shape);
    Rows(params->rows, threads, [&](std::size_t r) {
        const std::uint32_t row = params->first_row + std::uint32_t(r);
        for (std::uint32_t col = 0; col < params->width; ++col) {
            float& h = heights[r * params->width + col];
            if (IsNoData(*params, h)) {
                continue;
            }
            const tg_split_point& s = positions[r * params->width + col];
            const c32::Pt p{s.x, s.x_low, s.y, s.y_low, s.z, s.z_low};
            float relief = 0.0f;
            for (std::uint32_t e = lists[2 * row]; e < lists[2 * row] + lists[2 * row + 1]; ++e) {
                const std::uint32_t n = entries[e];
                const c32::Offset o = c32::BodySplitOffset(c32::Widen(frames[n]), p);
                relief += c32::CraterRelief(shape, data, c32::Widen(records[n]), bounds[n], o.dx,
                                            o.dy);
            }
            h = h + relief;
        }
    });
}

}  // extern "C"
