// SPDX-License-Identifier: BSD-3-Clause
//
// libtg_hip.so: the gfx950 kernels of the hip backend and their C ABI (include/tg_hip.h).
// Built with -ffp-contract=off, as the reference is (its CMakeLists.txt:32-38), so FP32 results
// depend only on the order of operations written here.

#include "boulder_device.h"
#include "tg_hip.h"

#include "crater_device.h"
#include "noise_device.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "internal.h"

using namespace tgi;

// The library's static scratch (tg_hip_batch_run_static): each device holds its own copy until the
// process ends. Host code never passes this name to a kernel or a copy: it stands in for the
// device address only on the device side. tg_hip_open looks up each device's address of it with
// hipGetSymbolAddress and keeps that in tg_dev::scratch, which is what host code uses.
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

namespace {


// dst[i] = value for i < count, grid-stride: a flat source's fill, ordered on a stream (and safe
// inside graph capture, where a memset node is not).
__global__ void FillValue(float* __restrict__ dst, float value, size_t count) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count;
         i += size_t(gridDim.x) * blockDim.x) {
        dst[i] = value;
    }
}

// Cells per block: 64 columns (one wave's lanes along a row, so stores coalesce) by 4 rows.
constexpr uint32_t kNoiseBlockX = 64;
constexpr uint32_t kNoiseBlockY = 4;

// noise_layer.slang's main(): heights[i] = height + EvaluateFractals(xs[col], ys[row], stack).
__global__ void __launch_bounds__(kNoiseBlockX * kNoiseBlockY)
    NoiseLayerStrip(float* __restrict__ heights, const float* __restrict__ xs,
                    const float* __restrict__ ys, const tg::NoisePlan* __restrict__ plan,
                    tg_kernel_params params) {
    const uint32_t col = blockIdx.x * kNoiseBlockX + threadIdx.x;
    const uint32_t row = blockIdx.y * kNoiseBlockY + threadIdx.y;
    if (col >= params.width || row >= params.rows) {
        return;
    }
    const size_t i = size_t(row) * params.width + col;
    const float height = heights[i];
    if (IsNoData(params, height)) {
        return;
    }
    heights[i] = height + tg::EvaluateStack(*plan, xs[col], ys[row]);
}

// SyntheticSource::NoiseAt: heights[i] = elevation + EvaluateFractals(xs[col], ys[row], stack).
__global__ void __launch_bounds__(kNoiseBlockX * kNoiseBlockY)
    NoiseFillStrip(float* __restrict__ heights, const float* __restrict__ xs,
                   const float* __restrict__ ys, const tg::NoisePlan* __restrict__ plan,
                   float elevation, tg_kernel_params params) {
    const uint32_t col = blockIdx.x * kNoiseBlockX + threadIdx.x;
    const uint32_t row = blockIdx.y * kNoiseBlockY + threadIdx.y;
    if (col >= params.width || row >= params.rows) {
        return;
    }
    heights[size_t(row) * params.width + col] =
        elevation + tg::EvaluateStack(*plan, xs[col], ys[row]);
}

// The noise on a body (SyntheticSource::NoiseAt / NoiseLayer::ApplySamples on a body,
// src/noise/surface_noise.cpp:9-20): heights[i] += EvaluateFractals3 at positions[i], or with
// `fill`, heights[i] = elevation + it.
__global__ void __launch_bounds__(256)
    NoiseBodyStrip(float* __restrict__ heights, const double* __restrict__ positions,
                   const tg::NoisePlan* __restrict__ plan, tg_kernel_params params, int fill,
                   float elevation) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= size_t(params.width) * params.rows) {
        return;
    }
    const float height = fill ? elevation : heights[i];
    if (!fill && IsNoData(params, height)) {
        return;
    }
    const double* p = positions + 3 * i;
    heights[i] = height + tg::EvaluateStack3(*plan, p[0], p[1], p[2]);
}

void LaunchNoise(const tg::NoisePlan* plan, float* heights, const float* xs, const float* ys,
                 const tg_kernel_params& params, hipStream_t stream, bool fill = false,
                 float elevation = 0.0f) {
    if (params.width == 0 || params.rows == 0) {
        return;
    }
    const dim3 block(kNoiseBlockX, kNoiseBlockY, 1);
    const dim3 grid((params.width + kNoiseBlockX - 1) / kNoiseBlockX,
                    (params.rows + kNoiseBlockY - 1) / kNoiseBlockY, 1);
    if (grid.y > 65535u * 64u) {
        throw HipError{"strip has too many rows for one launch"};
    }
    if (fill) {
        hipLaunchKernelGGL(NoiseFillStrip, grid, block, 0, stream, heights, xs, ys, plan,
                           elevation, params);
    } else {
        hipLaunchKernelGGL(NoiseLayerStrip, grid, block, 0, stream, heights, xs, ys, plan, params);
    }
    Check(hipGetLastError(), "launching the noise kernel");
}

// craters.slang's main(): each cell runs the crater list of its tile.
constexpr uint32_t kCraterTile = 16;

// The boulder layer a crater pass may carry (a plane boulder set's arrays, null when none): drawn
// over each cell after its craters, as the boulder kernel draws it over the craters' heights, so
// the grid is read and written once for both layers.
struct BoulderPass {
    tg_boulder_kernel kernel;
    const tg_boulder* __restrict__ boulders;
    const uint32_t* __restrict__ tiles;
    const uint32_t* __restrict__ entries;
    const float* __restrict__ cull;
};

// kAllFast: every crater of the set takes the hoisted MPD path (see BatchTiles).
// kBoulders: a boulder layer follows the craters (BoulderPass).
// kPrecise: some crater of the set takes that path in FP64 (a set without km-scale craters runs
// a variant without the FP64 code, whose registers it would otherwise hold).
template <bool kAllFast, bool kBoulders = false, bool kPrecise = false>
__global__ void __launch_bounds__(kCraterTile * kCraterTile)
    CraterStrip(float* __restrict__ heights, const float* __restrict__ xs,
                const float* __restrict__ ys, const tg_crater_shape* __restrict__ kernel,
                const tg_crater_record* __restrict__ records, const float* __restrict__ data,
                const uint32_t* __restrict__ tiles, const uint32_t* __restrict__ entries,
                const tg_crater_frame* __restrict__ frames,
                const tg_split_point* __restrict__ points, const float* __restrict__ lows,
                const float* __restrict__ cull, const float* __restrict__ harm,
                const float4* __restrict__ outline, const tg::MpdTerms* __restrict__ terms,
                const tg::PreciseTerms* __restrict__ precise, tg_kernel_params params,
                BoulderPass rocks) {
    // Each 64-lane wave covers an 8 x 8 square of the 16 x 16 tile, so a disc-shaped reach
    // leaves fewer waves partly active than rows of the tile would.
    const uint32_t t = threadIdx.x;
    const uint32_t wave = t >> 6u;
    const uint32_t lane = t & 63u;
    const uint32_t col = blockIdx.x * kCraterTile + (wave & 1u) * 8u + (lane & 7u);
    const uint32_t localRow = blockIdx.y * kCraterTile + (wave >> 1u) * 8u + (lane >> 3u);
    if (col >= params.width || localRow >= params.rows) {
        return;
    }
    const size_t i = size_t(localRow) * params.width + col;
    const float height = heights[i];
    if (IsNoData(params, height)) {
        return;
    }
    const uint32_t row = params.first_row + localRow;
    const uint32_t side = tiles[1];
    const uint32_t head = 3u + 2u * ((row / side) * tiles[0] + col / side);
    const uint32_t first = tiles[head];
    const uint32_t count = tiles[head + 1];
    const tg::CraterShape shape(*kernel);
    tg::PreciseSum cell{height};
    if (tiles[2] == 0u) {
        tg_split_point p;
        p.x = xs[col];
        p.x_low = lows[col];
        p.y = ys[localRow];
        p.y_low = lows[params.width + row];
        p.z = 0.0f;
        p.z_low = 0.0f;
        for (uint32_t e = first; e < first + count; ++e) {
            const uint32_t n = entries[e];
            const tg_crater_record& c = records[n];
            if constexpr (kPrecise) {
                if (terms[n].fast == tg::kFastPrecise) {
                    const tg::Offset64 o =
                        tg::PlaneOffset64(precise[n], p.x, p.x_low, p.y, p.y_low);
                    if (o.distance * o.distance < double(cull[n])) {
                        cell.pending += double(c.resolution) *
                                        tg::MpdPrecise(precise[n], c.ellipticity,
                                                       c.first_harmonic, c.harmonic_count,
                                                       harm + c.harmonic_offset, o);
                    }
                    continue;
                }
            }
            const float dx = tg::SplitDifference(p.x, p.x_low, c.x, c.x_low);
            const float dy = tg::SplitDifference(p.y, p.y_low, c.y, c.y_low);
            if (dx * dx + dy * dy >= cull[n]) {
                continue;  // beyond any reach its outline allows: ApplyCrater returns h as is
            }
            if (kAllFast || terms[n].fast != 0u) {
                cell.h = tg::ApplyCraterMpd(c, terms[n], cell.h, dx, dy, harm, outline[n],
                                            c.resolution);
            } else if constexpr (!kAllFast) {
                cell.h = tg::ApplyCrater(shape, data, c, cell.Folded(), dx, dy, harm, outline[n]);
            }
        }
    } else if constexpr (!kAllFast) {
        // On a body the offset is the cell's in the crater's frame, and the rest is the plane's.
        const tg_split_point p = points[i];
        for (uint32_t e = first; e < first + count; ++e) {
            const uint32_t n = entries[e];
            const tg_crater_record& c = records[n];
            if constexpr (kPrecise) {
                if (terms[n].fast == tg::kFastPrecise) {
                    const tg::Offset64 o = tg::BodyOffset64(precise[n], p);
                    if (o.distance * o.distance < double(cull[n])) {
                        cell.pending += double(c.resolution) *
                                        tg::MpdPrecise(precise[n], c.ellipticity,
                                                       c.first_harmonic, c.harmonic_count,
                                                       harm + c.harmonic_offset, o);
                    }
                    continue;
                }
            }
            const tg::Offset at = tg::BodySplitOffset(frames[n], p);
            if (at.dx * at.dx + at.dy * at.dy >= cull[n]) {
                continue;
            }
            if (terms[n].fast != 0u) {
                cell.h = tg::ApplyCraterMpd(c, terms[n], cell.h, at.dx, at.dy, harm, outline[n],
                                            c.resolution);
            } else {
                cell.h = tg::ApplyCrater(shape, data, c, cell.Folded(), at.dx, at.dy, harm,
                                         outline[n]);
            }
        }
    }
    float out = cell.Folded();
    if constexpr (kBoulders) {
        // BoulderStrip's loop over the cell's boulder tile, the craters' height as the base.
        const uint32_t bside = rocks.tiles[1];
        const uint32_t bhead = 3u + 2u * ((row / bside) * rocks.tiles[0] + col / bside);
        const uint32_t bfirst = rocks.tiles[bhead];
        const uint32_t bend = bfirst + rocks.tiles[bhead + 1];
        const float base = out;
        const float x = xs[col], xLow = lows[col];
        const float y = ys[localRow], yLow = lows[params.width + row];
        for (uint32_t e = bfirst; e < bend; ++e) {
            const uint32_t n = rocks.entries[e];
            const tg_boulder& b = rocks.boulders[n];
            const float dx = (x - b.x) + (xLow - b.x_low);
            const float dy = (y - b.y) + (yLow - b.y_low);
            if (dx * dx + dy * dy >= rocks.cull[n]) {
                continue;
            }
            out = tg::ApplyBoulder(b, rocks.kernel.width, out, base, dx, dy);
        }
    }
    heights[i] = out;
}

// CraterRelief (craters.h:1225-1238): the relief crater `c` adds at (dx, dy) from its centre
// without its datum, as datums are fitted to it: its resolution times the bare profile and the
// central peak, within the profile's reach; nothing beyond `bound` meters.
__device__ __forceinline__ float CraterRelief(const tg_crater_shape& k, const float* data,
                                              const tg_crater_record& c, float bound, float dx,
                                              float dy) {
    const float squared = dx * dx + dy * dy;
    if (!(squared < bound * bound)) {
        return 0.0f;
    }
    const float distance = tg::ShiftedDistance(data, c, dx, dy);
    if (!(distance < k.profile_reach_radii * 0.5f * c.diameter)) {
        return 0.0f;
    }
    const tg::CraterView view{
        c,
        {c.diameter, c.fresh_depth, c.fresh_rim, c.current_depth, c.current_rim},
        {c.diameter, c.fresh_depth, c.fresh_rim, c.fresh_depth, c.fresh_rim}};
    return c.resolution *
           (tg::UpslopeHeight(c, tg::ProfileHeight(k, data, distance, c.diameter, view), dx, dy) +
            tg::PeakRelief(k, data, c, distance, dx, dy));
}

// KernelCellsNear (craters.h:1256-1264): the cells of a band's row or column within one cell (and
// the float slack of a cell lookup) of position `p`, clamped to the band.
__device__ __forceinline__ int2 CellsNear(float p, float size, float slack, float start,
                                          int32_t count) {
    const int32_t first = int32_t(floorf((p - size - slack) / size) - start);
    const int32_t last = int32_t(floorf((p + size + slack) / size) - start);
    return int2{first > 0 ? first : 0, last < count - 1 ? last : count - 1};
}

// crater_datums.slang's main(): a row per fitted crater (`targets`), a column per lattice point of
// its datum footprint; each lattice point adds KernelDatumRelief (craters.h:1266-1305), the relief
// of the craters applied before its crater, found in the size bands of the crater index. One wave
// per row: a row's points read the same cells and entries, so the walk is uniform across the wave.
__global__ __launch_bounds__(64) void CraterDatumStrip(
    float* __restrict__ heights, const tg_crater_datum_kernel* __restrict__ kernel,
    const tg_crater_record* __restrict__ records, const float* __restrict__ data,
    const float* __restrict__ bandFloats, const uint32_t* __restrict__ bandCells,
    const uint32_t* __restrict__ cells, const uint32_t* __restrict__ entries,
    const uint32_t* __restrict__ targets, tg_kernel_params params) {
    const uint32_t col = blockIdx.y * 64u + threadIdx.x;
    const uint32_t localRow = blockIdx.x;
    if (col >= params.width || localRow >= params.rows) {
        return;
    }
    const size_t at = size_t(localRow) * params.width + col;
    const tg_crater_datum_kernel& kd = *kernel;
    const uint32_t k = targets[params.first_row + localRow];
    const tg_crater_record& target = records[k];
    const uint32_t side = kd.lattice_side;
    const float fitRadius = kd.fit_radii * 0.5f * target.diameter;
    // KernelDatumOffset (craters.h:1241-1243) of the point's column and row in the lattice.
    const float dx = fitRadius * ((2.0f * float(col % side) + 1.0f) / float(side) - 1.0f);
    const float dy = fitRadius * ((2.0f * float(col / side) + 1.0f) / float(side) - 1.0f);
    float relief = 0.0f;
    if (dx * dx + dy * dy < fitRadius * fitRadius) {
        const float x = target.x, xLow = target.x_low + dx;
        const float y = target.y, yLow = target.y_low + dy;
        const float px = x + xLow;
        const float py = y + yLow;
        const float slackX = 4.8e-7f * fabsf(px);  // kDatumCellSlack (craters.h:1246)
        const float slackY = 4.8e-7f * fabsf(py);
        for (uint32_t b = 0u; b < kd.band_count; ++b) {
            const float size = bandFloats[3u * b];
            const int32_t across = int32_t(bandCells[3u * b]);
            const int2 columns = CellsNear(px, size, slackX, bandFloats[3u * b + 1u], across);
            const int2 rows = CellsNear(py, size, slackY, bandFloats[3u * b + 2u],
                                        int32_t(bandCells[3u * b + 1u]));
            for (int32_t j = rows.x; j <= rows.y; ++j) {
                for (int32_t i = columns.x; i <= columns.y; ++i) {
                    const uint32_t cell = bandCells[3u * b + 2u] + uint32_t(j * across + i);
                    const uint32_t first = cells[2u * cell];
                    const uint32_t end = first + cells[2u * cell + 1u];
                    for (uint32_t e = first; e < end; ++e) {
                        const uint32_t n = entries[e];
                        if (n >= k) {
                            break;
                        }
                        const tg_crater_record& c = records[n];
                        relief += CraterRelief(kd.shape, data, c, size,
                                               tg::SplitDifference(x, xLow, c.x, c.x_low),
                                               tg::SplitDifference(y, yLow, c.y, c.y_low));
                    }
                }
            }
        }
    }
    heights[at] = heights[at] + relief;
}

// crater_body_datums.slang's main(): at each lattice point of a fitted crater's footprint on a
// body (a row per crater, a column per point), KernelBodyDatumRelief (craters.h:1313-1326): the
// relief of the row's listed craters, in list order, each read in its frame within its bound.
// Points with no data are left as they are.
__global__ __launch_bounds__(64) void CraterBodyDatumStrip(
    float* __restrict__ heights, const tg_crater_body_datum_kernel* __restrict__ kernel,
    const tg_crater_record* __restrict__ records, const float* __restrict__ data,
    const tg_crater_frame* __restrict__ frames, const float* __restrict__ bounds,
    const uint32_t* __restrict__ lists, const uint32_t* __restrict__ entries,
    const tg_split_point* __restrict__ positions, tg_kernel_params params) {
    const uint32_t col = blockIdx.y * 64u + threadIdx.x;
    const uint32_t localRow = blockIdx.x;
    if (col >= params.width || localRow >= params.rows) {
        return;
    }
    const size_t at = size_t(localRow) * params.width + col;
    const float height = heights[at];
    if (IsNoData(params, height)) {
        return;
    }
    const uint32_t row = params.first_row + localRow;
    const uint32_t first = lists[2u * row];
    const uint32_t count = lists[2u * row + 1u];
    const tg_split_point p = positions[at];
    float relief = 0.0f;
    for (uint32_t e = first; e < first + count; ++e) {
        const uint32_t n = entries[e];
        const tg::Offset o = tg::BodySplitOffset(frames[n], p);
        relief += CraterRelief(kernel->shape, data, records[n], bounds[n], o.dx, o.dy);
    }
    heights[at] = height + relief;
}

void LaunchCraters(const tg_crater_set& set, float* heights, const float* xs, const float* ys,
                   const tg_kernel_params& params, hipStream_t stream,
                   const tg_boulder_set* rocks = nullptr) {
    if (params.width == 0 || params.rows == 0) {
        return;
    }
    const dim3 block(kCraterTile * kCraterTile, 1, 1);
    const dim3 grid((params.width + kCraterTile - 1) / kCraterTile,
                    (params.rows + kCraterTile - 1) / kCraterTile, 1);
    BoulderPass pass{};
    if (rocks != nullptr) {
        pass = BoulderPass{rocks->kernel, rocks->boulders, rocks->tiles, rocks->entries,
                           rocks->cull};
    }
    // The compact kernel for a plane set whose craters all take the hoisted path; the FP64
    // path's code only for a set with km-scale craters.
    const bool fast = set.allFast && !set.onBody;
    using Strip = decltype(&CraterStrip<true>);
    const Strip variants[2][2][2] = {
        {{CraterStrip<false, false, false>, CraterStrip<false, false, true>},
         {CraterStrip<false, true, false>, CraterStrip<false, true, true>}},
        {{CraterStrip<true, false, false>, CraterStrip<true, false, true>},
         {CraterStrip<true, true, false>, CraterStrip<true, true, true>}}};
    Strip fn = variants[fast ? 1 : 0][rocks != nullptr ? 1 : 0][set.anyPrecise ? 1 : 0];
    hipLaunchKernelGGL(fn, grid, block, 0, stream, heights, xs, ys, set.kernel,
                       set.records, set.data, set.tiles, set.entries, set.frames, set.points,
                       set.lows, set.cull, set.harm, set.outline, set.terms, set.precise,
                       params, pass);
    Check(hipGetLastError(), "launching the crater kernel");
}

int CreateCraterSet(tg_dev* dev, const tg_crater_shape* kernel, const tg_crater_record* records,
                    size_t recordCount, const float* data, size_t dataCount,
                    const uint32_t* tiles, size_t tileWords, const uint32_t* entries,
                    size_t entryCount, const tg_crater_frame* frames, size_t frameCount,
                    const tg_split_point* points, size_t pointCount, const float* lows,
                    size_t lowCount, tg_crater_set** out, char* err, size_t errLen);
void CheckLaunch(const tg_crater_set& set, const tg_kernel_params& params);


}  // namespace

namespace {

// The element count of buffer `b` holding items of `size` bytes, checked to be whole.
size_t Items(const tg_buf& b, size_t size, const char* what) {
    if (b.bytes % size != 0 || (b.bytes > 0 && b.data == nullptr)) {
        throw HipError{std::string("tg_hip_run: the ") + what + " buffer is malformed"};
    }
    return b.bytes / size;
}

// crater_datums.slang and crater_body_datums.slang through RunKernel's binding: every index the
// kernels follow checked on the host, the buffers uploaded, one wave per row launched, the strip's
// heights copied back.
int RunDatums(tg_dev* dev, int kernelId, const void* kernel, float* heights,
              const tg_kernel_params& params, const tg_buf* b, char* err, size_t errLen) {
    std::lock_guard<std::mutex> lock(dev->mutex);
    return Guard(err, errLen, [&] {
        const bool body = kernelId == TG_KERNEL_CRATER_BODY_DATUMS;
        const size_t cells = size_t(params.width) * params.rows;
        const size_t recordCount = Items(b[0], sizeof(tg_crater_record), "records");
        const size_t dataCount = Items(b[1], sizeof(float), "data");
        const auto* records = static_cast<const tg_crater_record*>(b[0].data);
        const auto* shape = static_cast<const tg_crater_shape*>(kernel);  // the first member
        const uint32_t noTiles[5] = {1u, 16u, 0u, 0u, 0u};
        ValidateCraterBatch(*shape, records, recordCount, dataCount, noTiles, 5, nullptr, 0, 0);
        const size_t lastRow = size_t(params.first_row) + params.rows;
        const auto fail = [](const char* what) { throw HipError{std::string("tg_hip_run: ") + what}; };
        const size_t entryCount = Items(b[5], sizeof(uint32_t), "entries");
        const auto* entries = static_cast<const uint32_t*>(b[5].data);
        for (size_t e = 0; e < entryCount; ++e) {
            if (entries[e] >= recordCount) {
                fail("an entry names a record past the end");
            }
        }
        Check(hipSetDevice(dev->ordinal), "hipSetDevice");
        DeviceArray<float> dHeights(cells);
        dHeights.Upload(heights, cells * sizeof(float));
        DeviceArray<tg_crater_record> dRecords(std::max<size_t>(recordCount, 1));
        dRecords.Upload(records, recordCount * sizeof(tg_crater_record));
        DeviceArray<float> dData(std::max<size_t>(dataCount, 1));
        dData.Upload(b[1].data, dataCount * sizeof(float));
        DeviceArray<uint32_t> dEntries(std::max<size_t>(entryCount, 1));
        dEntries.Upload(entries, entryCount * sizeof(uint32_t));
        // Rows along x, which takes any count; a row's points (at most a lattice) along y.
        const dim3 grid(params.rows, (params.width + 63u) / 64u, 1);
        if (!body) {
            const auto& kd = *static_cast<const tg_crater_datum_kernel*>(kernel);
            if (kd.lattice_side == 0) {
                fail("the datum lattice needs a side");
            }
            const size_t bandFloats = Items(b[2], sizeof(float), "band floats");
            const size_t bandCells = Items(b[3], sizeof(uint32_t), "band cells");
            const size_t cellWords = Items(b[4], sizeof(uint32_t), "cells");
            const size_t targetCount = Items(b[6], sizeof(uint32_t), "targets");
            const auto* bc = static_cast<const uint32_t*>(b[3].data);
            const auto* cl = static_cast<const uint32_t*>(b[4].data);
            const auto* tg = static_cast<const uint32_t*>(b[6].data);
            if (bandFloats < 3ull * kd.band_count || bandCells < 3ull * kd.band_count ||
                cellWords % 2 != 0) {
                fail("the crater index holds fewer bands than the kernel names");
            }
            for (uint32_t band = 0; band < kd.band_count; ++band) {
                const uint64_t across = bc[3 * band], down = bc[3 * band + 1];
                if (across == 0 || down == 0 || across * down > 0x7fffffffull ||
                    bc[3 * band + 2] + across * down > cellWords / 2) {
                    fail("a band's cells run past the cell list");
                }
            }
            for (size_t c = 0; c < cellWords / 2; ++c) {
                if (uint64_t(cl[2 * c]) + cl[2 * c + 1] > entryCount) {
                    fail("a cell's entries run past the entry list");
                }
            }
            if (targetCount < lastRow) {
                fail("the targets hold fewer rows than the strip");
            }
            for (size_t r = params.first_row; r < lastRow; ++r) {
                if (tg[r] >= recordCount) {
                    fail("a target names a record past the end");
                }
            }
            DeviceArray<tg_crater_datum_kernel> dKernel(1);
            dKernel.Upload(&kd, sizeof kd);
            DeviceArray<float> dBandFloats(std::max<size_t>(bandFloats, 1));
            dBandFloats.Upload(b[2].data, bandFloats * sizeof(float));
            DeviceArray<uint32_t> dBandCells(std::max<size_t>(bandCells, 1));
            dBandCells.Upload(bc, bandCells * sizeof(uint32_t));
            DeviceArray<uint32_t> dCells(std::max<size_t>(cellWords, 1));
            dCells.Upload(cl, cellWords * sizeof(uint32_t));
            DeviceArray<uint32_t> dTargets(std::max<size_t>(targetCount, 1));
            dTargets.Upload(tg, targetCount * sizeof(uint32_t));
            hipLaunchKernelGGL(CraterDatumStrip, grid, dim3(64), 0, nullptr, dHeights.ptr,
                               dKernel.ptr, dRecords.ptr, dData.ptr, dBandFloats.ptr,
                               dBandCells.ptr, dCells.ptr, dEntries.ptr, dTargets.ptr, params);
        } else {
            const size_t frameCount = Items(b[2], sizeof(tg_crater_frame), "frames");
            const size_t boundCount = Items(b[3], sizeof(float), "bounds");
            const size_t listWords = Items(b[4], sizeof(uint32_t), "lists");
            const size_t positionCount = Items(b[6], sizeof(tg_split_point), "positions");
            const auto* lists = static_cast<const uint32_t*>(b[4].data);
            if (frameCount < recordCount || boundCount < recordCount) {
                fail("a body datum batch needs a frame and a bound per record");
            }
            if (listWords < 2 * lastRow) {
                fail("the lists hold fewer rows than the strip");
            }
            for (size_t r = params.first_row; r < lastRow; ++r) {
                if (uint64_t(lists[2 * r]) + lists[2 * r + 1] > entryCount) {
                    fail("a row's entries run past the entry list");
                }
            }
            if (positionCount < cells) {
                fail("the body datum kernel binds one position per point of the strip");
            }
            DeviceArray<tg_crater_body_datum_kernel> dKernel(1);
            dKernel.Upload(kernel, sizeof(tg_crater_body_datum_kernel));
            DeviceArray<tg_crater_frame> dFrames(std::max<size_t>(frameCount, 1));
            dFrames.Upload(b[2].data, frameCount * sizeof(tg_crater_frame));
            DeviceArray<float> dBounds(std::max<size_t>(boundCount, 1));
            dBounds.Upload(b[3].data, boundCount * sizeof(float));
            DeviceArray<uint32_t> dLists(std::max<size_t>(listWords, 1));
            dLists.Upload(lists, listWords * sizeof(uint32_t));
            DeviceArray<tg_split_point> dPositions(cells);
            dPositions.Upload(b[6].data, cells * sizeof(tg_split_point));
            hipLaunchKernelGGL(CraterBodyDatumStrip, grid, dim3(64), 0, nullptr, dHeights.ptr,
                               dKernel.ptr, dRecords.ptr, dData.ptr, dFrames.ptr, dBounds.ptr,
                               dLists.ptr, dEntries.ptr, dPositions.ptr, params);
        }
        Check(hipGetLastError(), "launching the datum kernel");
        Check(hipDeviceSynchronize(), "running the datum kernel");
        Check(hipMemcpy(heights, dHeights.ptr, cells * sizeof(float), hipMemcpyDeviceToHost),
              "hipMemcpy to host");
    });
}

}  // namespace

extern "C" {

int tg_hip_open(int ordinal, tg_dev** out, char* err, size_t errLen) {
    static std::mutex openMutex;
    static std::vector<std::unique_ptr<tg_dev>> devices;
    if (out == nullptr) {
        SetErr(err, errLen, "tg_hip_open: null output");
        return TG_ERR_ARGUMENT;
    }
    *out = nullptr;
    int count = 0;
    const hipError_t status = hipGetDeviceCount(&count);
    if (status != hipSuccess || count == 0) {
        SetErr(err, errLen,
               status != hipSuccess ? std::string("no HIP device: ") + hipGetErrorString(status)
                                    : std::string("no HIP device"));
        return TG_ERR_NO_DEVICE;
    }
    if (ordinal < 0 || ordinal >= count) {
        SetErr(err, errLen, "HIP device ordinal " + std::to_string(ordinal) + " out of range");
        return TG_ERR_NO_DEVICE;
    }
    std::lock_guard<std::mutex> lock(openMutex);
    for (const auto& d : devices) {
        if (d->ordinal == ordinal) {
            *out = d.get();
            return TG_OK;
        }
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, ordinal) != hipSuccess) {
        SetErr(err, errLen, "hipGetDeviceProperties failed");
        return TG_ERR_DEVICE;
    }
    auto dev = std::make_unique<tg_dev>();
    dev->ordinal = ordinal;
    // The marketing name where the runtime reports one, else a generic one; then the target.
    char marketing[256] = {0};
    std::string name;
    if (hipDeviceGetName(marketing, sizeof marketing, ordinal) == hipSuccess) {
        name = marketing;
    }
    if (name.empty()) {
        name = prop.name;
    }
    if (name.empty()) {
        name = "AMD GPU " + std::to_string(ordinal);
    }
    dev->name = name + " (" + prop.gcnArchName + ")";
    // This device's address of the static scratch, with the device current.
    void* at = nullptr;
    hipError_t found = hipSetDevice(ordinal);
    if (found == hipSuccess) {
        found = hipGetSymbolAddress(&at, HIP_SYMBOL(scratch));
    }
    if (found == hipSuccess && at != nullptr) {
        dev->scratch = static_cast<unsigned char*>(at);
    } else {
        dev->scratchError = std::string("the static scratch has no device address: ") +
                            hipGetErrorString(found);
        (void)hipGetLastError();
    }
    *out = dev.get();
    devices.push_back(std::move(dev));
    return TG_OK;
}

const char* tg_hip_device_name(const tg_dev* dev) {
    return dev != nullptr ? dev->name.c_str() : "";
}

int tg_hip_can_run(const tg_dev* dev, int kernelId) {
    return dev != nullptr && (kernelId == TG_KERNEL_NOISE_LAYER || kernelId == TG_KERNEL_CRATERS ||
                              kernelId == TG_KERNEL_BOULDERS || kernelId == TG_KERNEL_NOISE_BODY ||
                              kernelId == TG_KERNEL_CRATER_DATUMS ||
                              kernelId == TG_KERNEL_CRATER_BODY_DATUMS);
}

int tg_hip_close(tg_dev* dev) {
    (void)dev;
    return TG_OK;
}

int tg_hip_run(tg_dev* dev, int kernelId, const void* kernel, size_t kernelBytes, float* heights,
               const float* xs, const float* ys, const tg_kernel_params* params,
               const tg_buf* extra, size_t nExtra, char* err, size_t errLen) {
    if (dev == nullptr || kernel == nullptr || params == nullptr) {
        SetErr(err, errLen, "tg_hip_run: null argument");
        return TG_ERR_ARGUMENT;
    }
    const size_t width = params->width;
    const size_t rows = params->rows;
    const size_t cells = width * rows;
    if (cells > 0 && (heights == nullptr || xs == nullptr || ys == nullptr)) {
        SetErr(err, errLen, "tg_hip_run: null grid buffer");
        return TG_ERR_ARGUMENT;
    }
    if (kernelId == TG_KERNEL_NOISE_LAYER || kernelId == TG_KERNEL_NOISE_BODY) {
        if (kernelBytes != sizeof(tg_noise_layer_kernel)) {
            SetErr(err, errLen, "tg_hip_run: a NoiseLayerKernel is 420 bytes");
            return TG_ERR_ARGUMENT;
        }
        if (kernelId == TG_KERNEL_NOISE_BODY &&
            (nExtra != 1 || extra == nullptr ||
             (cells > 0 && (extra[0].data == nullptr || extra[0].bytes < cells * 3 * sizeof(double))))) {
            SetErr(err, errLen, "tg_hip_run: the body noise binds one position (3 doubles) per cell");
            return TG_ERR_ARGUMENT;
        }
    } else if (kernelId == TG_KERNEL_CRATERS) {
        if (kernelBytes != sizeof(tg_crater_shape)) {
            SetErr(err, errLen, "tg_hip_run: a CraterShapeKernel is 188 bytes");
            return TG_ERR_ARGUMENT;
        }
        if (nExtra != 7 || extra == nullptr) {
            SetErr(err, errLen,
                   "tg_hip_run: the crater kernel binds records, data, tiles, entries, frames, "
                   "points and lows");
            return TG_ERR_ARGUMENT;
        }
    } else if (kernelId == TG_KERNEL_BOULDERS) {
        if (kernelBytes != sizeof(tg_boulder_kernel)) {
            SetErr(err, errLen, "tg_hip_run: a tg_boulder_kernel is 16 bytes");
            return TG_ERR_ARGUMENT;
        }
        if (nExtra != 4 || extra == nullptr) {
            SetErr(err, errLen, "tg_hip_run: the boulder kernel binds boulders, tiles, entries "
                                "and lows");
            return TG_ERR_ARGUMENT;
        }
    } else if (kernelId == TG_KERNEL_CRATER_DATUMS || kernelId == TG_KERNEL_CRATER_BODY_DATUMS) {
        const bool body = kernelId == TG_KERNEL_CRATER_BODY_DATUMS;
        if (kernelBytes != (body ? sizeof(tg_crater_body_datum_kernel)
                                 : sizeof(tg_crater_datum_kernel))) {
            SetErr(err, errLen, body ? "tg_hip_run: a CraterBodyDatumKernel is 188 bytes"
                                     : "tg_hip_run: a CraterDatumKernel is 200 bytes");
            return TG_ERR_ARGUMENT;
        }
        if (nExtra != 7 || extra == nullptr) {
            SetErr(err, errLen,
                   body ? "tg_hip_run: the body datum kernel binds records, data, frames, bounds, "
                          "lists, entries and positions"
                        : "tg_hip_run: the datum kernel binds records, data, band floats, band "
                          "cells, cells, entries and targets");
            return TG_ERR_ARGUMENT;
        }
    } else {
        SetErr(err, errLen, "tg_hip_run: unknown kernel id " + std::to_string(kernelId));
        return TG_ERR_ARGUMENT;
    }
    if (cells == 0) {
        return TG_OK;
    }
    if (kernelId == TG_KERNEL_CRATER_DATUMS || kernelId == TG_KERNEL_CRATER_BODY_DATUMS) {
        return RunDatums(dev, kernelId, kernel, heights, *params, extra, err, errLen);
    }
    std::lock_guard<std::mutex> lock(dev->mutex);
    return Guard(err, errLen, [&] {
        Check(hipSetDevice(dev->ordinal), "hipSetDevice");
        DeviceArray<float> dHeights(cells);
        DeviceArray<float> dXs(width);
        DeviceArray<float> dYs(rows);
        dHeights.Upload(heights, cells * sizeof(float));
        dXs.Upload(xs, width * sizeof(float));
        dYs.Upload(ys, rows * sizeof(float));
        if (kernelId == TG_KERNEL_NOISE_LAYER) {
            tg_noise_layer_kernel k;
            std::memcpy(&k, kernel, sizeof k);
            auto plan = std::make_unique<tg::NoisePlan>();
            tg::MakeNoisePlan(k, *plan);
            DeviceArray<tg::NoisePlan> dPlan(1);
            dPlan.Upload(plan.get(), sizeof(tg::NoisePlan));
            LaunchNoise(dPlan.ptr, dHeights.ptr, dXs.ptr, dYs.ptr, *params, nullptr);
        } else if (kernelId == TG_KERNEL_NOISE_BODY) {
            tg_noise_layer_kernel k;
            std::memcpy(&k, kernel, sizeof k);
            auto plan = std::make_unique<tg::NoisePlan>();
            tg::MakeNoisePlan(k, *plan);
            DeviceArray<tg::NoisePlan> dPlan(1);
            dPlan.Upload(plan.get(), sizeof(tg::NoisePlan));
            DeviceArray<double> dPositions(cells * 3);
            dPositions.Upload(extra[0].data, cells * 3 * sizeof(double));
            const uint32_t block = 256;
            hipLaunchKernelGGL(NoiseBodyStrip, dim3(uint32_t((cells + block - 1) / block)),
                               dim3(block), 0, nullptr, dHeights.ptr, dPositions.ptr, dPlan.ptr,
                               *params, 0, 0.0f);
            Check(hipGetLastError(), "launching the body noise kernel");
        } else if (kernelId == TG_KERNEL_BOULDERS) {
            const tg_buf* b = extra;
            tg_boulder_set* set = nullptr;
            char inner[512] = {0};
            const int code = tg_hip_boulder_set_create(
                dev, static_cast<const tg_boulder_kernel*>(kernel),
                static_cast<const tg_boulder*>(b[0].data), b[0].bytes / sizeof(tg_boulder),
                static_cast<const uint32_t*>(b[1].data), b[1].bytes / sizeof(uint32_t),
                static_cast<const uint32_t*>(b[2].data), b[2].bytes / sizeof(uint32_t),
                static_cast<const float*>(b[3].data), b[3].bytes / sizeof(float), &set, inner,
                sizeof inner);
            if (code != TG_OK) {
                throw HipError{inner};
            }
            std::unique_ptr<tg_boulder_set, int (*)(tg_boulder_set*)> owner(
                set, tg_hip_boulder_set_free);
            LaunchBoulders(*set, dHeights.ptr, dXs.ptr, dYs.ptr, *params, nullptr);
        } else {
            // Binding order of craters.slang: records, data, tiles, entries, frames, points, lows.
            const tg_buf* b = extra;
            const size_t frameCount = b[4].bytes / sizeof(tg_crater_frame);
            const size_t pointCount = b[5].bytes / sizeof(tg_split_point);
            tg_crater_set* set = nullptr;
            char inner[512] = {0};
            const int code = CreateCraterSet(
                dev, static_cast<const tg_crater_shape*>(kernel),
                static_cast<const tg_crater_record*>(b[0].data), b[0].bytes / sizeof(tg_crater_record),
                static_cast<const float*>(b[1].data), b[1].bytes / sizeof(float),
                static_cast<const uint32_t*>(b[2].data), b[2].bytes / sizeof(uint32_t),
                static_cast<const uint32_t*>(b[3].data), b[3].bytes / sizeof(uint32_t),
                static_cast<const tg_crater_frame*>(b[4].data), frameCount,
                static_cast<const tg_split_point*>(b[5].data), pointCount,
                static_cast<const float*>(b[6].data), b[6].bytes / sizeof(float), &set, inner,
                sizeof inner);
            if (code != TG_OK) {
                throw HipError{inner};
            }
            std::unique_ptr<tg_crater_set, int (*)(tg_crater_set*)> owner(set,
                                                                          tg_hip_crater_set_free);
            CheckLaunch(*set, *params);
            LaunchCraters(*set, dHeights.ptr, dXs.ptr, dYs.ptr, *params, nullptr);
        }
        Check(hipDeviceSynchronize(), "running the kernel");
        Check(hipMemcpy(heights, dHeights.ptr, cells * sizeof(float), hipMemcpyDeviceToHost),
              "hipMemcpy to host");
    });
}

int tg_hip_noise_plan_create(tg_dev* dev, const tg_noise_layer_kernel* kernel, tg_noise_plan** out,
                             char* err, size_t errLen) {
    if (dev == nullptr || kernel == nullptr || out == nullptr) {
        SetErr(err, errLen, "tg_hip_noise_plan_create: null argument");
        return TG_ERR_ARGUMENT;
    }
    *out = nullptr;
    return Guard(err, errLen, [&] {
        Check(hipSetDevice(dev->ordinal), "hipSetDevice");
        auto plan = std::make_unique<tg_noise_plan>();
        plan->dev = dev;
        tg::MakeNoisePlan(*kernel, plan->host);
        Check(hipMalloc(reinterpret_cast<void**>(&plan->device), sizeof(tg::NoisePlan)),
              "hipMalloc");
        const hipError_t copied =
            hipMemcpy(plan->device, &plan->host, sizeof(tg::NoisePlan), hipMemcpyHostToDevice);
        if (copied != hipSuccess) {
            (void)hipFree(plan->device);
            Check(copied, "hipMemcpy to device");
        }
        *out = plan.release();
    });
}

int tg_hip_noise_plan_free(tg_noise_plan* plan) {
    if (plan == nullptr) {
        return TG_OK;
    }
    (void)hipSetDevice(plan->dev->ordinal);
    (void)hipFree(plan->device);
    delete plan;
    return TG_OK;
}

int tg_hip_noise_launch(tg_dev* dev, const tg_noise_plan* plan, float* heights, const float* xs,
                        const float* ys, const tg_kernel_params* params, void* stream, char* err,
                        size_t errLen) {
    if (dev == nullptr || plan == nullptr || params == nullptr || heights == nullptr ||
        xs == nullptr || ys == nullptr) {
        SetErr(err, errLen, "tg_hip_noise_launch: null argument");
        return TG_ERR_ARGUMENT;
    }
    return Guard(err, errLen, [&] {
        LaunchNoise(plan->device, heights, xs, ys, *params, static_cast<hipStream_t>(stream));
    });
}

}  // extern "C"

namespace {

template <class T> T* Upload(const T* src, size_t count, std::vector<void*>& owned) {
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

int CreateCraterSet(tg_dev* dev, const tg_crater_shape* kernel, const tg_crater_record* records,
                    size_t recordCount, const float* data, size_t dataCount,
                    const uint32_t* tiles, size_t tileWords, const uint32_t* entries,
                    size_t entryCount, const tg_crater_frame* frames, size_t frameCount,
                    const tg_split_point* points, size_t pointCount, const float* lows,
                    size_t lowCount, tg_crater_set** out, char* err, size_t errLen) {
    if (dev == nullptr || kernel == nullptr || out == nullptr || tiles == nullptr ||
        (recordCount > 0 && records == nullptr) || (entryCount > 0 && entries == nullptr)) {
        SetErr(err, errLen, "crater set: null argument");
        return TG_ERR_ARGUMENT;
    }
    *out = nullptr;
    std::vector<void*> owned;
    const int code = Guard(err, errLen, [&] {
        ValidateCraterBatch(*kernel, records, recordCount, dataCount, tiles, tileWords, entries,
                            entryCount, frames != nullptr ? frameCount : 0);
        Check(hipSetDevice(dev->ordinal), "hipSetDevice");
        auto set = std::make_unique<tg_crater_set>();
        set->dev = dev;
        set->kernel = Upload(kernel, 1, owned);
        set->records = Upload(records, recordCount, owned);
        set->data = Upload(data, dataCount, owned);
        set->tiles = Upload(tiles, tileWords, owned);
        set->entries = Upload(entries, entryCount, owned);
        set->frames = Upload(frames, frames != nullptr ? frameCount : 0, owned);
        set->points = Upload(points, points != nullptr ? pointCount : 0, owned);
        set->pointCount = points != nullptr ? pointCount : 0;
        set->lows = Upload(lows, lowCount, owned);
        set->lowCount = lowCount;
        const std::vector<float> cull = CullRadii(*kernel, records, recordCount, data, dataCount);
        set->cull = Upload(cull.data(), cull.size(), owned);
        std::vector<float> harm;
        std::vector<float4> outline;
        OutlineTerms(records, recordCount, data, dataCount, harm, outline);
        set->harm = Upload(harm.data(), harm.size(), owned);
        set->outline = Upload(outline.data(), outline.size(), owned);
        set->terms = DeviceMpdTerms(set->kernel, set->records, recordCount);
        owned.push_back(set->terms);
        set->precise = DevicePreciseTerms(set->kernel, set->records,
                                          tiles[2] != 0 && frames != nullptr ? set->frames
                                                                             : nullptr,
                                          recordCount);
        owned.push_back(set->precise);
        set->allFast = AllMpdFast(*kernel, records, recordCount);
        set->anyPrecise = AnyPrecise(*kernel, records, recordCount);
        set->tileSide = tiles[1];
        set->tilesAcross = tiles[0];
        set->tileRows = uint32_t((tileWords - 3) / 2 / tiles[0]);
        set->onBody = tiles[2] != 0;
        owned.clear();
        *out = set.release();
    });
    for (void* p : owned) {
        (void)hipFree(p);
    }
    return code;
}

// A launch's grid against the set's tiles and per-column/row buffers.
void CheckLaunch(const tg_crater_set& set, const tg_kernel_params& params) {
    if (params.rows == 0 || params.width == 0) {
        return;
    }
    const uint64_t lastRow = uint64_t(params.first_row) + params.rows - 1;
    if ((params.width - 1) / set.tileSide >= set.tilesAcross ||
        lastRow / set.tileSide >= set.tileRows) {
        throw HipError{"the strip reaches past the crater tiles"};
    }
    if (set.onBody) {
        if (set.pointCount < size_t(params.rows) * params.width) {
            throw HipError{"a body strip needs one split point per cell"};
        }
    } else if (set.lowCount < size_t(params.width) + lastRow + 1) {
        throw HipError{"lows must hold every column, then every row up to the strip's last"};
    }
}

}  // namespace

extern "C" {

int tg_hip_crater_set_create(tg_dev* dev, const tg_crater_shape* kernel,
                             const tg_crater_record* records, size_t recordCount,
                             const float* data, size_t dataCount, const uint32_t* tiles,
                             size_t tileWords, const uint32_t* entries, size_t entryCount,
                             const tg_crater_frame* frames, size_t frameCount, const float* lows,
                             size_t lowCount, tg_crater_set** out, char* err, size_t errLen) {
    return CreateCraterSet(dev, kernel, records, recordCount, data, dataCount, tiles, tileWords,
                           entries, entryCount, frames, frameCount, nullptr, 0, lows, lowCount,
                           out, err, errLen);
}

int tg_hip_crater_set_free(tg_crater_set* set) {
    if (set == nullptr) {
        return TG_OK;
    }
    (void)hipSetDevice(set->dev->ordinal);
    for (void* p : {static_cast<void*>(set->kernel), static_cast<void*>(set->records),
                    static_cast<void*>(set->data), static_cast<void*>(set->tiles),
                    static_cast<void*>(set->entries), static_cast<void*>(set->frames),
                    static_cast<void*>(set->points), static_cast<void*>(set->lows),
                    static_cast<void*>(set->cull), static_cast<void*>(set->harm),
                    static_cast<void*>(set->outline), static_cast<void*>(set->terms),
                    static_cast<void*>(set->precise)}) {
        (void)hipFree(p);
    }
    delete set;
    return TG_OK;
}

int tg_hip_crater_launch(tg_dev* dev, const tg_crater_set* set, float* heights, const float* xs,
                         const float* ys, const tg_kernel_params* params, void* stream, char* err,
                         size_t errLen) {
    if (dev == nullptr || set == nullptr || params == nullptr || heights == nullptr ||
        xs == nullptr || ys == nullptr) {
        SetErr(err, errLen, "tg_hip_crater_launch: null argument");
        return TG_ERR_ARGUMENT;
    }
    if (set->onBody) {
        SetErr(err, errLen, "tg_hip_crater_launch: the set is binned on a body; use the batched path");
        return TG_ERR_ARGUMENT;
    }
    return Guard(err, errLen, [&] {
        CheckLaunch(*set, *params);
        LaunchCraters(*set, heights, xs, ys, *params, static_cast<hipStream_t>(stream));
    });
}

int tg_hip_crater_boulder_launch(tg_dev* dev, const tg_crater_set* set,
                                 const tg_boulder_set* boulders, float* heights, const float* xs,
                                 const float* ys, const tg_kernel_params* params, void* stream,
                                 char* err, size_t errLen) {
    if (dev == nullptr || set == nullptr || boulders == nullptr || params == nullptr ||
        heights == nullptr || xs == nullptr || ys == nullptr) {
        SetErr(err, errLen, "tg_hip_crater_boulder_launch: null argument");
        return TG_ERR_ARGUMENT;
    }
    if (set->onBody) {
        SetErr(err, errLen,
               "tg_hip_crater_boulder_launch: the set is binned on a body; use the batched path");
        return TG_ERR_ARGUMENT;
    }
    return Guard(err, errLen, [&] {
        CheckLaunch(*set, *params);
        // The boulder set's tiles and lows must cover the strip as its own launch requires.
        if (uint64_t(boulders->tilesAcross) * boulders->tileSide < params->width ||
            uint64_t(boulders->tileRows) * boulders->tileSide <
                uint64_t(params->first_row) + params->rows ||
            boulders->lowCount < size_t(params->width) + params->first_row + params->rows) {
            throw HipError{"the boulder tiles or lows do not cover the grid"};
        }
        LaunchCraters(*set, heights, xs, ys, *params, static_cast<hipStream_t>(stream),
                      boulders);
    });
}

int tg_hip_noise_fill(tg_dev* dev, const tg_noise_plan* plan, float elevation, float* heights,
                      const float* xs, const float* ys, const tg_kernel_params* params,
                      void* stream, char* err, size_t errLen) {
    if (dev == nullptr || plan == nullptr || params == nullptr || heights == nullptr ||
        xs == nullptr || ys == nullptr) {
        SetErr(err, errLen, "tg_hip_noise_fill: null argument");
        return TG_ERR_ARGUMENT;
    }
    return Guard(err, errLen, [&] {
        LaunchNoise(plan->device, heights, xs, ys, *params, static_cast<hipStream_t>(stream),
                    true, elevation);
    });
}

int tg_hip_malloc(tg_dev* dev, size_t bytes, void** out, char* err, size_t errLen) {
    if (dev == nullptr || out == nullptr) {
        SetErr(err, errLen, "tg_hip_malloc: null argument");
        return TG_ERR_ARGUMENT;
    }
    *out = nullptr;
    return Guard(err, errLen, [&] {
        Check(hipSetDevice(dev->ordinal), "hipSetDevice");
        Check(hipMalloc(out, bytes > 0 ? bytes : 1), "hipMalloc");
    });
}

int tg_hip_free(tg_dev* dev, void* ptr) {
    if (dev == nullptr || ptr == nullptr) {
        return TG_OK;
    }
    (void)hipSetDevice(dev->ordinal);
    return hipFree(ptr) == hipSuccess ? TG_OK : TG_ERR_DEVICE;
}

int tg_hip_memcpy_h2d(tg_dev* dev, void* dst, const void* src, size_t bytes, char* err,
                      size_t errLen) {
    if (dev == nullptr || (bytes > 0 && (dst == nullptr || src == nullptr))) {
        SetErr(err, errLen, "tg_hip_memcpy_h2d: null argument");
        return TG_ERR_ARGUMENT;
    }
    return Guard(err, errLen, [&] {
        Check(hipSetDevice(dev->ordinal), "hipSetDevice");
        if (bytes > 0) {
            Check(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy to device");
        }
    });
}

int tg_hip_memcpy_d2h(tg_dev* dev, void* dst, const void* src, size_t bytes, char* err,
                      size_t errLen) {
    if (dev == nullptr || (bytes > 0 && (dst == nullptr || src == nullptr))) {
        SetErr(err, errLen, "tg_hip_memcpy_d2h: null argument");
        return TG_ERR_ARGUMENT;
    }
    return Guard(err, errLen, [&] {
        Check(hipSetDevice(dev->ordinal), "hipSetDevice");
        if (bytes > 0) {
            Check(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "hipMemcpy to host");
        }
    });
}

int tg_hip_memset(tg_dev* dev, void* dst, int value, size_t bytes, char* err, size_t errLen) {
    if (dev == nullptr || (bytes > 0 && dst == nullptr)) {
        SetErr(err, errLen, "tg_hip_memset: null argument");
        return TG_ERR_ARGUMENT;
    }
    return Guard(err, errLen, [&] {
        Check(hipSetDevice(dev->ordinal), "hipSetDevice");
        if (bytes > 0) {
            Check(hipMemset(dst, value, bytes), "hipMemset");
        }
    });
}

int tg_hip_stream_create(tg_dev* dev, void** out, char* err, size_t errLen) {
    if (dev == nullptr || out == nullptr) {
        SetErr(err, errLen, "tg_hip_stream_create: null argument");
        return TG_ERR_ARGUMENT;
    }
    return Guard(err, errLen, [&] {
        Check(hipSetDevice(dev->ordinal), "hipSetDevice");
        hipStream_t s = nullptr;
        Check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
        *out = s;
    });
}

int tg_hip_stream_destroy(tg_dev* dev, void* stream) {
    if (dev == nullptr || stream == nullptr) {
        return TG_OK;
    }
    (void)hipSetDevice(dev->ordinal);
    return hipStreamDestroy(static_cast<hipStream_t>(stream)) == hipSuccess ? TG_OK
                                                                             : TG_ERR_DEVICE;
}

int tg_hip_stream_synchronize(tg_dev* dev, void* stream, char* err, size_t errLen) {
    if (dev == nullptr) {
        SetErr(err, errLen, "tg_hip_stream_synchronize: null device");
        return TG_ERR_ARGUMENT;
    }
    return Guard(err, errLen, [&] {
        Check(hipSetDevice(dev->ordinal), "hipSetDevice");
        Check(hipStreamSynchronize(static_cast<hipStream_t>(stream)), "hipStreamSynchronize");
    });
}

int tg_hip_event_create(tg_dev* dev, void** out, char* err, size_t errLen) {
    if (dev == nullptr || out == nullptr) {
        SetErr(err, errLen, "tg_hip_event_create: null argument");
        return TG_ERR_ARGUMENT;
    }
    return Guard(err, errLen, [&] {
        Check(hipSetDevice(dev->ordinal), "hipSetDevice");
        hipEvent_t e = nullptr;
        Check(hipEventCreate(&e), "hipEventCreate");
        *out = e;
    });
}

int tg_hip_event_destroy(tg_dev* dev, void* event) {
    if (dev == nullptr || event == nullptr) {
        return TG_OK;
    }
    (void)hipSetDevice(dev->ordinal);
    return hipEventDestroy(static_cast<hipEvent_t>(event)) == hipSuccess ? TG_OK : TG_ERR_DEVICE;
}

int tg_hip_event_record(tg_dev* dev, void* event, void* stream, char* err, size_t errLen) {
    if (dev == nullptr || event == nullptr) {
        SetErr(err, errLen, "tg_hip_event_record: null argument");
        return TG_ERR_ARGUMENT;
    }
    return Guard(err, errLen, [&] {
        Check(hipEventRecord(static_cast<hipEvent_t>(event), static_cast<hipStream_t>(stream)),
              "hipEventRecord");
    });
}

int tg_hip_event_elapsed(tg_dev* dev, void* start, void* stop, float* ms, char* err,
                         size_t errLen) {
    if (dev == nullptr || start == nullptr || stop == nullptr || ms == nullptr) {
        SetErr(err, errLen, "tg_hip_event_elapsed: null argument");
        return TG_ERR_ARGUMENT;
    }
    return Guard(err, errLen, [&] {
        Check(hipEventSynchronize(static_cast<hipEvent_t>(stop)), "hipEventSynchronize");
        Check(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start),
                                  static_cast<hipEvent_t>(stop)),
              "hipEventElapsedTime");
    });
}

int tg_hip_host_alloc(tg_dev* dev, size_t bytes, void** out, char* err, size_t errLen) {
    if (dev == nullptr || out == nullptr) {
        SetErr(err, errLen, "tg_hip_host_alloc: null argument");
        return TG_ERR_ARGUMENT;
    }
    *out = nullptr;
    return Guard(err, errLen, [&] {
        Check(hipSetDevice(dev->ordinal), "hipSetDevice");
        Check(hipHostMalloc(out, bytes > 0 ? bytes : 1, hipHostMallocDefault), "hipHostMalloc");
    });
}

int tg_hip_host_free(tg_dev* dev, void* ptr) {
    if (dev == nullptr) {
        return TG_ERR_ARGUMENT;
    }
    if (ptr != nullptr) {
        (void)hipSetDevice(dev->ordinal);
        (void)hipHostFree(ptr);
    }
    return TG_OK;
}

int tg_hip_memcpy_d2h_async(tg_dev* dev, void* dst, const void* src, size_t bytes, void* stream,
                            char* err, size_t errLen) {
    if (dev == nullptr || (bytes > 0 && (dst == nullptr || src == nullptr))) {
        SetErr(err, errLen, "tg_hip_memcpy_d2h_async: null argument");
        return TG_ERR_ARGUMENT;
    }
    return Guard(err, errLen, [&] {
        Check(hipSetDevice(dev->ordinal), "hipSetDevice");
        if (bytes > 0) {
            Check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost,
                                 static_cast<hipStream_t>(stream)),
                  "hipMemcpyAsync to host");
        }
    });
}

int tg_hip_memcpy_h2d_async(tg_dev* dev, void* dst, const void* src, size_t bytes, void* stream,
                            char* err, size_t errLen) {
    if (dev == nullptr || (bytes > 0 && (dst == nullptr || src == nullptr))) {
        SetErr(err, errLen, "tg_hip_memcpy_h2d_async: null argument");
        return TG_ERR_ARGUMENT;
    }
    return Guard(err, errLen, [&] {
        Check(hipSetDevice(dev->ordinal), "hipSetDevice");
        if (bytes > 0) {
            Check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice,
                                 static_cast<hipStream_t>(stream)),
                  "hipMemcpyAsync to device");
        }
    });
}

int tg_hip_memcpy_d2d_async(tg_dev* dev, void* dst, const void* src, size_t bytes, void* stream,
                            char* err, size_t errLen) {
    if (dev == nullptr || (bytes > 0 && (dst == nullptr || src == nullptr))) {
        SetErr(err, errLen, "tg_hip_memcpy_d2d_async: null argument");
        return TG_ERR_ARGUMENT;
    }
    return Guard(err, errLen, [&] {
        Check(hipSetDevice(dev->ordinal), "hipSetDevice");
        if (bytes > 0) {
            Check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice,
                                 static_cast<hipStream_t>(stream)),
                  "hipMemcpyAsync on the device");
        }
    });
}

int tg_hip_fill(tg_dev* dev, float* dst, float value, size_t count, void* stream, char* err,
                size_t errLen) {
    if (dev == nullptr || (count > 0 && dst == nullptr)) {
        SetErr(err, errLen, "tg_hip_fill: null argument");
        return TG_ERR_ARGUMENT;
    }
    return Guard(err, errLen, [&] {
        Check(hipSetDevice(dev->ordinal), "hipSetDevice");
        if (count == 0) {
            return;
        }
        const size_t blocks = std::min<size_t>((count + 255) / 256, size_t(1) << 20);
        hipLaunchKernelGGL(FillValue, dim3(uint32_t(blocks)), dim3(256), 0,
                           static_cast<hipStream_t>(stream), dst, value, count);
        Check(hipGetLastError(), "launching the fill kernel");
    });
}

int tg_hip_event_query(tg_dev* dev, void* event, char* err, size_t errLen) {
    if (dev == nullptr || event == nullptr) {
        SetErr(err, errLen, "tg_hip_event_query: null argument");
        return TG_ERR_ARGUMENT;
    }
    const hipError_t status = hipEventQuery(static_cast<hipEvent_t>(event));
    if (status == hipSuccess) {
        return TG_OK;
    }
    if (status == hipErrorNotReady) {
        return TG_NOT_READY;
    }
    SetErr(err, errLen, std::string("hipEventQuery: ") + hipGetErrorString(status));
    return TG_ERR_DEVICE;
}

}  // extern "C"

namespace {

__global__ void MakeMpdTermsKernel(const tg_crater_shape* __restrict__ kernel,
                                   const tg_crater_record* __restrict__ records, uint32_t n,
                                   tg::MpdTerms* __restrict__ terms) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        terms[i] = tg::MakeMpdTerms(*kernel, records[i]);
    }
}

__global__ void MakePreciseTermsKernel(const tg_crater_shape* __restrict__ kernel,
                                       const tg_crater_record* __restrict__ records,
                                       const tg_crater_frame* __restrict__ frames, uint32_t n,
                                       tg::PreciseTerms* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        out[i] = tg::MakePreciseTerms(*kernel, records[i], frames != nullptr ? frames + i : nullptr);
    }
}

}  // namespace

namespace tgi {

tg::PreciseTerms* DevicePreciseTerms(const tg_crater_shape* kernelOnDevice,
                                     const tg_crater_record* recordsOnDevice,
                                     const tg_crater_frame* framesOnDevice, size_t n) {
    if (n > 0xFFFFFFFFull) {
        throw HipError{"too many craters for the terms kernel"};
    }
    void* p = nullptr;
    Check(hipMalloc(&p, (n > 0 ? n : 1) * sizeof(tg::PreciseTerms)), "hipMalloc");
    if (n > 0) {
        const uint32_t threads = 256;
        const uint32_t blocks = uint32_t((n + threads - 1) / threads);
        hipLaunchKernelGGL(MakePreciseTermsKernel, dim3(blocks), dim3(threads), 0, nullptr,
                           kernelOnDevice, recordsOnDevice, framesOnDevice, uint32_t(n),
                           static_cast<tg::PreciseTerms*>(p));
        hipError_t status = hipGetLastError();
        if (status == hipSuccess) {
            status = hipDeviceSynchronize();
        }
        if (status != hipSuccess) {
            (void)hipFree(p);
            Check(status, "computing the precise crater terms");
        }
    }
    return static_cast<tg::PreciseTerms*>(p);
}

tg::MpdTerms* DeviceMpdTerms(const tg_crater_shape* kernelOnDevice,
                             const tg_crater_record* recordsOnDevice, size_t n) {
    if (n > 0xFFFFFFFFull) {
        throw HipError{"too many craters for the terms kernel"};
    }
    void* p = nullptr;
    Check(hipMalloc(&p, (n > 0 ? n : 1) * sizeof(tg::MpdTerms)), "hipMalloc");
    if (n > 0) {
        const uint32_t threads = 256;
        const uint32_t blocks = uint32_t((n + threads - 1) / threads);
        hipLaunchKernelGGL(MakeMpdTermsKernel, dim3(blocks), dim3(threads), 0, nullptr,
                           kernelOnDevice, recordsOnDevice, uint32_t(n),
                           static_cast<tg::MpdTerms*>(p));
        hipError_t status = hipGetLastError();
        if (status == hipSuccess) {
            status = hipDeviceSynchronize();
        }
        if (status != hipSuccess) {
            (void)hipFree(p);
            Check(status, "computing the crater terms");
        }
    }
    return static_cast<tg::MpdTerms*>(p);
}

bool AllMpdFast(const tg_crater_shape& k, const tg_crater_record* records, size_t n) {
    if (!tg::MpdShapeOnly(k)) {
        return n == 0;
    }
    for (size_t i = 0; i < n; ++i) {
        if (!tg::MpdRecordFast(records[i])) {
            return false;
        }
    }
    return true;
}

bool AnyPrecise(const tg_crater_shape& k, const tg_crater_record* records, size_t n) {
    if (!tg::MpdShapeOnly(k)) {
        return false;
    }
    for (size_t i = 0; i < n; ++i) {
        if (tg::MpdRecordFast(records[i]) && records[i].diameter >= tg::kPreciseDiameter) {
            return true;
        }
    }
    return false;
}

// The squared round distance from a crater's centre beyond which ApplyCrater leaves every height
// unchanged (craters.h:1186-1193): reach x radius x an upper bound of its outline scale
// (KernelOutlineScaleAt, craters.h:772-808: the ellipse stretches by at most sqrt(ellipticity),
// the wobble by at most 1 + sum|A_k| + max|table| + |table - broad|), with relative slack far
// above FP32 rounding. Infinite, so never culled, for craters whose distance other terms can
// shorten: terraces, marks and floor shifts.
std::vector<float> CullRadii(const tg_crater_shape& k, const tg_crater_record* records, size_t n,
                             const float* data, size_t dataCount) {
    std::vector<float> cull(n > 0 ? n : 1, INFINITY);
    for (size_t i = 0; i < n; ++i) {
        const tg_crater_record& c = records[i];
        if (c.terrace_count > 0 || c.mark_count > 0 || c.shift_x != 0.0f || c.shift_y != 0.0f) {
            continue;
        }
        double amplitudes = 0.0;
        for (uint32_t h = 0; h < c.harmonic_count; ++h) {
            const size_t at = size_t(c.harmonic_offset) + 2 * h;
            if (at >= dataCount) {
                amplitudes = INFINITY;
                break;
            }
            amplitudes += std::fabs(double(data[at]));
        }
        double table = 0.0;
        for (uint32_t t = 0; t < c.table_count; ++t) {
            const size_t at = size_t(c.table_offset) + t;
            table = at < dataCount ? std::fmax(table, std::fabs(double(data[at]))) : INFINITY;
        }
        const double wobble =
            c.table_count > 0 ? 1.0 + 2.0 * amplitudes + 2.0 * table : 1.0 + amplitudes;
        const double scale = std::sqrt(std::fmax(double(c.ellipticity), 1.0)) * wobble;
        const double r = double(k.reach_radii) * 0.5 * double(c.diameter) * scale * (1.0 + 1e-4);
        if (std::isfinite(r)) {
            cull[i] = float(r * r);
        }
    }
    return cull;
}

}  // namespace tgi

namespace tgi {

// Per-crater outline terms the fast path reads: (A cos phi, A sin phi) in each harmonic's data slot
// and (cos h, sin h, sqrt(ellipticity), 1 / sqrt(ellipticity)) per record, in FP64 then rounded.
// The harmonics carry tg::kHarmonicBlock floats of zeros past the end: the kernels load them a
// block at a time.
void OutlineTerms(const tg_crater_record* records, size_t n, const float* data, size_t dataCount,
                  std::vector<float>& harm, std::vector<float4>& outline) {
    harm.assign(dataCount + tg::kHarmonicBlock, 0.0f);
    outline.assign(n > 0 ? n : 1, float4{1.0f, 0.0f, 1.0f, 1.0f});
    for (size_t i = 0; i < n; ++i) {
        const tg_crater_record& c = records[i];
        for (uint32_t k = 0; k < c.harmonic_count; ++k) {
            const size_t at = size_t(c.harmonic_offset) + 2 * k;
            const double a = data[at], phi = data[at + 1];
            harm[at] = float(a * std::cos(phi));
            harm[at + 1] = float(a * std::sin(phi));
        }
        const double along = std::sqrt(double(c.ellipticity));
        outline[i] = float4{float(std::cos(double(c.heading))), float(std::sin(double(c.heading))),
                            float(along), float(1.0 / along)};
    }
}

}  // namespace tgi

namespace tgi {

// Every index the crater kernel follows, checked on the host so a malformed batch is an error,
// never an out-of-bounds read: tile heads against the entries, entries against the records,
// each record's data ranges against the data (craters.h:772-1182 read them), frames on a body.
void ValidateCraterBatch(const tg_crater_shape& k, const tg_crater_record* records, size_t n,
                         size_t dataCount, const uint32_t* tiles, size_t tileWords,
                         const uint32_t* entries, size_t entryCount, size_t frameCount) {
    const auto fail = [](const std::string& what) { throw HipError{"invalid crater batch: " + what}; };
    if (tileWords < 3 || tiles[0] == 0 || tiles[1] == 0 || (tileWords - 3) % 2 != 0 ||
        ((tileWords - 3) / 2) % tiles[0] != 0) {
        fail("tile block must be [tiles across, side, body flag, (first, count) per tile]");
    }
    for (size_t t = 3; t < tileWords; t += 2) {
        if (uint64_t(tiles[t]) + tiles[t + 1] > entryCount) {
            fail("a tile's entries run past the entry list");
        }
    }
    for (size_t e = 0; e < entryCount; ++e) {
        if (entries[e] >= n && !(n == 0 && entryCount == 1)) {
            fail("an entry names a record past the end");
        }
    }
    if (tiles[2] != 0 && frameCount < n) {
        fail("a body batch needs one frame per record");
    }
    const tg_crater_profile& p = k.profile;
    const auto span = [&](uint64_t offset, uint64_t count, const char* what) {
        if (count > 0 && offset + count > dataCount) {
            fail(std::string(what) + " run past the data");
        }
    };
    if (p.diffused == 1u || p.simple == TG_PROFILE_TABULATED) {
        if (p.table_cells == 0) {
            fail("tabulated profiles need table cells");
        }
    }
    if (p.diffused == 1u) {
        if (p.age_count < 2) {
            fail("diffused profiles need at least two ages");
        }
        span(p.ages_offset, p.age_count, "diffusion ages");
    }
    for (size_t i = 0; i < n; ++i) {
        const tg_crater_record& c = records[i];
        span(c.harmonic_offset, 2ull * c.harmonic_count, "outline harmonics");
        span(c.table_offset, c.table_count, "outline tables");
        if (c.mark_count > 0) {
            span(c.mark_offset, 3ull + c.mark_count, "outline marks");
        }
        if (c.terrace_count > 0) {
            span(c.terrace_offset,
                 20ull + 8ull * (c.terrace_count - 1) + 3ull * c.terrace_count, "terraces");
        }
        if (c.peak_count > 0) {
            span(c.peak_offset, 6ull + 11ull * c.peak_count, "peak massifs");
        }
        if (p.simple == TG_PROFILE_TABULATED) {
            span(c.profile_offset, uint64_t(p.table_cells) + 1, "a tabulated profile row");
        } else if (p.diffused == 1u && c.diffusion > 0.0f) {
            span(c.profile_offset, uint64_t(p.age_count) * (uint64_t(p.table_cells) + 1),
                 "a diffusion table");
        }
    }
}

}  // namespace tgi
