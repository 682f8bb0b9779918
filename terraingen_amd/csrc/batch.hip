// SPDX-License-Identifier: BSD-3-Clause
//
// Batched tiles: one persistent launch evaluates many tiles of uneven size and pixel size. The
// work is a queue of bins (16 x 16 cells of one tile); each wave takes the next bin with one
// atomic add, evaluates its cells in registers (FP64 cell centres narrowed to FP32 as RunKernel
// narrows them, the source's noise, then the bin's craters in catalogue order) and writes the
// heights once. A wave keeps taking bins until the queue is empty, so tiles of any mix of sizes
// fill every CU without a launch per tile, and nothing is read back or allocated between tiles.

#include "boulder_device.h"
#include "crater_device.h"
#include "internal.h"
#include "noise_device.h"
#include "sphere_math.h"

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <cstring>
#include <vector>

using namespace tgi;

struct tg_catalogue {
    tg_dev* dev = nullptr;
    tg_crater_shape* kernel = nullptr;
    tg_crater_record* records = nullptr;
    float* data = nullptr;
    float* cull = nullptr;
    float* harm = nullptr;
    float4* outline = nullptr;
    double* diameters = nullptr;
    tg::MpdTerms* terms = nullptr;
    tg::HotCrater* hot = nullptr;  // the fast path's record per crater
    float* hotHarm = nullptr;      // its first tg::kHotHarmonics outline harmonics
    tg_crater_frame* frames = nullptr;  // MakeCraterFrame per crater, for tiles on a sphere
    tg::PreciseTerms* precise = nullptr;  // FP64 terms of km-scale craters (and frame axes)
    size_t count = 0;
    bool allFast = false;  // every crater takes the hoisted MPD path
    bool anyPrecise = false;  // some crater takes it in FP64 (tg::kFastPrecise)
    // Changes whenever what a launch reads of the catalogue changes (creation, frames).
    uint64_t generation = tgi::NextGeneration();
};

struct tg_boulder_catalogue {
    tg_dev* dev = nullptr;
    tg_boulder* records = nullptr;    // geometry: split centre, radius, exposed height, facets
    double* diameters = nullptr;      // FP64 diameters, for each tile's weight and reach
    tg_crater_frame* frames = nullptr;  // per boulder on a sphere, else null
    size_t count = 0;
    uint64_t generation = tgi::NextGeneration();
};

struct tg_batch {
    tg_dev* dev = nullptr;
    hipStream_t stream = nullptr;
    tg_tile_desc* tiles = nullptr;
    tg_bin* bins = nullptr;
    uint32_t* entries = nullptr;
    uint32_t* counters = nullptr;  // the queues' next offsets, then the bin count (kCountSlot)
    uint32_t* hostCounts = nullptr;  // pinned: the counters a staging copies in
    size_t maxTiles = 0, maxBins = 0, maxEntries = 0;
    // What the staged batch needs of a run: the largest entry and output end.
    size_t tileCount = 0, binCount = 0;
    uint64_t maxEntry = 0, outEnd = 0;
    bool anyEntries = false;
    // The captured graph and the arguments it was captured for.
    hipGraphExec_t graph = nullptr;
    uint64_t graphPlan = 0;       // generations of the plan and catalogue it was captured with
    uint64_t graphCatalogue = 0;
    float* graphOut = nullptr;
    float graphElevation = 0.0f;
    uint32_t graphGroups = 0;
    uint32_t graphWalk = 0;
    bool graphHoisted = true;
    uint32_t graphChunk = 0;
    // Launch choices (tg_hip_batch_tune).
    uint32_t walk = TG_WALK_CULLED;
    uint32_t groupsPerCu = 8;
    bool hoisted = true;        // the hoisted MPD path where a crater allows it
    uint32_t chunk = 8;         // bins a wave takes per atomic add
    // The tiles' surface (tg_hip_batch_surface): 0 for a plane, else a sphere of this radius.
    double radius = 0.0;
    double graphRadius = 0.0;
    // Cells per bin side of the staged batch (16 from tg_hip_batch_stage; 8, 16 or 32 from
    // tg_hip_batch_stage_windows), and entries a culled walk tests per pass (32, 64 or 128).
    uint32_t binSide = 16;
    uint32_t cullDepth = 64;
    uint32_t graphBinSide = 0;
    uint32_t graphCullDepth = 0;
    size_t entryCount = 0;  // entries staged
    // Device binning's buffers, grown as batches need them.
    tg_window* windows = nullptr;
    size_t windowCapacity = 0;
    uint32_t* tileWindows = nullptr;  // (first, count) per tile, maxTiles pairs
    uint32_t* rowFirsts = nullptr;    // per tile its first bin row in the batch, maxTiles
    uint32_t* binCounts = nullptr;    // per bin, then their exclusive sums
    uint32_t* binFirsts = nullptr;
    void* scanTemp = nullptr;
    size_t scanBytes = 0;
    // Boulder lists of the staged bins (tg_hip_batch_stage_boulders), dropped by a restage.
    tg_bin* boulderBins = nullptr;
    uint32_t* boulderEntries = nullptr;
    size_t boulderEntryCapacity = 0;
    const tg_boulder_catalogue* boulders = nullptr;  // null: no boulders staged
    uint64_t boulderGeneration = 0;  // of the lists' buffers and catalogue a graph reads
    uint64_t graphBoulders = 0;
    // Each staged tile's bins, for checking boulder lists against them.
    std::vector<uint32_t> tileBinFirst, tileBinCount;
};

namespace {

constexpr uint32_t kBinSide = 16;
constexpr uint32_t kGroupThreads = 256;  // four independent waves

// Waves per SIMD the all-fast variants are built for: 8 without the FP64 path, and with it the
// fastest of 5-8 measured on the C4 sphere batch at one launch configuration (5: 50.01 ms, 96
// VGPRs; 6: 49.06 ms, 80 VGPRs and 20 B of scratch a lane; 7: 50.54 ms, 72 and 60 B; 8: 50.76 ms,
// 64 and 88 B): fewer waves spill less of the FP64 terms but hide less latency.
#ifndef TG_PRECISE_WAVES
#define TG_PRECISE_WAVES 6
#endif
constexpr int kPreciseWaves = TG_PRECISE_WAVES;

// Waves per SIMD the full-evaluation (general) variants are built for, the fastest on the C3 grid
// as one tile: left to the compiler (130 VGPRs, 3 waves) 3.52 ms, 4 waves (126 VGPRs, no
// scratch) 2.96 ms, 5 waves 3.03 ms.
#ifndef TG_GENERAL_WAVES
#define TG_GENERAL_WAVES 4
#endif

// The bins are split into kQueues contiguous ranges, one per XCD: workgroups are dispatched to the
// XCDs in turn, so workgroup g starts on range g % kQueues, and neighbouring bins (which list the
// same craters) are shaped by waves sharing one L2. Each range has its own counter in its own
// 128-byte line; a wave takes `chunk` bins per atomic add, and when its range runs dry it moves to
// the next one, so every bin is taken once and the device stays busy to the end.
constexpr uint32_t kQueues = 8;
constexpr uint32_t kQueueStride = 32;
constexpr uint32_t kCountSlot = kQueues * kQueueStride;
constexpr uint32_t kCounterWords = kCountSlot + 8;
static_assert(kCounterWords * sizeof(uint32_t) <= tgi::kStaticCounterBytes,
              "the static scratch holds the counters");

// The counters of a batch in the static scratch: every queue offset cleared, the bin count set.
__global__ void StaticCounters(uint32_t* counters, uint32_t binCount) {
    if (threadIdx.x < kQueues) {
        counters[threadIdx.x * kQueueStride] = 0u;
    }
    if (threadIdx.x == 0) {
        counters[kCountSlot] = binCount;
    }
}

// ---- Binning on the device ---------------------------------------------------------------------
// A batch's bins from its tiles' crater windows (tgx_index_windows): a count pass gives each bin
// its number of craters, an exclusive sum over the bins places each bin's list, and a fill pass
// writes the lists (BinWindowRows below).

// The tile of global bin row R: the last tile whose first row is at or before it.
__device__ __forceinline__ uint32_t TileOfRow(const uint32_t* rowFirsts, uint32_t tileCount,
                                              uint32_t R) {
    uint32_t lo = 0u, hi = tileCount - 1u;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1u) >> 1u;
        if (rowFirsts[mid] <= R) {
            lo = mid;
        } else {
            hi = mid - 1u;
        }
    }
    return lo;
}

// Bins of up to this many across are one segment of a row; wider rows take several passes.
constexpr uint32_t kRowSegment = 512;

// One wave per row of bins of one tile: the tile's windows (its candidates, in catalogue order)
// are tested 64 at a time against the row's cells, and each window that meets them joins every
// bin of the row its columns cover, windows one after another in list order, the lanes taking the
// window's bins. So each bin lists its craters in catalogue order, as tgx_index_bin does, and a
// tile's windows are read once per row rather than once per bin. kFill false: each bin's count
// into `counts`; true: the entries at the bin's exclusive sum `firsts` and the bin records.
template <bool kFill>
__global__ void __launch_bounds__(256)
    BinWindowRows(const tg_tile_desc* __restrict__ tiles, uint32_t tileCount,
                  const uint32_t* __restrict__ rowFirsts, uint32_t rowCount,
                  const uint32_t* __restrict__ tileWindows, const tg_window* __restrict__ windows,
                  uint32_t side, uint32_t* __restrict__ counts,
                  const uint32_t* __restrict__ firsts, tg_bin* __restrict__ bins,
                  uint32_t* __restrict__ entries) {
    __shared__ uint32_t cursors[4][kRowSegment];
    const uint32_t R = blockIdx.x * 4u + (threadIdx.x >> 6u);
    const uint32_t lane = threadIdx.x & 63u;
    if (R >= rowCount) {
        return;  // uniform across the wave
    }
    uint32_t* cursor = cursors[threadIdx.x >> 6u];
    const uint32_t t = TileOfRow(rowFirsts, tileCount, R);
    const tg_tile_desc tile = tiles[t];
    const uint32_t row = R - rowFirsts[t];
    const uint32_t r0 = row * side;
    const uint32_t r1 = min(r0 + side, tile.height) - 1u;
    const uint32_t across = tile.bins_across;
    const uint32_t rowBin = tile.bin_first + row * across;
    const uint32_t first = tileWindows[2u * t];
    const uint32_t end = first + tileWindows[2u * t + 1u];
    for (uint32_t seg0 = 0; seg0 < across; seg0 += kRowSegment) {
        const uint32_t seg1 = min(seg0 + kRowSegment, across);
        for (uint32_t j = seg0 + lane; j < seg1; j += 64u) {
            cursor[j - seg0] = kFill ? firsts[rowBin + j] : 0u;
        }
        for (uint32_t base = first; base < end; base += 64u) {
            const uint32_t e = base + lane;
            bool hit = false;
            uint32_t crater = 0u, a = 0u, z = 0u;
            if (e < end) {
                const tg_window w = windows[e];
                crater = w.crater;
                a = max(w.col0 / side, seg0);
                z = min(w.col1 / side + 1u, seg1);
                hit = w.row0 <= r1 && w.row1 >= r0 && a < z;
            }
            uint64_t mask = __ballot(hit);
            while (mask != 0ull) {
                const int i = int(__builtin_ctzll(mask));
                mask &= mask - 1ull;
                const uint32_t n = uint32_t(__builtin_amdgcn_readlane(int(crater), i));
                const uint32_t from = uint32_t(__builtin_amdgcn_readlane(int(a), i));
                const uint32_t to = uint32_t(__builtin_amdgcn_readlane(int(z), i));
                for (uint32_t j = from + lane; j < to; j += 64u) {
                    const uint32_t at = cursor[j - seg0];
                    if (kFill) {
                        entries[at] = n;
                    }
                    cursor[j - seg0] = at + 1u;
                }
            }
        }
        for (uint32_t j = seg0 + lane; j < seg1; j += 64u) {
            if (kFill) {
                const uint32_t at = firsts[rowBin + j];
                bins[rowBin + j] = tg_bin{t, at, cursor[j - seg0] - at, 0u};
            } else {
                counts[rowBin + j] = cursor[j - seg0];
            }
        }
    }
}

__global__ void ResetQueue(uint32_t* counters) {
    if (threadIdx.x < kQueues) {
        counters[threadIdx.x * kQueueStride] = 0u;
    }
}

// ResolutionWeight (src/layers/fade.cpp:48-52) with the smoothstep fade from 2 to 4 samples
// across, in FP64 as the reference computes it.
__device__ __forceinline__ float ResolutionAt(double diameter, double pixel) {
    // D >= 4 p (D <= 2 p) exactly puts D / p at or above 4 (at or below 2): the scalings by 4 and
    // 2 are exact and division rounds monotonically. Most craters in reach fall there, and skip
    // the FP64 division.
    if (diameter >= 4.0 * pixel) {
        return 1.0f;
    }
    if (diameter <= 2.0 * pixel) {
        return 0.0f;
    }
    const double x = diameter / pixel;
    double falloff;
    if (x <= 2.0) {
        falloff = 1.0;
    } else if (x >= 4.0) {
        falloff = 0.0;
    } else {
        double t = (x - 2.0) / (4.0 - 2.0);
        t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
        falloff = 1.0 - t * t * (3.0 - 2.0 * t);
    }
    return float(1.0 - falloff);
}

// The crater walks of a bin's list: how a wave visits the (cell, crater) pairs. Every walk applies
// each cell's craters in list order, so all give the same heights.
enum CraterWalk : uint32_t {
    // The wave steps through the list together; a crater is applied whenever any lane is in its
    // reach (the lanes outside idle through it).
    kWalkTogether = TG_WALK_TOGETHER,
    // Each lane scans the list for its own next crater in reach and applies it; the wave repeats
    // until no lane finds another.
    kWalkPerLane = TG_WALK_PER_LANE,
    // The wave tests 64 entries at once, one per lane, against the box of its 8 x 8 cell centres
    // and steps only through the entries that can reach one of them (a list of at most
    // kShortList entries is stepped through as kWalkTogether does). One gather then replaces a
    // chain of dependent scalar loads per entry.
    kWalkCulled = TG_WALK_CULLED,
};

constexpr uint32_t kShortList = 4u;

// Crater n's outline harmonics from its first: its slot of the hot table, or the padded harmonic
// table when the outline has more than a slot holds.
__device__ __forceinline__ const float* Coefficients(const tg::HotCrater& c, uint32_t n,
                                                    const float* hotHarm, const float* harm) {
    return c.harmonic_count <= tg::kHotHarmonics
               ? hotHarm + size_t(n) * (2u * tg::kHotHarmonics)
               : harm + c.harmonic_offset;
}

// A km-scale crater's FP64 contribution at a plane cell, before it joins the cell's precise sum.
__device__ __forceinline__ double PlanePrecise(uint32_t n, float x, float xLow, float y,
                                               float yLow, double pixel, const tg::HotCrater* hot,
                                               const float* hotHarm, const float* harm,
                                               const tg::PreciseTerms* precise) {
    const tg::HotCrater& c = hot[n];
    const tg::Offset64 o = tg::PlaneOffset64(precise[n], x, xLow, y, yLow);
    if (!(o.distance * o.distance < double(c.cull))) {
        return 0.0;
    }
    return double(ResolutionAt(c.diameter, pixel)) *
           tg::MpdPrecise(precise[n], c.ellipticity, c.first_harmonic, c.harmonic_count,
                          Coefficients(c, n, hotHarm, harm), o);
}

// The same at a body-fixed cell point, the offset taken in the crater's frame.
__device__ __forceinline__ double BodyPrecise(uint32_t n, const tg_split_point& p, double spacing,
                                              const tg::HotCrater* hot, const float* hotHarm,
                                              const float* harm,
                                              const tg::PreciseTerms* precise) {
    const tg::HotCrater& c = hot[n];
    const tg::Offset64 o = tg::BodyOffset64(precise[n], p);
    if (!(o.distance * o.distance < double(c.cull))) {
        return 0.0;
    }
    return double(ResolutionAt(c.diameter, spacing)) *
           tg::MpdPrecise(precise[n], c.ellipticity, c.first_harmonic, c.harmonic_count,
                          Coefficients(c, n, hotHarm, harm), o);
}

// Crater n at the cell (x + xLow, y + yLow), as the strip kernel applies it: nothing beyond the
// crater's cull radius (or for a lane whose cell lies past the tile); a km-scale crater's
// contribution in FP64 into the cell's precise sum; else the hoisted MPD evaluation where the
// crater allows it and the full one, after folding the precise sum in, otherwise.
template <bool kAllFast, bool kPrecise>
__device__ __forceinline__ void VisitCrater(uint32_t n, bool inside, tg::PreciseSum& cell, float x,
                                            float xLow, float y, float yLow, double pixel,
                                            bool hoisted, const tg::HotCrater* hot,
                                            const float* hotHarm, const tg_crater_shape* kernel,
                                            const float* data, const tg_crater_record* records,
                                            const float* harm, const float4* outline,
                                            const tg::PreciseTerms* precise) {
    const tg::HotCrater& c = hot[n];
    if constexpr (kPrecise) {
        if (c.terms.fast == tg::kFastPrecise) {
            if (inside) {
                cell.pending +=
                    PlanePrecise(n, x, xLow, y, yLow, pixel, hot, hotHarm, harm, precise);
            }
            return;
        }
    }
    const float dx = tg::SplitDifference(x, xLow, c.x, c.x_low);
    const float dy = tg::SplitDifference(y, yLow, c.y, c.y_low);
    if (inside && dx * dx + dy * dy < c.cull) {
        const float weight = ResolutionAt(c.diameter, pixel);
        if (kAllFast || (hoisted && c.terms.fast != 0u)) {
            cell.h = tg::ApplyMpd(c.terms, c.ellipticity, c.first_harmonic, c.harmonic_count,
                                  Coefficients(c, n, hotHarm, harm), c.ra, cell.h, dx, dy, weight);
        } else if constexpr (!kAllFast) {
            const tg::CraterShape shape(*kernel);
            // A crater outside the hoisted path may read the height: the precise sum joins it
            // first. A hoisted one only adds, as its full evaluation here does.
            const float h = c.terms.fast == 0u ? cell.Folded() : cell.h;
            cell.h = tg::ApplyCrater(shape, data, records[n], h, dx, dy, harm, outline[n], weight);
        }
    }
}

// The boulder layer over a cell after its craters (BoulderLayer::AddRelief without a fillet,
// boulder_relief.cpp:244-262): over the bin's boulders, each drawn at the tile's `spacing` as
// BoulderLayer::PrepareRelief (boulder_relief.cpp:373-423) prepares it for that grid: its
// resolution weight at the spacing, the dome averaged over a band one spacing wide and reaching
// half of (diameter + spacing). The reach, weight and band are narrowed from FP64 as the host's
// preparation narrows them for the plain boulder kernel, whose cull test this repeats, so a tile
// here is bit-equal to its craters then its boulders through the plain kernels. `offsetOf(n, b,
// dx, dy)` gives the cell's offset east and north of boulder n. The maximum over the boulders does
// not depend on their order.
template <class OffsetOf>
__device__ __forceinline__ float DrawBoulders(float base, const tg_bin& list,
                                              const uint32_t* entries, const tg_boulder* boulders,
                                              const double* diameters, double spacing,
                                              const OffsetOf& offsetOf) {
    float h = base;
    const float width = float(spacing);
    const uint32_t end = list.first + list.count;
    for (uint32_t e = list.first; e < end; ++e) {
        const uint32_t n = entries[e];
        const double diameter = diameters[n];
        const float weight = ResolutionAt(diameter, spacing);
        if (weight == 0.0f) {
            continue;  // unresolved at this spacing (uniform across the wave)
        }
        const float reach = float(0.5 * (diameter + spacing));
        const double slack = double(reach) * (1.0 + 1e-4);
        const float cull = float(slack * slack);
        const tg_boulder& b = boulders[n];
        float dx, dy;
        offsetOf(n, b, dx, dy);
        if (dx * dx + dy * dy < cull) {
            h = tg::ApplyBoulder(b, width, weight, h, base, dx, dy);
        }
    }
    return h;
}

// Takes bins off the queues until every range is empty, calling `bin(b)` for each bin b taken:
// lane 0 takes `chunk` bins with one atomic add and the first is broadcast with readfirstlane, so
// the calls are uniform across the wave and no wave waits for another.
template <class F>
__device__ __forceinline__ void DrainBins(uint32_t* counters, uint32_t chunk, F&& bin) {
    const uint32_t binCount = counters[kCountSlot];
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t queue = blockIdx.x % kQueues;
    uint32_t drained = 0u;  // ranges found empty; a range never refills
    for (;;) {
        const uint32_t lo = uint32_t(uint64_t(binCount) * queue / kQueues);
        const uint32_t hi = uint32_t(uint64_t(binCount) * (queue + 1u) / kQueues);
        uint32_t taken = 0u;
        if (lane == 0u) {
            taken = atomicAdd(&counters[queue * kQueueStride], chunk);
        }
        const uint64_t first = uint64_t(lo) + __builtin_amdgcn_readfirstlane(taken);
        if (first >= hi) {
            if (++drained == kQueues) {
                return;  // every range is empty, so every wave ends here
            }
            queue = (queue + 1u) % kQueues;
            continue;
        }
        const uint32_t last = uint32_t(std::min<uint64_t>(first + chunk, hi));
        for (uint32_t b = uint32_t(first); b < last; ++b) {
            bin(b);
        }
    }
}

// The culled walk over entries [first, end): `depth` entries (32, 64 or 128) are tested per pass,
// one or two per lane (`test(e, n)` reads entry e's crater into n and says whether it can reach
// the wave's cells), the passing ones gathered with a ballot and visited in list order. Every depth
// visits the same craters in the same order, so it changes only the time.
// kOneSite: `visit` is called from one place, so a large visit (the full crater evaluation) is
// inlined once rather than made a function call; a small one is inlined at each of the
// depths' gathers, which keeps the depth-64 loop tighter.
template <bool kOneSite, class Test, class Visit>
__device__ __forceinline__ void CulledWalk(uint32_t first, uint32_t end, uint32_t depth,
                                           uint32_t lane, const Test& test, const Visit& visit) {
    if constexpr (!kOneSite) {
        const auto gather = [&](uint64_t candidates, uint32_t n) {
            while (candidates != 0ull) {
                const uint32_t j = uint32_t(__builtin_ctzll(candidates));
                candidates &= candidates - 1ull;
                visit(uint32_t(__builtin_amdgcn_readlane(int(n), int(j))));
            }
        };
        if (depth == 128u) {
            for (uint32_t base = first; base < end; base += 128u) {
                // Both halves' loads and tests before either ballot, for more loads in flight.
                uint32_t n0 = 0u, n1 = 0u;
                const bool near0 = test(base + lane, n0);
                const bool near1 = test(base + 64u + lane, n1);
                const uint64_t c0 = __ballot(near0);
                const uint64_t c1 = __ballot(near1);
                gather(c0, n0);
                gather(c1, n1);
            }
            return;
        }
        const uint32_t step = depth == 32u ? 32u : 64u;
        for (uint32_t base = first; base < end; base += step) {
            uint32_t n = 0u;
            const bool near = lane < step && test(base + lane, n);
            gather(__ballot(near), n);
        }
        return;
    }
    const uint32_t step = depth == 128u ? 128u : (depth == 32u ? 32u : 64u);
    for (uint32_t base = first; base < end; base += step) {
        // Both halves' loads and tests before either ballot, for more loads in flight.
        uint32_t n0 = 0u, n1 = 0u;
        const bool near0 = lane < step && test(base + lane, n0);
        const bool near1 = step == 128u && test(base + 64u + lane, n1);
        uint64_t low = __ballot(near0);
        uint64_t high = __ballot(near1);
        while ((low | high) != 0ull) {
            const bool fromLow = low != 0ull;
            const uint32_t j = uint32_t(__builtin_ctzll(fromLow ? low : high));
            const uint32_t n =
                uint32_t(__builtin_amdgcn_readlane(int(fromLow ? n0 : n1), int(j)));
            if (fromLow) {
                low &= low - 1ull;
            } else {
                high &= high - 1ull;
            }
            visit(n);
        }
    }
}

// One bin per wave: lane 0 takes the next bin with one atomic add and the index is broadcast
// with readfirstlane, so the bin loop is uniform across the wave and no wave waits for another
// (the kernel has no block barrier). The wave covers the bin as four 8 x 8 squares, the shape the
// strip kernel uses, so its crater-reach tests stay coherent across lanes.
// kAllFast: every crater of the catalogue takes the hoisted MPD path, so the full evaluation is
// not compiled in; the kernel then holds far fewer registers and keeps more waves resident.
// kPrecise: some crater takes that path in FP64 (a catalogue without km-scale craters runs the
// variant without the FP64 code, whose registers cost the other variants residency).
// kBoulders: a boulder layer follows the craters (its code only in the variants that draw it).
template <uint32_t kWalk, bool kAllFast, bool kPrecise, bool kBoulders = false>
__global__ void __launch_bounds__(kGroupThreads, kAllFast ? (kPrecise ? kPreciseWaves : 8) : TG_GENERAL_WAVES)
    BatchTiles(const tg_tile_desc* __restrict__ tiles, const tg_bin* __restrict__ bins,
               const uint32_t* __restrict__ entries, uint32_t* __restrict__ counters,
               const tg::NoisePlan* __restrict__ plan, float elevation,
               const tg_crater_shape* __restrict__ kernel,
               const tg_crater_record* __restrict__ records, const float* __restrict__ data,
               const float* __restrict__ harm, const float4* __restrict__ outline,
               const tg::HotCrater* __restrict__ hot, const float* __restrict__ hotHarm,
               const tg::PreciseTerms* __restrict__ precise,
               const tg_bin* __restrict__ boulderBins,
               const uint32_t* __restrict__ boulderEntries,
               const tg_boulder* __restrict__ boulders,
               const double* __restrict__ boulderDiameters, uint32_t hoisted, uint32_t chunk,
               uint32_t binSide, uint32_t cullDepth, float* __restrict__ out) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t squaresAcross = binSide / 8u;
    DrainBins(counters, chunk, [&](uint32_t b) __attribute__((always_inline)) {
        const tg_bin bin = bins[b];
        const tg_tile_desc tile = tiles[bin.tile];
        const uint32_t local = b - tile.bin_first;
        const uint32_t col0 = (local % tile.bins_across) * binSide;
        const uint32_t row0 = (local / tile.bins_across) * binSide;
        for (uint32_t square = 0; square < squaresAcross * squaresAcross; ++square) {
            const uint32_t squareCol = col0 + (square % squaresAcross) * 8u;
            const uint32_t squareRow = row0 + (square / squaresAcross) * 8u;
            if (squareCol >= tile.width || squareRow >= tile.height) {
                continue;  // the square lies past the tile's edge (uniform across the wave)
            }
            const uint32_t col = squareCol + (lane & 7u);
            const uint32_t row = squareRow + (lane >> 3u);
            const bool inside = col < tile.width && row < tile.height;
            // Lanes past the tile's edge evaluate its last column or row and store nothing, so
            // the whole wave takes part in the noise's shared corner hashing and in the walks'
            // ballots.
            const uint32_t cellCol = std::min(col, tile.width - 1u);
            const uint32_t cellRow = std::min(row, tile.height - 1u);
            // GeoTransform::PixelToWorld of the cell centre in FP64, narrowed and split as
            // ColumnOffsets/RowOffsets and MakeSplitPoint do (src/gpu/kernel.cpp:55-70,
            // crater_kernel.cpp:229-236): the source origin is {0, 0}.
            const double cx = tile.origin_x + (double(cellCol) + 0.5) * tile.pixel + 0.0;
            const double cy = tile.origin_y + (double(cellRow) + 0.5) * -tile.pixel;
            const float x = float(cx);
            const float y = float(cy);
            tg::PreciseSum cell{plan != nullptr
                                    ? elevation + tg::EvaluateStack<true>(*plan, x, y, lane)
                                    : elevation};
            if (kernel != nullptr && bin.count > 0u) {
                const float xLow = float(cx - double(x));
                const float yLow = float(cy - double(y));
                const uint32_t end = bin.first + bin.count;
                const auto visit = [&](uint32_t n) __attribute__((always_inline)) {
                    VisitCrater<kAllFast, kPrecise>(n, inside, cell, x, xLow, y, yLow, tile.pixel,
                                          hoisted != 0u, hot, hotHarm, kernel, data, records, harm,
                                          outline, precise);
                };
                if constexpr (kWalk == kWalkPerLane) {
                    uint32_t e = bin.first;
                    for (;;) {
                        uint32_t n = 0u;
                        bool hit = false;
                        while (inside && e < end) {
                            const uint32_t m = entries[e];
                            ++e;
                            const tg::HotCrater& c = hot[m];
                            const float ex = tg::SplitDifference(x, xLow, c.x, c.x_low);
                            const float ey = tg::SplitDifference(y, yLow, c.y, c.y_low);
                            if (ex * ex + ey * ey < c.cull) {
                                n = m;
                                hit = true;
                                break;
                            }
                        }
                        if (__ballot(hit) == 0ull) {
                            break;  // no lane of the wave has a crater left
                        }
                        if (hit) {
                            visit(n);
                        }
                    }
                } else if constexpr (kWalk == kWalkTogether) {
                    for (uint32_t e = bin.first; e < end; ++e) {
                        visit(entries[e]);
                    }
                } else if (kAllFast && bin.count <= kShortList) {
                    for (uint32_t e = bin.first; e < end; ++e) {
                        visit(entries[e]);
                    }
                } else {
                    // Without kAllFast a short list is visited whole through the same walk, every
                    // entry passing, in the order the loop above visits (the visit's own test
                    // decides what each cell takes), so the full evaluation is inlined once.
                    const bool whole = bin.count <= kShortList;
                    // The box of the square's cell centres (the same FP64 centres as above).
                    const uint32_t lastCol = std::min(squareCol + 7u, tile.width - 1u);
                    const uint32_t lastRow = std::min(squareRow + 7u, tile.height - 1u);
                    const double west = tile.origin_x + (double(squareCol) + 0.5) * tile.pixel + 0.0;
                    const double east = tile.origin_x + (double(lastCol) + 0.5) * tile.pixel + 0.0;
                    const double north = tile.origin_y + (double(squareRow) + 0.5) * -tile.pixel;
                    const double south = tile.origin_y + (double(lastRow) + 0.5) * -tile.pixel;
                    // Entry e of the list near the square's box. The centre's split sum is exact
                    // in FP64. A cell passing the FP32 test of the visit lies within the cull
                    // radius to a relative 1e-6 or so, and the box is no farther than any of its
                    // cells, so the margin keeps every entry a cell of the square can take.
                    const auto test = [&](uint32_t e, uint32_t& n) {
                        if (e >= end) {
                            return false;
                        }
                        n = entries[e];
                        if (whole) {
                            return true;
                        }
                        const tg::HotCrater& c = hot[n];
                        const double px = double(c.x) + double(c.x_low);
                        const double py = double(c.y) + double(c.y_low);
                        const double ex = fmax(0.0, fmax(west - px, px - east));
                        const double ey = fmax(0.0, fmax(south - py, py - north));
                        return ex * ex + ey * ey <= double(c.cull) * (1.0 + 1e-4) + 1e-6;
                    };
                    CulledWalk<!kAllFast>(bin.first, end, whole ? 64u : cullDepth, lane, test,
                                          visit);
                }
            }
            float height = cell.Folded();
            if constexpr (kBoulders) {
                const tg_bin list = boulderBins[b];
                if (list.count > 0u) {
                    const float xLow = float(cx - double(x));
                    const float yLow = float(cy - double(y));
                    height = DrawBoulders(
                        height, list, boulderEntries, boulders, boulderDiameters, tile.pixel,
                        [&](uint32_t, const tg_boulder& r, float& dx, float& dy) {
                            dx = (x - r.x) + (xLow - r.x_low);
                            dy = (y - r.y) + (yLow - r.y_low);
                        });
                }
            }
            if (inside) {
                out[tile.out_offset + size_t(row) * tile.width + col] = height;
            }
        }
    });
}

// The largest of v over the wave (every lane active).
__device__ __forceinline__ double WaveMax(double v) {
    for (int offset = 32; offset > 0; offset >>= 1) {
        v = fmax(v, __shfl_xor(v, offset, 64));
    }
    return v;
}

// Crater n at the body-fixed cell point p, as the strip kernel applies it on a body: the offset in
// the crater's frame (BodySplitOffset), nothing beyond the crater's cull radius (or for a lane
// whose cell lies past the tile); a km-scale crater's contribution in FP64 (offset included) into
// the cell's precise sum; else the hoisted MPD evaluation where the crater allows it and the full
// one otherwise, at the tile's resolution weight.
template <bool kAllFast, bool kPrecise>
__device__ __forceinline__ void VisitBodyCrater(uint32_t n, bool inside, tg::PreciseSum& cell,
                                                const tg_split_point& p, double spacing,
                                                bool hoisted, const tg::HotCrater* hot,
                                                const float* hotHarm,
                                                const tg_crater_frame* frames,
                                                const tg_crater_shape* kernel, const float* data,
                                                const tg_crater_record* records,
                                                const float* harm, const float4* outline,
                                                const tg::PreciseTerms* precise) {
    const tg::HotCrater& c = hot[n];
    if constexpr (kPrecise) {
        if (c.terms.fast == tg::kFastPrecise) {
            if (inside) {
                cell.pending += BodyPrecise(n, p, spacing, hot, hotHarm, harm, precise);
            }
            return;
        }
    }
    const tg::Offset at = tg::BodySplitOffset(frames[n], p);
    if (inside && at.dx * at.dx + at.dy * at.dy < c.cull) {
        const float weight = ResolutionAt(c.diameter, spacing);
        if (kAllFast || (hoisted && c.terms.fast != 0u)) {
            cell.h = tg::ApplyMpd(c.terms, c.ellipticity, c.first_harmonic, c.harmonic_count,
                                  Coefficients(c, n, hotHarm, harm), c.ra, cell.h, at.dx, at.dy,
                                  weight);
        } else if constexpr (!kAllFast) {
            const tg::CraterShape shape(*kernel);
            const float h = c.terms.fast == 0u ? cell.Folded() : cell.h;
            cell.h = tg::ApplyCrater(shape, data, records[n], h, at.dx, at.dy, harm, outline[n],
                                     weight);
        }
    }
}

// BatchTiles for longitude-latitude tiles on a sphere of `radius` meters (a tile's origin and
// pixel in degrees). Each cell's body-fixed position is computed in FP64 (Surface::Position,
// src/geometry/surface.cpp:45-50,261-267, with sphere_math.h's sine and cosine), the noise is the
// stack's at that position (EvaluateFractals3, src/noise/surface_noise.cpp:9-20), and the craters
// take the position split into FP32 pairs (MakeSplitPoint, crater_kernel.cpp:303-310) and their
// offsets in each crater's frame (craters.h:168-209), at the resolution weight of the tile's
// spacing, pixel times unitLength (RasterGrid::Spacing, src/raster/grid.cpp:28-33). The crater
// lists are walked 64 entries at a time: an entry is visited only when its centre lies within its
// cull radius (a bound on the arc, and so on the chord, to any cell it shapes) of the ball around
// the wave's 8 x 8 cells.
// kSmall: every lattice coordinate of the launch lies within 2^31 of the origin (checked on the
// host from the radius and the plan's highest frequency).
// kGeneral: the plan holds warped fractals or bases other than Perlin (one variant, with the
// general crater walk and 64-bit lattice indices, carries their code).
// kPrecise, kBoulders: as for BatchTiles.
template <bool kAllFast, bool kSmall, bool kGeneral, bool kPrecise, bool kBoulders = false>
__global__ void __launch_bounds__(kGroupThreads, kAllFast ? (kPrecise ? kPreciseWaves : 8) : TG_GENERAL_WAVES)
    BatchBodyTiles(const tg_tile_desc* __restrict__ tiles, const tg_bin* __restrict__ bins,
                   const uint32_t* __restrict__ entries, uint32_t* __restrict__ counters,
                   const tg::NoisePlan* __restrict__ plan, float elevation, double radius,
                   double unitLength, const tg_crater_shape* __restrict__ kernel,
                   const tg_crater_record* __restrict__ records, const float* __restrict__ data,
                   const float* __restrict__ harm, const float4* __restrict__ outline,
                   const tg::HotCrater* __restrict__ hot, const float* __restrict__ hotHarm,
                   const tg_crater_frame* __restrict__ frames,
                   const tg::PreciseTerms* __restrict__ precise,
                   const tg_bin* __restrict__ boulderBins,
                   const uint32_t* __restrict__ boulderEntries,
                   const tg_boulder* __restrict__ boulders,
                   const double* __restrict__ boulderDiameters,
                   const tg_crater_frame* __restrict__ boulderFrames, uint32_t hoisted,
                   uint32_t chunk, uint32_t binSide, uint32_t cullDepth,
                   float* __restrict__ out) {
    const uint32_t lane = threadIdx.x & 63u;
    const tg::CubeSteps steps = tg::MakeCubeSteps(lane);
    const uint32_t squaresAcross = binSide / 8u;
    DrainBins(counters, chunk, [&](uint32_t b) __attribute__((always_inline)) {
        const tg_bin bin = bins[b];
        const tg_tile_desc tile = tiles[bin.tile];
        const uint32_t local = b - tile.bin_first;
        const uint32_t col0 = (local % tile.bins_across) * binSide;
        const uint32_t row0 = (local / tile.bins_across) * binSide;
        const double spacing = tile.pixel * unitLength;
        for (uint32_t square = 0; square < squaresAcross * squaresAcross; ++square) {
            const uint32_t squareCol = col0 + (square % squaresAcross) * 8u;
            const uint32_t squareRow = row0 + (square / squaresAcross) * 8u;
            if (squareCol >= tile.width || squareRow >= tile.height) {
                continue;  // the square lies past the tile's edge (uniform across the wave)
            }
            const uint32_t col = squareCol + (lane & 7u);
            const uint32_t row = squareRow + (lane >> 3u);
            const bool inside = col < tile.width && row < tile.height;
            // Lanes past the tile's edge evaluate its last column or row and store nothing.
            // A cell's longitude depends on its column and its latitude on its row, so lanes 0-7
            // take the sine and cosine of the square's columns, lanes 8-15 of its rows, and every
            // lane gathers its own: the position is CellPosition's, bit for bit.
            const uint32_t which = lane & 15u;
            const double angle =
                which < 8u
                    ? tile.origin_x +
                          (double(std::min(squareCol + which, tile.width - 1u)) + 0.5) *
                              tile.pixel +
                          0.0
                    : tile.origin_y +
                          (double(std::min(squareRow + which - 8u, tile.height - 1u)) + 0.5) *
                              -tile.pixel;
            const tg::SinCos mine = tg::SinCosDegrees(angle);
            const tg::SinCos lon{__shfl(mine.sin, int(lane & 7u), 64),
                                 __shfl(mine.cos, int(lane & 7u), 64)};
            const tg::SinCos lat{__shfl(mine.sin, int(8u + (lane >> 3u)), 64),
                                 __shfl(mine.cos, int(8u + (lane >> 3u)), 64)};
            const tg::Position at = tg::SpherePositionOf(lon, lat, radius);
            tg::PreciseSum cell{plan != nullptr
                                    ? elevation + tg::EvaluateStack3<true, kSmall, kGeneral>(
                                                      *plan, at.x, at.y, at.z, lane, steps)
                                    : elevation};
            // MakeSplitPoint of the cell's position (crater_kernel.cpp:303-310).
            const auto split = [&] {
                tg_split_point p;
                p.x = float(at.x);
                p.x_low = float(at.x - double(p.x));
                p.y = float(at.y);
                p.y_low = float(at.y - double(p.y));
                p.z = float(at.z);
                p.z_low = float(at.z - double(p.z));
                return p;
            };
            if (kernel != nullptr && bin.count > 0u) {
                const tg_split_point p = split();
                const uint32_t end = bin.first + bin.count;
                const auto visit = [&](uint32_t n) __attribute__((always_inline)) {
                    VisitBodyCrater<kAllFast, kPrecise>(n, inside, cell, p, spacing, hoisted != 0u, hot,
                                              hotHarm, frames, kernel, data, records, harm,
                                              outline, precise);
                };
                if (kAllFast && bin.count <= kShortList) {
                    for (uint32_t e = bin.first; e < end; ++e) {
                        visit(entries[e]);
                    }
                } else {
                    // Without kAllFast a short list is visited whole through the same walk, every
                    // entry passing, in the order the loop above visits (the visit's own test
                    // decides what each cell takes), so the full evaluation is inlined once.
                    const bool whole = bin.count <= kShortList;
                    // The ball around the square's cells: centred on lane 27's cell (row 3,
                    // column 3), its radius the farthest of the wave's cells from it (only for
                    // a list the walk tests).
                    double ox = 0.0, oy = 0.0, oz = 0.0, reach = 0.0;
                    if (!whole) {
                        ox = __shfl(at.x, 27, 64);
                        oy = __shfl(at.y, 27, 64);
                        oz = __shfl(at.z, 27, 64);
                        reach = WaveMax(sqrt((at.x - ox) * (at.x - ox) +
                                             (at.y - oy) * (at.y - oy) +
                                             (at.z - oz) * (at.z - oz)));
                    }
                    // Entry e of the list near the ball. A cell passing the FP32 test lies within
                    // the cull radius of the centre along the sphere to a relative 1e-6 or so,
                    // and the chord is no longer than the arc; the margins keep every entry a
                    // cell of the square can take.
                    const auto test = [&](uint32_t e, uint32_t& n) {
                        if (e >= end) {
                            return false;
                        }
                        n = entries[e];
                        if (whole) {
                            return true;
                        }
                        const tg_crater_frame& f = frames[n];
                        const double qx = double(f.x) + double(f.x_low) - ox;
                        const double qy = double(f.y) + double(f.y_low) - oy;
                        const double qz = double(f.z) + double(f.z_low) - oz;
                        const double d = sqrt(qx * qx + qy * qy + qz * qz);
                        return d <= sqrt(double(hot[n].cull)) * (1.0 + 1e-4) + 1e-3 + reach;
                    };
                    CulledWalk<!kAllFast>(bin.first, end, whole ? 64u : cullDepth, lane, test,
                                          visit);
                }
            }
            float height = cell.Folded();
            if constexpr (kBoulders) {
                const tg_bin list = boulderBins[b];
                if (list.count > 0u) {
                    // The offset in the boulder's frame: its east and north at its centre, the
                    // tangent offset rescaled to the arc (Surface::SphereOffset,
                    // surface.cpp:214-229, to FP32 rounding at a boulder's few meters).
                    const tg_split_point p = split();
                    height = DrawBoulders(
                        height, list, boulderEntries, boulders, boulderDiameters, spacing,
                        [&](uint32_t n, const tg_boulder&, float& dx, float& dy) {
                            const tg::Offset o = tg::BodySplitOffset(boulderFrames[n], p);
                            dx = o.dx;
                            dy = o.dy;
                        });
                }
            }
            if (inside) {
                out[tile.out_offset + size_t(row) * tile.width + col] = height;
            }
        }
    });
}

// One lane per crater: the hot record and harmonic slot from the catalogue's arrays.
__global__ void MakeHotCraters(const tg_crater_record* __restrict__ records,
                               const float* __restrict__ cull, const double* __restrict__ diameters,
                               const tg::MpdTerms* __restrict__ terms,
                               const float4* __restrict__ outline, const float* __restrict__ harm,
                               size_t count, tg::HotCrater* __restrict__ hot,
                               float* __restrict__ hotHarm) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= count) {
        return;
    }
    const tg_crater_record& c = records[i];
    tg::HotCrater h;
    h.x = c.x;
    h.y = c.y;
    h.x_low = c.x_low;
    h.y_low = c.y_low;
    h.cull = cull[i];
    h.ellipticity = c.ellipticity;
    h.first_harmonic = c.first_harmonic;
    h.harmonic_count = c.harmonic_count;
    h.diameter = diameters[i];
    h.harmonic_offset = c.harmonic_offset;
    h.pad = 0u;
    h.ra = outline[i];
    h.terms = terms[i];
    hot[i] = h;
    float* slot = hotHarm + i * (2u * tg::kHotHarmonics);
    const uint32_t held = 2u * std::min(c.harmonic_count, tg::kHotHarmonics);
    for (uint32_t k = 0; k < 2u * tg::kHotHarmonics; ++k) {
        slot[k] = k < held ? harm[c.harmonic_offset + k] : 0.0f;
    }
}

template <class T> T* Alloc(size_t count) {
    void* p = nullptr;
    Check(hipMalloc(&p, (count > 0 ? count : 1) * sizeof(T)), "hipMalloc");
    return static_cast<T*>(p);
}

template <class T> T* UploadArray(const T* src, size_t count) {
    T* p = Alloc<T>(count);
    if (src != nullptr && count > 0) {
        const hipError_t status = hipMemcpy(p, src, count * sizeof(T), hipMemcpyHostToDevice);
        if (status != hipSuccess) {
            (void)hipFree(p);
            Check(status, "hipMemcpy to device");
        }
    } else {
        const hipError_t status = hipMemset(p, 0, (count > 0 ? count : 1) * sizeof(T));
        if (status != hipSuccess) {
            (void)hipFree(p);
            Check(status, "hipMemset");
        }
    }
    return p;
}

void FreeCatalogue(tg_catalogue* c) {
    for (void* p : {static_cast<void*>(c->kernel), static_cast<void*>(c->records),
                    static_cast<void*>(c->data), static_cast<void*>(c->cull),
                    static_cast<void*>(c->harm), static_cast<void*>(c->outline),
                    static_cast<void*>(c->diameters), static_cast<void*>(c->terms),
                    static_cast<void*>(c->hot), static_cast<void*>(c->hotHarm),
                    static_cast<void*>(c->frames), static_cast<void*>(c->precise)}) {
        if (p != nullptr) {
            (void)hipFree(p);
        }
    }
}

// One wave per bin up to a few workgroups per CU: every CU stays busy and each wave keeps
// taking bins until the queue is empty.
uint32_t GroupsFor(const tg_batch& b) {
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b.dev->ordinal);
    const size_t wavesPerGroup = kGroupThreads / 64;
    return uint32_t(std::max<size_t>(
        1, std::min<size_t>((b.binCount + wavesPerGroup - 1) / wavesPerGroup,
                            size_t(cus) * b.groupsPerCu)));
}

// The bins a wave takes per atomic add: the tuned chunk, cut down on small batches so that each
// wave gets at least four takes (a batch of one tile then spreads over as many waves as it has
// bins, which keeps its latency low).
uint32_t ChunkFor(const tg_batch& b, uint32_t groups) {
    const uint64_t waves = uint64_t(groups) * (kGroupThreads / 64);
    return uint32_t(std::clamp<uint64_t>(b.binCount / (waves * 4u), 1u, b.chunk));
}

using PlaneKernelFn = decltype(&BatchTiles<kWalkCulled, true, false>);
using BodyKernelFn = decltype(&BatchBodyTiles<true, true, false, false>);

template <uint32_t kWalk>
PlaneKernelFn PlaneKernelOf(bool allFast, bool precise) {
    return allFast ? (precise ? BatchTiles<kWalk, true, true> : BatchTiles<kWalk, true, false>)
                   : (precise ? BatchTiles<kWalk, false, true> : BatchTiles<kWalk, false, false>);
}

// With boulders: the culled walk and the full crater evaluation (which takes the hoisted path
// crater by crater), so that only two variants carry the boulder code.
PlaneKernelFn PlaneKernel(uint32_t walk, bool allFast, bool precise, bool boulders) {
    if (boulders) {
        return precise ? BatchTiles<kWalkCulled, false, true, true>
                       : BatchTiles<kWalkCulled, false, false, true>;
    }
    return walk == TG_WALK_PER_LANE ? PlaneKernelOf<kWalkPerLane>(allFast, precise)
           : walk == TG_WALK_CULLED ? PlaneKernelOf<kWalkCulled>(allFast, precise)
                                    : PlaneKernelOf<kWalkTogether>(allFast, precise);
}

template <bool kPrecise>
BodyKernelFn BodyKernelOf(bool allFast, bool small, bool general, bool boulders) {
    if (boulders) {  // the full crater evaluation and 64-bit lattice indices, as kGeneral's
        return general ? BatchBodyTiles<false, false, true, kPrecise, true>
                       : BatchBodyTiles<false, false, false, kPrecise, true>;
    }
    if (general) {  // other bases or warped fractals: one variant carries their code
        return BatchBodyTiles<false, false, true, kPrecise>;
    }
    return allFast ? (small ? BatchBodyTiles<true, true, false, kPrecise>
                            : BatchBodyTiles<true, false, false, kPrecise>)
                   : (small ? BatchBodyTiles<false, true, false, kPrecise>
                            : BatchBodyTiles<false, false, false, kPrecise>);
}

BodyKernelFn BodyKernel(bool allFast, bool small, bool general, bool precise, bool boulders) {
    return precise ? BodyKernelOf<true>(allFast, small, general, boulders)
                   : BodyKernelOf<false>(allFast, small, general, boulders);
}

void Launch(tg_batch& b, const tg_noise_plan* plan, float elevation, const tg_catalogue* cat,
            float* out) {
    const tg_boulder_catalogue* rocks = b.boulders;
    const tg_bin* boulderBins = rocks != nullptr ? b.boulderBins : nullptr;
    const uint32_t* boulderEntries = rocks != nullptr ? b.boulderEntries : nullptr;
    const tg_boulder* boulderRecords = rocks != nullptr ? rocks->records : nullptr;
    const double* boulderDiameters = rocks != nullptr ? rocks->diameters : nullptr;
    const uint32_t groups = GroupsFor(b);
    const uint32_t chunk = ChunkFor(b, groups);
    hipLaunchKernelGGL(ResetQueue, dim3(1), dim3(64), 0, b.stream, b.counters);
    // The compact kernel when every crater takes the hoisted path (or there are none).
    const bool allFast = b.hoisted && (cat == nullptr || cat->allFast);
    if (b.radius > 0.0) {
        // The length of a degree of arc, as Surface::Sphere sets it (surface.cpp:49).
        const double unitLength = b.radius * (3.141592653589793 / 180.0);
        // Cell positions lie on the sphere of the batch's radius.
        const bool small =
            plan == nullptr || tg::BodyLatticeBound(plan->host, b.radius) < 2147483648.0;
        const bool precise = cat != nullptr && cat->anyPrecise;
        auto* body = BodyKernel(allFast, small, plan != nullptr && !plan->host.bodyPerlin,
                                precise, rocks != nullptr);
        hipLaunchKernelGGL(body, dim3(groups), dim3(kGroupThreads), 0, b.stream, b.tiles, b.bins,
                           b.entries, b.counters, plan != nullptr ? plan->device : nullptr,
                           elevation, b.radius, unitLength,
                           cat != nullptr ? cat->kernel : nullptr,
                           cat != nullptr ? cat->records : nullptr,
                           cat != nullptr ? cat->data : nullptr,
                           cat != nullptr ? cat->harm : nullptr,
                           cat != nullptr ? cat->outline : nullptr,
                           cat != nullptr ? cat->hot : nullptr,
                           cat != nullptr ? cat->hotHarm : nullptr,
                           cat != nullptr ? cat->frames : nullptr,
                           cat != nullptr ? cat->precise : nullptr, boulderBins, boulderEntries,
                           boulderRecords, boulderDiameters,
                           rocks != nullptr ? rocks->frames : nullptr, b.hoisted ? 1u : 0u, chunk,
                           b.binSide, b.cullDepth, out);
        Check(hipGetLastError(), "launching the batched body tile kernel");
        return;
    }
    auto* fn = PlaneKernel(b.walk, allFast, cat != nullptr && cat->anyPrecise, rocks != nullptr);
    hipLaunchKernelGGL(fn, dim3(groups), dim3(kGroupThreads), 0, b.stream, b.tiles, b.bins,
                       b.entries, b.counters, plan != nullptr ? plan->device : nullptr,
                       elevation, cat != nullptr ? cat->kernel : nullptr,
                       cat != nullptr ? cat->records : nullptr,
                       cat != nullptr ? cat->data : nullptr, cat != nullptr ? cat->harm : nullptr,
                       cat != nullptr ? cat->outline : nullptr,
                       cat != nullptr ? cat->hot : nullptr,
                       cat != nullptr ? cat->hotHarm : nullptr,
                       cat != nullptr ? cat->precise : nullptr, boulderBins, boulderEntries,
                       boulderRecords, boulderDiameters, b.hoisted ? 1u : 0u, chunk, b.binSide,
                       b.cullDepth, out);
    Check(hipGetLastError(), "launching the batched tile kernel");
}

void DropGraph(tg_batch& b) {
    if (b.graph != nullptr) {
        (void)hipGraphExecDestroy(b.graph);
        b.graph = nullptr;
    }
}

}  // namespace

extern "C" {

int tg_hip_catalogue_create(tg_dev* dev, const tg_crater_shape* kernel,
                            const tg_crater_record* records, size_t count, const float* data,
                            size_t dataCount, const double* diameters, tg_catalogue** out,
                            char* err, size_t errLen) {
    if (dev == nullptr || kernel == nullptr || out == nullptr ||
        (count > 0 && (records == nullptr || diameters == nullptr))) {
        SetErr(err, errLen, "tg_hip_catalogue_create: null argument");
        return TG_ERR_ARGUMENT;
    }
    *out = nullptr;
    auto cat = std::make_unique<tg_catalogue>();
    const int code = Guard(err, errLen, [&] {
        // The same data checks as a crater set, with one bin holding every crater.
        std::vector<uint32_t> tiles{1u, kBinSide, 0u, 0u, uint32_t(count)};
        std::vector<uint32_t> all(count > 0 ? count : 1, 0u);
        for (size_t i = 0; i < count; ++i) {
            all[i] = uint32_t(i);
        }
        ValidateCraterBatch(*kernel, records, count, dataCount, tiles.data(), tiles.size(),
                            all.data(), count > 0 ? count : 1, 0);
        Check(hipSetDevice(dev->ordinal), "hipSetDevice");
        cat->dev = dev;
        cat->count = count;
        cat->kernel = UploadArray(kernel, 1);
        cat->records = UploadArray(records, count);
        cat->data = UploadArray(data, dataCount);
        const std::vector<float> cull = CullRadii(*kernel, records, count, data, dataCount);
        cat->cull = UploadArray(cull.data(), cull.size());
        std::vector<float> harm;
        std::vector<float4> outline;
        OutlineTerms(records, count, data, dataCount, harm, outline);
        cat->harm = UploadArray(harm.data(), harm.size());
        cat->outline = UploadArray(outline.data(), outline.size());
        cat->diameters = UploadArray(diameters, count);
        cat->terms = DeviceMpdTerms(cat->kernel, cat->records, count);
        cat->precise = DevicePreciseTerms(cat->kernel, cat->records, nullptr, count);
        cat->allFast = AllMpdFast(*kernel, records, count);
        cat->anyPrecise = AnyPrecise(*kernel, records, count);
        cat->hot = Alloc<tg::HotCrater>(count);
        cat->hotHarm = Alloc<float>(count * 2u * tg::kHotHarmonics);
        if (count > 0) {
            const uint32_t block = 256;
            hipLaunchKernelGGL(MakeHotCraters, dim3(uint32_t((count + block - 1) / block)),
                               dim3(block), 0, nullptr, cat->records, cat->cull, cat->diameters,
                               cat->terms, cat->outline, cat->harm, count, cat->hot,
                               cat->hotHarm);
            Check(hipGetLastError(), "launching the hot crater records");
            Check(hipDeviceSynchronize(), "building the hot crater records");
        }
    });
    if (code != TG_OK) {
        FreeCatalogue(cat.get());
        return code;
    }
    *out = cat.release();
    return TG_OK;
}

int tg_hip_catalogue_free(tg_catalogue* catalogue) {
    if (catalogue == nullptr) {
        return TG_OK;
    }
    (void)hipSetDevice(catalogue->dev->ordinal);
    FreeCatalogue(catalogue);
    delete catalogue;
    return TG_OK;
}

int tg_hip_batch_create(tg_dev* dev, size_t maxTiles, size_t maxBins, size_t maxEntries,
                        tg_batch** out, char* err, size_t errLen) {
    if (dev == nullptr || out == nullptr) {
        SetErr(err, errLen, "tg_hip_batch_create: null argument");
        return TG_ERR_ARGUMENT;
    }
    if (maxBins > 0xFFFFFFFFull || maxEntries > 0xFFFFFFFFull || maxTiles > 0xFFFFFFFFull) {
        SetErr(err, errLen, "tg_hip_batch_create: capacities must fit 32-bit indices");
        return TG_ERR_ARGUMENT;
    }
    *out = nullptr;
    auto b = std::make_unique<tg_batch>();
    const int code = Guard(err, errLen, [&] {
        Check(hipSetDevice(dev->ordinal), "hipSetDevice");
        b->dev = dev;
        Check(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "hipStreamCreate");
        b->tiles = Alloc<tg_tile_desc>(maxTiles);
        b->bins = Alloc<tg_bin>(maxBins);
        b->entries = Alloc<uint32_t>(maxEntries);
        b->counters = UploadArray<uint32_t>(nullptr, kCounterWords);
        void* counts = nullptr;
        Check(hipHostMalloc(&counts, kCounterWords * sizeof(uint32_t), hipHostMallocDefault),
              "hipHostMalloc");
        b->hostCounts = static_cast<uint32_t*>(counts);
        std::memset(b->hostCounts, 0, kCounterWords * sizeof(uint32_t));
        b->maxTiles = maxTiles;
        b->maxBins = maxBins;
        b->maxEntries = maxEntries;
    });
    if (code != TG_OK) {
        tg_hip_batch_free(b.release());
        return code;
    }
    *out = b.release();
    return TG_OK;
}

int tg_hip_batch_free(tg_batch* b) {
    if (b == nullptr) {
        return TG_OK;
    }
    if (b->dev != nullptr) {
        (void)hipSetDevice(b->dev->ordinal);
    }
    if (b->stream != nullptr) {
        (void)hipStreamSynchronize(b->stream);
    }
    DropGraph(*b);
    for (void* p : {static_cast<void*>(b->tiles), static_cast<void*>(b->bins),
                    static_cast<void*>(b->entries), static_cast<void*>(b->counters),
                    static_cast<void*>(b->windows), static_cast<void*>(b->tileWindows),
                    static_cast<void*>(b->rowFirsts),
                    static_cast<void*>(b->binCounts), static_cast<void*>(b->binFirsts),
                    b->scanTemp, static_cast<void*>(b->boulderBins),
                    static_cast<void*>(b->boulderEntries)}) {
        if (p != nullptr) {
            (void)hipFree(p);
        }
    }
    if (b->hostCounts != nullptr) {
        (void)hipHostFree(b->hostCounts);
    }
    if (b->stream != nullptr) {
        (void)hipStreamDestroy(b->stream);
    }
    delete b;
    return TG_OK;
}

}  // extern "C"

namespace {

// What a run needs to know of a checked batch: the end of its output, its largest entry and
// whether any bin lists a crater.
struct Checked {
    uint64_t outEnd = 0;
    uint64_t maxEntry = 0;
    bool any = false;
};

// Checks every tile against the bin list's length: its shape, its bins for bins of `side` cells,
// and its output; the end of the output.
uint64_t CheckTiles(const tg_tile_desc* tiles, size_t tileCount, size_t binCount, uint32_t side) {
    uint64_t outEnd = 0;
    for (size_t i = 0; i < tileCount; ++i) {
        const tg_tile_desc& t = tiles[i];
        const uint64_t cells = uint64_t(t.width) * t.height;
        if (t.width == 0 || t.height == 0 ||
            t.bins_across != (t.width + side - 1) / side || !(t.pixel > 0.0) ||
            !std::isfinite(t.origin_x) || !std::isfinite(t.origin_y) ||
            t.out_offset > ~uint64_t(0) - cells) {
            throw HipError{"tile " + std::to_string(i) + " is malformed"};
        }
        const uint64_t binsDown = (t.height + side - 1) / side;
        if (uint64_t(t.bin_first) + uint64_t(t.bins_across) * binsDown > binCount) {
            throw HipError{"tile " + std::to_string(i) + "'s bins run past the bin list"};
        }
        outEnd = std::max<uint64_t>(outEnd, t.out_offset + cells);
    }
    return outEnd;
}

// Checks every index of a batch's tiles, bins and entries against the others.
Checked CheckBatch(const tg_tile_desc* tiles, size_t tileCount, const tg_bin* bins,
                   size_t binCount, const uint32_t* entries, size_t entryCount) {
    const uint64_t outEnd = CheckTiles(tiles, tileCount, binCount, kBinSide);
    uint64_t maxEntry = 0;
    bool any = false;
    for (size_t i = 0; i < binCount; ++i) {
        const tg_bin& bin = bins[i];
        if (bin.tile >= tileCount || uint64_t(bin.first) + bin.count > entryCount) {
            throw HipError{"bin " + std::to_string(i) + " is malformed"};
        }
        const tg_tile_desc& t = tiles[bin.tile];
        const uint64_t binsDown = (t.height + kBinSide - 1) / kBinSide;
        if (i < t.bin_first || i >= t.bin_first + uint64_t(t.bins_across) * binsDown) {
            throw HipError{"bin " + std::to_string(i) + " lies outside its tile's bins"};
        }
        any = any || bin.count > 0;
    }
    for (size_t i = 0; i < entryCount; ++i) {
        maxEntry = std::max<uint64_t>(maxEntry, entries[i]);
    }
    return {outEnd, maxEntry, entryCount > 0 && any};
}

// Keeps each staged tile's bins on the host, for checking boulder lists against them, and drops
// the boulder lists of the previous staging.
void KeepTileBins(tg_batch& b, const tg_tile_desc* tiles, size_t tileCount) {
    b.tileBinFirst.resize(tileCount);
    b.tileBinCount.resize(tileCount);
    for (size_t t = 0; t < tileCount; ++t) {
        b.tileBinFirst[t] = tiles[t].bin_first;
        b.tileBinCount[t] = tiles[t].bins_across * ((tiles[t].height + b.binSide - 1) / b.binSide);
    }
    b.boulders = nullptr;
}

// Validates a staging and copies it to the batch: synchronously from any host memory, or, with
// `async`, queued on the batch's stream from memory the caller keeps unchanged until the batch's
// next run has completed.
int Stage(tg_batch* b, const tg_tile_desc* tiles, size_t tileCount, const tg_bin* bins,
          size_t binCount, const uint32_t* entries, size_t entryCount, bool async, char* err,
          size_t errLen) {
    if (b == nullptr || (tileCount > 0 && tiles == nullptr) || (binCount > 0 && bins == nullptr) ||
        (entryCount > 0 && entries == nullptr)) {
        SetErr(err, errLen, "staging a batch: null argument");
        return TG_ERR_ARGUMENT;
    }
    return Guard(err, errLen, [&] {
        if (tileCount > b->maxTiles || binCount > b->maxBins || entryCount > b->maxEntries) {
            throw HipError{"batch exceeds the capacities it was created with"};
        }
        const Checked checked = CheckBatch(tiles, tileCount, bins, binCount, entries, entryCount);
        Check(hipSetDevice(b->dev->ordinal), "hipSetDevice");
        if (async) {
            // The pinned counts are rewritten here, so no earlier staging's copy may be pending.
            const hipError_t idle = hipStreamQuery(b->stream);
            if (idle == hipErrorNotReady) {
                throw HipError{"the batch is still running: wait for it before staging again"};
            }
            Check(idle, "hipStreamQuery");
            const auto put = [&](void* dst, const void* src, size_t bytes, const char* what) {
                if (bytes > 0) {
                    Check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, b->stream), what);
                }
            };
            put(b->tiles, tiles, tileCount * sizeof(tg_tile_desc), "staging tiles");
            put(b->bins, bins, binCount * sizeof(tg_bin), "staging bins");
            put(b->entries, entries, entryCount * sizeof(uint32_t), "staging entries");
            std::memset(b->hostCounts, 0, kCounterWords * sizeof(uint32_t));
            b->hostCounts[kCountSlot] = uint32_t(binCount);
            put(b->counters, b->hostCounts, kCounterWords * sizeof(uint32_t),
                "staging the bin count");
        } else {
            Check(hipStreamSynchronize(b->stream), "waiting for the batch's previous run");
            if (tileCount > 0) {
                Check(hipMemcpy(b->tiles, tiles, tileCount * sizeof(tg_tile_desc),
                                hipMemcpyHostToDevice),
                      "staging tiles");
            }
            if (binCount > 0) {
                Check(hipMemcpy(b->bins, bins, binCount * sizeof(tg_bin), hipMemcpyHostToDevice),
                      "staging bins");
            }
            if (entryCount > 0) {
                Check(hipMemcpy(b->entries, entries, entryCount * sizeof(uint32_t),
                                hipMemcpyHostToDevice),
                      "staging entries");
            }
            uint32_t counts[kCounterWords] = {};
            counts[kCountSlot] = uint32_t(binCount);
            Check(hipMemcpy(b->counters, counts, sizeof counts, hipMemcpyHostToDevice),
                  "staging the bin count");
        }
        b->tileCount = tileCount;
        b->binCount = binCount;
        b->maxEntry = checked.maxEntry;
        b->anyEntries = checked.any;
        b->outEnd = checked.outEnd;
        b->binSide = kBinSide;
        b->entryCount = entryCount;
        KeepTileBins(*b, tiles, tileCount);
    });
}

// Checks a run's output and catalogue against the staged batch.
void CheckRun(const tg_batch& b, const tg_catalogue* cat, size_t outCapacity) {
    if (b.outEnd > outCapacity) {
        throw HipError{"the batch's tiles run past the output buffer"};
    }
    if (b.anyEntries && (cat == nullptr || b.maxEntry >= cat->count)) {
        throw HipError{"the batch's bins name craters the catalogue does not hold"};
    }
    if (b.radius > 0.0 && b.anyEntries && cat->frames == nullptr) {
        throw HipError{"tiles on a sphere need the catalogue's crater frames "
                       "(tg_hip_catalogue_frames)"};
    }
    if (b.boulders != nullptr && (b.radius > 0.0) != (b.boulders->frames != nullptr)) {
        throw HipError{b.radius > 0.0
                           ? "boulders on a sphere need their frames"
                           : "boulders with frames are for tiles on a sphere"};
    }
}

// What a captured graph read of a batch's boulders: a generation taken at each boulder staging,
// which is when their lists' buffers and catalogue can change.
uint64_t BouldersKey(const tg_batch& b) {
    return b.boulders != nullptr ? b.boulderGeneration : 0;
}

}  // namespace

extern "C" {

int tg_hip_batch_stage(tg_batch* b, const tg_tile_desc* tiles, size_t tileCount,
                       const tg_bin* bins, size_t binCount, const uint32_t* entries,
                       size_t entryCount, char* err, size_t errLen) {
    return Stage(b, tiles, tileCount, bins, binCount, entries, entryCount, false, err, errLen);
}

int tg_hip_batch_stage_async(tg_batch* b, const tg_tile_desc* tiles, size_t tileCount,
                             const tg_bin* bins, size_t binCount, const uint32_t* entries,
                             size_t entryCount, char* err, size_t errLen) {
    return Stage(b, tiles, tileCount, bins, binCount, entries, entryCount, true, err, errLen);
}

int tg_hip_batch_stage_windows(tg_batch* b, const tg_tile_desc* tiles, size_t tileCount,
                               uint32_t binSide, const uint32_t* tileWindows,
                               const tg_window* windows, size_t windowCount, size_t* entryCount,
                               char* err, size_t errLen) {
    if (b == nullptr || (tileCount > 0 && (tiles == nullptr || tileWindows == nullptr)) ||
        (windowCount > 0 && windows == nullptr)) {
        SetErr(err, errLen, "tg_hip_batch_stage_windows: null argument");
        return TG_ERR_ARGUMENT;
    }
    if (binSide != 8 && binSide != 16 && binSide != 32) {
        SetErr(err, errLen, "tg_hip_batch_stage_windows: the bin side must be 8, 16 or 32");
        return TG_ERR_ARGUMENT;
    }
    return Guard(err, errLen, [&] {
        if (tileCount > b->maxTiles) {
            throw HipError{"batch exceeds the capacities it was created with"};
        }
        // The tiles as CheckBatch checks them, their bins counted from the descriptors.
        uint64_t binCount = 0;
        for (size_t t = 0; t < tileCount; ++t) {
            if (tiles[t].bin_first != binCount) {
                throw HipError{"tile " + std::to_string(t) + "'s bins do not follow the last's"};
            }
            binCount += uint64_t(tiles[t].bins_across) *
                        ((tiles[t].height + binSide - 1) / binSide);
        }
        if (binCount > b->maxBins || binCount > 0x7fffffffull) {
            throw HipError{"batch exceeds the capacities it was created with"};
        }
        const uint64_t outEnd = CheckTiles(tiles, tileCount, binCount, binSide);
        // The windows: each tile's within the list, inside the tile, craters ascending.
        uint64_t maxEntry = 0;
        for (size_t t = 0; t < tileCount; ++t) {
            const uint64_t first = tileWindows[2 * t], count = tileWindows[2 * t + 1];
            if (first + count > windowCount) {
                throw HipError{"tile " + std::to_string(t) + "'s windows run past the list"};
            }
            for (uint64_t e = first; e < first + count; ++e) {
                const tg_window& w = windows[e];
                if (w.col0 > w.col1 || w.col1 >= tiles[t].width || w.row0 > w.row1 ||
                    w.row1 >= tiles[t].height || (e > first && w.crater <= windows[e - 1].crater)) {
                    throw HipError{"window " + std::to_string(e) + " is malformed"};
                }
                maxEntry = std::max<uint64_t>(maxEntry, w.crater);
            }
        }
        Check(hipSetDevice(b->dev->ordinal), "hipSetDevice");
        Check(hipStreamSynchronize(b->stream), "waiting for the batch's previous run");
        if (windowCount > b->windowCapacity) {
            if (b->windows != nullptr) {
                Check(hipFree(b->windows), "hipFree");
                b->windows = nullptr;
                b->windowCapacity = 0;
            }
            b->windows = Alloc<tg_window>(windowCount + windowCount / 4);
            b->windowCapacity = windowCount + windowCount / 4;
        }
        if (b->tileWindows == nullptr) {
            b->tileWindows = Alloc<uint32_t>(2 * std::max<size_t>(b->maxTiles, 1));
            b->rowFirsts = Alloc<uint32_t>(std::max<size_t>(b->maxTiles, 1));
            b->binCounts = Alloc<uint32_t>(std::max<size_t>(b->maxBins, 1));
            b->binFirsts = Alloc<uint32_t>(std::max<size_t>(b->maxBins, 1));
            Check(hipcub::DeviceScan::ExclusiveSum(nullptr, b->scanBytes, b->binCounts,
                                                   b->binFirsts, int(std::max<size_t>(b->maxBins, 1)),
                                                   b->stream),
                  "sizing the bin scan");
            Check(hipMalloc(&b->scanTemp, std::max<size_t>(b->scanBytes, 1)), "hipMalloc");
        }
        const auto put = [&](void* dst, const void* src, size_t bytes, const char* what) {
            if (bytes > 0) {
                Check(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), what);
            }
        };
        put(b->tiles, tiles, tileCount * sizeof(tg_tile_desc), "staging tiles");
        put(b->tileWindows, tileWindows, 2 * tileCount * sizeof(uint32_t), "staging windows");
        // Each tile's first bin row, for the binning's one wave per row.
        std::vector<uint32_t> rowFirsts(tileCount);
        uint64_t rowCount = 0;
        for (size_t t = 0; t < tileCount; ++t) {
            rowFirsts[t] = uint32_t(rowCount);
            rowCount += (tiles[t].height + binSide - 1) / binSide;
        }
        put(b->rowFirsts, rowFirsts.data(), tileCount * sizeof(uint32_t), "staging rows");
        put(b->windows, windows, windowCount * sizeof(tg_window), "staging windows");
        size_t total = 0;
        if (binCount > 0) {
            const dim3 grid(uint32_t((rowCount + 3) / 4));
            hipLaunchKernelGGL(BinWindowRows<false>, grid, dim3(256), 0, b->stream, b->tiles,
                               uint32_t(tileCount), b->rowFirsts, uint32_t(rowCount),
                               b->tileWindows, b->windows, binSide, b->binCounts, nullptr,
                               nullptr, nullptr);
            Check(hipGetLastError(), "launching the bin count");
            size_t bytes = b->scanBytes;
            Check(hipcub::DeviceScan::ExclusiveSum(b->scanTemp, bytes, b->binCounts, b->binFirsts,
                                                   int(binCount), b->stream),
                  "summing the bin counts");
            uint32_t tail[2] = {0u, 0u};
            Check(hipMemcpyAsync(&tail[0], b->binFirsts + binCount - 1, sizeof(uint32_t),
                                 hipMemcpyDeviceToHost, b->stream), "reading the entry count");
            Check(hipMemcpyAsync(&tail[1], b->binCounts + binCount - 1, sizeof(uint32_t),
                                 hipMemcpyDeviceToHost, b->stream), "reading the entry count");
            Check(hipStreamSynchronize(b->stream), "binning the batch");
            total = size_t(tail[0]) + tail[1];
            if (total > b->maxEntries) {
                throw HipError{"the batch's bins list " + std::to_string(total) +
                               " entries, past the " + std::to_string(b->maxEntries) +
                               " it was created with"};
            }
            hipLaunchKernelGGL(BinWindowRows<true>, grid, dim3(256), 0, b->stream, b->tiles,
                               uint32_t(tileCount), b->rowFirsts, uint32_t(rowCount),
                               b->tileWindows, b->windows, binSide, nullptr, b->binFirsts,
                               b->bins, b->entries);
            Check(hipGetLastError(), "launching the bin fill");
        }
        hipLaunchKernelGGL(StaticCounters, dim3(1), dim3(64), 0, b->stream, b->counters,
                           uint32_t(binCount));
        Check(hipGetLastError(), "setting the queue counters");
        Check(hipStreamSynchronize(b->stream), "binning the batch");
        b->tileCount = tileCount;
        b->binCount = binCount;
        b->maxEntry = maxEntry;
        b->anyEntries = total > 0;
        b->outEnd = outEnd;
        b->binSide = binSide;
        b->entryCount = total;
        KeepTileBins(*b, tiles, tileCount);
        if (entryCount != nullptr) {
            *entryCount = total;
        }
    });
}

int tg_hip_batch_read_bins(tg_batch* b, tg_bin* bins, size_t binCapacity, uint32_t* entries,
                           size_t entryCapacity, size_t* binCount, size_t* entryCount, char* err,
                           size_t errLen) {
    if (b == nullptr || binCount == nullptr || entryCount == nullptr) {
        SetErr(err, errLen, "tg_hip_batch_read_bins: null argument");
        return TG_ERR_ARGUMENT;
    }
    return Guard(err, errLen, [&] {
        *binCount = b->binCount;
        *entryCount = b->entryCount;
        if (bins == nullptr && entries == nullptr) {
            return;  // the counts only
        }
        if (binCapacity < b->binCount || entryCapacity < b->entryCount || bins == nullptr ||
            entries == nullptr) {
            throw HipError{"tg_hip_batch_read_bins: the host arrays are too small"};
        }
        Check(hipSetDevice(b->dev->ordinal), "hipSetDevice");
        Check(hipStreamSynchronize(b->stream), "waiting for the batch");
        if (b->binCount > 0) {
            Check(hipMemcpy(bins, b->bins, b->binCount * sizeof(tg_bin), hipMemcpyDeviceToHost),
                  "reading the bins");
        }
        if (b->entryCount > 0) {
            Check(hipMemcpy(entries, b->entries, b->entryCount * sizeof(uint32_t),
                            hipMemcpyDeviceToHost),
                  "reading the entries");
        }
    });
}

int tg_hip_batch_cull_depth(tg_batch* b, uint32_t depth, char* err, size_t errLen) {
    if (b == nullptr || (depth != 32 && depth != 64 && depth != 128)) {
        SetErr(err, errLen, "tg_hip_batch_cull_depth: a batch and a depth of 32, 64 or 128");
        return TG_ERR_ARGUMENT;
    }
    b->cullDepth = depth;
    return TG_OK;
}

int tg_hip_batch_run(tg_batch* b, const tg_noise_plan* plan, float elevation,
                     const tg_catalogue* cat, float* out, size_t outCapacity, int useGraph,
                     char* err, size_t errLen) {
    if (b == nullptr || out == nullptr) {
        SetErr(err, errLen, "tg_hip_batch_run: null argument");
        return TG_ERR_ARGUMENT;
    }
    return Guard(err, errLen, [&] {
        CheckRun(*b, cat, outCapacity);
        if (b->binCount == 0) {
            return;
        }
        Check(hipSetDevice(b->dev->ordinal), "hipSetDevice");
        if (!useGraph) {
            Launch(*b, plan, elevation, cat, out);
            return;
        }
        // The graph holds the launch's arguments and grid, not the staged batch (the kernel reads
        // the bin count from device memory), so a restage of the same grid replays it.
        const uint32_t groups = GroupsFor(*b);
        const uint64_t planGeneration = plan != nullptr ? plan->generation : 0;
        const uint64_t catalogueGeneration = cat != nullptr ? cat->generation : 0;
        if (b->graph == nullptr || b->graphPlan != planGeneration ||
            b->graphCatalogue != catalogueGeneration ||
            b->graphOut != out || b->graphElevation != elevation || b->graphGroups != groups ||
            b->graphWalk != b->walk || b->graphHoisted != b->hoisted ||
            b->graphRadius != b->radius || b->graphBinSide != b->binSide ||
            b->graphCullDepth != b->cullDepth || b->graphBoulders != BouldersKey(*b) ||
            b->graphChunk != ChunkFor(*b, groups)) {
            DropGraph(*b);
            hipGraph_t graph = nullptr;
            Check(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal),
                  "hipStreamBeginCapture");
            try {
                Launch(*b, plan, elevation, cat, out);
            } catch (...) {
                (void)hipStreamEndCapture(b->stream, &graph);
                if (graph != nullptr) {
                    (void)hipGraphDestroy(graph);
                }
                throw;
            }
            Check(hipStreamEndCapture(b->stream, &graph), "hipStreamEndCapture");
            const hipError_t made = hipGraphInstantiate(&b->graph, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            Check(made, "hipGraphInstantiate");
            b->graphPlan = planGeneration;
            b->graphCatalogue = catalogueGeneration;
            b->graphOut = out;
            b->graphElevation = elevation;
            b->graphGroups = groups;
            b->graphWalk = b->walk;
            b->graphHoisted = b->hoisted;
            b->graphChunk = ChunkFor(*b, groups);
            b->graphRadius = b->radius;
            b->graphBinSide = b->binSide;
            b->graphCullDepth = b->cullDepth;
            b->graphBoulders = BouldersKey(*b);
        }
        Check(hipGraphLaunch(b->graph, b->stream), "hipGraphLaunch");
    });
}

int tg_hip_batch_tune(tg_batch* b, uint32_t crater_walk, uint32_t groups_per_cu,
                      uint32_t hoisted, uint32_t chunk, char* err, size_t errLen) {
    if (b == nullptr) {
        SetErr(err, errLen, "tg_hip_batch_tune: null batch");
        return TG_ERR_ARGUMENT;
    }
    if (crater_walk > TG_WALK_CULLED || groups_per_cu == 0 || groups_per_cu > 64 ||
        chunk == 0 || chunk > 256) {
        SetErr(err, errLen,
               "tg_hip_batch_tune: unknown crater walk, groups per CU not in 1-64 or chunk not "
               "in 1-256");
        return TG_ERR_ARGUMENT;
    }
    b->walk = crater_walk;
    b->groupsPerCu = groups_per_cu;
    b->hoisted = hoisted != 0u;
    b->chunk = chunk;
    return TG_OK;
}

int tg_hip_batch_surface(tg_batch* b, double sphere_radius, char* err, size_t errLen) {
    if (b == nullptr) {
        SetErr(err, errLen, "tg_hip_batch_surface: null batch");
        return TG_ERR_ARGUMENT;
    }
    if (!(sphere_radius >= 0.0) || !std::isfinite(sphere_radius)) {
        SetErr(err, errLen, "tg_hip_batch_surface: the radius must be 0 or positive and finite");
        return TG_ERR_ARGUMENT;
    }
    b->radius = sphere_radius;
    return TG_OK;
}

int tg_hip_batch_run_static(tg_dev* dev, const tg_tile_desc* tiles, size_t tileCount,
                            const tg_bin* bins, size_t binCount, const uint32_t* entries,
                            size_t entryCount, const tg_noise_plan* plan, float elevation,
                            const tg_catalogue* cat, double sphereRadius, float* out,
                            size_t outCapacity, void* stream, char* err, size_t errLen) {
    if (dev == nullptr || out == nullptr || (tileCount > 0 && tiles == nullptr) ||
        (binCount > 0 && bins == nullptr) || (entryCount > 0 && entries == nullptr)) {
        SetErr(err, errLen, "tg_hip_batch_run_static: null argument");
        return TG_ERR_ARGUMENT;
    }
    if (tileCount > TG_STATIC_MAX_TILES || binCount > TG_STATIC_MAX_BINS ||
        entryCount > TG_STATIC_MAX_ENTRIES) {
        SetErr(err, errLen,
               "tg_hip_batch_run_static: the batch exceeds the static scratch (at most " +
                   std::to_string(TG_STATIC_MAX_TILES) + " tiles, " +
                   std::to_string(TG_STATIC_MAX_BINS) + " bins, " +
                   std::to_string(TG_STATIC_MAX_ENTRIES) + " entries)");
        return TG_ERR_ARGUMENT;
    }
    if (!(sphereRadius >= 0.0) || !std::isfinite(sphereRadius)) {
        SetErr(err, errLen,
               "tg_hip_batch_run_static: the radius must be 0 or positive and finite");
        return TG_ERR_ARGUMENT;
    }
    if (dev->scratch == nullptr) {
        SetErr(err, errLen, "tg_hip_batch_run_static: " + dev->scratchError);
        return TG_ERR_DEVICE;
    }
    return Guard(err, errLen, [&] {
        const Checked checked = CheckBatch(tiles, tileCount, bins, binCount, entries, entryCount);
        // A batch view of the scratch, with the default launch choices; nothing of it is freed.
        tg_batch b;
        b.dev = dev;
        b.stream = static_cast<hipStream_t>(stream);
        b.counters = reinterpret_cast<uint32_t*>(dev->scratch);
        b.tiles = reinterpret_cast<tg_tile_desc*>(dev->scratch + kStaticTilesAt);
        b.bins = reinterpret_cast<tg_bin*>(dev->scratch + kStaticBinsAt);
        b.entries = reinterpret_cast<uint32_t*>(dev->scratch + kStaticEntriesAt);
        b.tileCount = tileCount;
        b.binCount = binCount;
        b.maxEntry = checked.maxEntry;
        b.anyEntries = checked.any;
        b.outEnd = checked.outEnd;
        b.radius = sphereRadius;
        CheckRun(b, cat, outCapacity);
        if (binCount == 0) {
            return;
        }
        Check(hipSetDevice(dev->ordinal), "hipSetDevice");
        const auto put = [&](void* dst, const void* src, size_t bytes, const char* what) {
            Check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, b.stream), what);
        };
        put(b.tiles, tiles, tileCount * sizeof(tg_tile_desc), "staging tiles");
        put(b.bins, bins, binCount * sizeof(tg_bin), "staging bins");
        if (entryCount > 0) {
            put(b.entries, entries, entryCount * sizeof(uint32_t), "staging entries");
        }
        // The counters from the previous call are cleared and the bin count set by a kernel.
        hipLaunchKernelGGL(StaticCounters, dim3(1), dim3(64), 0, b.stream, b.counters,
                           uint32_t(binCount));
        Check(hipGetLastError(), "setting the static queue counters");
        Launch(b, plan, elevation, cat, out);
    });
}

int tg_hip_catalogue_frames(tg_catalogue* cat, const tg_crater_frame* frames, size_t count,
                            char* err, size_t errLen) {
    if (cat == nullptr || (count > 0 && frames == nullptr)) {
        SetErr(err, errLen, "tg_hip_catalogue_frames: null argument");
        return TG_ERR_ARGUMENT;
    }
    if (count != cat->count) {
        SetErr(err, errLen, "tg_hip_catalogue_frames: one frame per crater of the catalogue");
        return TG_ERR_ARGUMENT;
    }
    if (cat->frames != nullptr) {
        // Captured graphs hold the frames' address, so a catalogue's frames are set once.
        SetErr(err, errLen, "tg_hip_catalogue_frames: the catalogue already has its frames");
        return TG_ERR_ARGUMENT;
    }
    return Guard(err, errLen, [&] {
        for (size_t i = 0; i < count; ++i) {
            const tg_crater_frame& f = frames[i];
            if (!(f.sphere_radius > 0.0f) || !std::isfinite(f.x) || !std::isfinite(f.y) ||
                !std::isfinite(f.z)) {
                throw HipError{"frame " + std::to_string(i) + " is not a frame on a sphere"};
            }
        }
        Check(hipSetDevice(cat->dev->ordinal), "hipSetDevice");
        tg_crater_frame* uploaded = UploadArray(frames, count);
        tg::PreciseTerms* precise = nullptr;
        try {
            precise = DevicePreciseTerms(cat->kernel, cat->records, uploaded, count);
        } catch (...) {
            (void)hipFree(uploaded);
            throw;
        }
        (void)hipFree(cat->precise);  // the plane's terms, without frame axes
        cat->frames = uploaded;
        cat->precise = precise;
        cat->generation = NextGeneration();  // graphs captured without the frames are stale
    });
}

int tg_hip_boulder_catalogue_create(tg_dev* dev, const tg_boulder* boulders, size_t count,
                                    const double* diameters, const tg_crater_frame* frames,
                                    tg_boulder_catalogue** out, char* err, size_t errLen) {
    if (dev == nullptr || out == nullptr ||
        (count > 0 && (boulders == nullptr || diameters == nullptr))) {
        SetErr(err, errLen, "tg_hip_boulder_catalogue_create: null argument");
        return TG_ERR_ARGUMENT;
    }
    *out = nullptr;
    auto cat = std::make_unique<tg_boulder_catalogue>();
    const int code = Guard(err, errLen, [&] {
        for (size_t i = 0; i < count; ++i) {
            const tg_boulder& b = boulders[i];
            // The record's radius is half its diameter, narrowed: the dome and the tile's reach
            // must describe one rock.
            if (!(diameters[i] > 0.0) || !std::isfinite(diameters[i]) ||
                b.radius != float(0.5 * diameters[i]) || !std::isfinite(b.exposed) ||
                !std::isfinite(b.x) || !std::isfinite(b.y) || b.facet_count > 2) {
                throw HipError{"boulder " + std::to_string(i) + " is malformed"};
            }
            if (frames != nullptr &&
                (!(frames[i].sphere_radius > 0.0f) || !std::isfinite(frames[i].x) ||
                 !std::isfinite(frames[i].y) || !std::isfinite(frames[i].z))) {
                throw HipError{"boulder frame " + std::to_string(i) + " is not a frame on a sphere"};
            }
        }
        Check(hipSetDevice(dev->ordinal), "hipSetDevice");
        cat->dev = dev;
        cat->count = count;
        cat->records = UploadArray(boulders, count);
        cat->diameters = UploadArray(diameters, count);
        if (frames != nullptr) {
            cat->frames = UploadArray(frames, count);
        }
    });
    if (code != TG_OK) {
        tg_hip_boulder_catalogue_free(cat.release());
        return code;
    }
    *out = cat.release();
    return TG_OK;
}

int tg_hip_boulder_catalogue_free(tg_boulder_catalogue* cat) {
    if (cat == nullptr) {
        return TG_OK;
    }
    if (cat->dev != nullptr) {
        (void)hipSetDevice(cat->dev->ordinal);
    }
    for (void* p : {static_cast<void*>(cat->records), static_cast<void*>(cat->diameters),
                    static_cast<void*>(cat->frames)}) {
        if (p != nullptr) {
            (void)hipFree(p);
        }
    }
    delete cat;
    return TG_OK;
}

int tg_hip_batch_stage_boulders(tg_batch* b, const tg_boulder_catalogue* cat, const tg_bin* bins,
                                size_t binCount, const uint32_t* entries, size_t entryCount,
                                int async, char* err, size_t errLen) {
    if (b == nullptr || cat == nullptr || (binCount > 0 && bins == nullptr) ||
        (entryCount > 0 && entries == nullptr)) {
        SetErr(err, errLen, "tg_hip_batch_stage_boulders: null argument");
        return TG_ERR_ARGUMENT;
    }
    return Guard(err, errLen, [&] {
        if (cat->dev != b->dev) {
            throw HipError{"the boulders are on another device than the batch"};
        }
        if (binCount != b->binCount || b->binSide != kBinSide) {
            throw HipError{"boulder lists need one bin per staged bin, of 16 x 16 cells"};
        }
        if (entryCount > 0xFFFFFFFFull) {
            throw HipError{"boulder entries must fit 32-bit indices"};
        }
        // Every bin where the staged tiles put it, its list within the entries; every entry a
        // boulder of the catalogue.
        for (size_t t = 0; t < b->tileBinFirst.size(); ++t) {
            for (uint64_t i = b->tileBinFirst[t];
                 i < uint64_t(b->tileBinFirst[t]) + b->tileBinCount[t]; ++i) {
                if (bins[i].tile != t || uint64_t(bins[i].first) + bins[i].count > entryCount) {
                    throw HipError{"boulder bin " + std::to_string(i) + " is malformed"};
                }
            }
        }
        for (size_t e = 0; e < entryCount; ++e) {
            if (entries[e] >= cat->count) {
                throw HipError{"boulder entry " + std::to_string(e) +
                               " names a boulder the catalogue does not hold"};
            }
        }
        Check(hipSetDevice(b->dev->ordinal), "hipSetDevice");
        // Copies waited for are not ordered with the batch's stream, so its work must be done;
        // queued ones follow it on the stream.
        if (!async) {
            Check(hipStreamSynchronize(b->stream), "waiting for the batch's previous run");
        }
        if (b->boulderBins == nullptr) {
            b->boulderBins = Alloc<tg_bin>(std::max<size_t>(b->maxBins, 1));
        }
        if (entryCount > b->boulderEntryCapacity) {
            // The entries' buffer is replaced, so nothing queued may still read it.
            Check(hipStreamSynchronize(b->stream), "waiting for the batch's previous run");
            if (b->boulderEntries != nullptr) {
                Check(hipFree(b->boulderEntries), "hipFree");
                b->boulderEntries = nullptr;
                b->boulderEntryCapacity = 0;
            }
            const size_t capacity = entryCount + entryCount / 4 + 1;
            b->boulderEntries = Alloc<uint32_t>(capacity);
            b->boulderEntryCapacity = capacity;
        }
        const auto put = [&](void* dst, const void* src, size_t bytes, const char* what) {
            if (bytes == 0) {
                return;
            }
            Check(async ? hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, b->stream)
                        : hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice),
                  what);
        };
        put(b->boulderBins, bins, binCount * sizeof(tg_bin), "staging boulder bins");
        put(b->boulderEntries, entries, entryCount * sizeof(uint32_t), "staging boulder entries");
        b->boulders = cat;
        b->boulderGeneration = NextGeneration();
    });
}

int tg_hip_batch_wait(tg_batch* b, char* err, size_t errLen) {
    if (b == nullptr) {
        SetErr(err, errLen, "tg_hip_batch_wait: null batch");
        return TG_ERR_ARGUMENT;
    }
    return Guard(err, errLen, [&] {
        Check(hipSetDevice(b->dev->ordinal), "hipSetDevice");
        Check(hipStreamSynchronize(b->stream), "running the batch");
    });
}

void* tg_hip_batch_stream(tg_batch* b) {
    return b != nullptr ? b->stream : nullptr;
}

}  // extern "C"
