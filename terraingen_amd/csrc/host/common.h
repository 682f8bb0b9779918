// SPDX-License-Identifier: BSD-3-Clause
//
// Shared pieces of the host library: grids and pixel windows in the reference's conventions
// (include/terraingen/raster/{geo_transform,cell_loops}.h), error reporting across the C API.
#pragma once

#include "terraingen_amd.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <exception>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace tgh {

// A north-up grid (GeoTransform::NorthUp: pixelWidth = pixel, pixelHeight = -pixel).
struct Grid {
    std::size_t width = 0;
    std::size_t height = 0;
    double originX = 0.0;
    double originY = 0.0;
    double pixel = 1.0;

    static Grid From(const tgx_grid& g) {
        return {g.width, g.height, g.origin_x, g.origin_y, g.pixel};
    }

    // GeoTransform::PixelToWorld of the cell's center (raster.h:187-190).
    double CenterX(std::size_t col) const {
        return originX + (static_cast<double>(col) + 0.5) * pixel + 0.0;
    }
    double CenterY(std::size_t row) const {
        return originY + (static_cast<double>(row) + 0.5) * -pixel;
    }

    // GeoTransform::WorldToPixel (geo_transform.h:72-81) without rotation terms changing bits:
    // det = w * h - 0 * 0, col = (dx * h - dy * 0) / det, row = (dy * w - dx * 0) / det.
    void WorldToPixel(double x, double y, double& col, double& row) const {
        const double w = pixel, h = -pixel;
        const double det = w * h - 0.0 * 0.0;
        const double dx = x - originX;
        const double dy = y - originY;
        col = (dx * h - dy * 0.0) / det;
        row = (dy * w - dx * 0.0) / det;
    }
};

struct PixelWindow {
    std::size_t colBegin = 0, colEnd = 0, rowBegin = 0, rowEnd = 0;
    bool Empty() const {
        return colBegin >= colEnd || rowBegin >= rowEnd;
    }
};

// ClampIndex(value, limit + 1) of cell_loops.cpp:14-16 with bits.h:42-44.
inline std::size_t ClampTo(double value, std::size_t limit) {
    return static_cast<std::size_t>(std::clamp(value, 0.0, static_cast<double>(limit)));
}

// PixelWindowIn (src/raster/cell_loops.cpp:20-27): the box's corners mapped to pixel
// coordinates, floored and ceiled, clamped to the grid.
inline PixelWindow PixelWindowIn(const Grid& g, double minX, double minY, double maxX,
                                 double maxY) {
    // On a north-up grid a corner's column depends on its x alone and its row on its y alone (the
    // rotation terms are products with zero, which change at most the sign of a zero), so two
    // opposite corners give the four corners' columns and rows.
    double c0, r0, c1, r1;
    g.WorldToPixel(minX, minY, c0, r0);
    g.WorldToPixel(maxX, maxY, c1, r1);
    const double lo[2] = {std::fmin(c0, c1), std::fmin(r0, r1)};
    const double hi[2] = {std::fmax(c0, c1), std::fmax(r0, r1)};
    return {ClampTo(std::floor(lo[0]), g.width), ClampTo(std::ceil(hi[0]), g.width),
            ClampTo(std::floor(lo[1]), g.height), ClampTo(std::ceil(hi[1]), g.height)};
}

// Copies `v` into a malloc'd array the caller frees with tgx_free.
template <class T> T* Export(const std::vector<T>& v) {
    T* out = static_cast<T*>(std::malloc(std::max<std::size_t>(1, v.size()) * sizeof(T)));
    if (out == nullptr) {
        throw std::bad_alloc();
    }
    if (!v.empty()) {
        std::memcpy(out, v.data(), v.size() * sizeof(T));
    }
    return out;
}

inline void SetErr(char* err, std::size_t len, const std::string& message) {
    if (err != nullptr && len > 0) {
        std::snprintf(err, len, "%s", message.c_str());
    }
}

// Runs `body`, turning exceptions into a nonzero return and a message.
template <class F> int Guard(char* err, std::size_t len, F body) {
    try {
        body();
        return 0;
    } catch (const std::invalid_argument& e) {
        SetErr(err, len, e.what());
        return 1;
    } catch (const std::exception& e) {
        SetErr(err, len, e.what());
        return 2;
    }
}

// The thread count host routines use when asked for 0 ("all"): tgx_set_host_threads, else every
// hardware thread.
inline std::atomic<unsigned>& HostThreads() {
    static std::atomic<unsigned> threads{0};
    return threads;
}

// Runs task(n) for n in [0, count) on `threads` threads (0: HostThreads' count).
template <class Task>
void ForEachIndex(std::size_t count, unsigned threads, const Task& task) {
    if (threads == 0) {
        threads = HostThreads().load();
    }
    if (threads == 0) {
        threads = std::max(1u, std::thread::hardware_concurrency());
    }
    threads = static_cast<unsigned>(std::min<std::size_t>(threads, count));
    if (threads <= 1) {
        for (std::size_t n = 0; n < count; ++n) {
            task(n);
        }
        return;
    }
    std::atomic<std::size_t> next{0};
    std::exception_ptr failure;
    std::mutex failed;
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < threads; ++t) {
        pool.emplace_back([&] {
            for (std::size_t n = next++; n < count; n = next++) {
                try {
                    task(n);
                } catch (...) {
                    std::lock_guard<std::mutex> lock(failed);
                    if (!failure) {
                        failure = std::current_exception();
                    }
                    next = count;  // the rest are skipped
                }
            }
        });
    }
    for (std::thread& th : pool) {
        th.join();
    }
    if (failure) {
        std::rethrow_exception(failure);
    }
}

// Sorts `v` by `less`, which must be a strict total order (so the result is the one std::sort
// gives): runs sorted on threads, then merged pairwise, each level's merges on threads.
template <class T, class Less>
void ParallelSort(std::vector<T>& v, const Less& less, unsigned threads) {
    constexpr std::size_t kRun = std::size_t{1} << 16;
    const std::size_t n = v.size();
    if (n <= kRun) {
        std::sort(v.begin(), v.end(), less);
        return;
    }
    const auto at = [&](std::vector<T>& w, std::size_t k) {
        return w.begin() + static_cast<std::ptrdiff_t>(std::min(k, n));
    };
    ForEachIndex((n + kRun - 1) / kRun, threads, [&](std::size_t r) {
        std::sort(at(v, r * kRun), at(v, (r + 1) * kRun), less);
    });
    std::vector<T> other(n);
    for (std::size_t width = kRun; width < n; width *= 2) {
        ForEachIndex((n + 2 * width - 1) / (2 * width), threads, [&](std::size_t p) {
            const std::size_t lo = p * 2 * width;
            std::merge(at(v, lo), at(v, lo + width), at(v, lo + width), at(v, lo + 2 * width),
                       at(other, lo), less);
        });
        v.swap(other);
    }
}

}  // namespace tgh
