// SPDX-License-Identifier: BSD-3-Clause
//
// TEST INFRASTRUCTURE ONLY. A CPU stand-in for the kernel library (libtg_hip.so), with the entry
// points the tile server opens (terraingen_amd/csrc/host/hip_api.h), so the server's threads can
// run under ThreadSanitizer on a machine without a GPU. "Device" memory is host memory; each
// batch's stream is a worker thread running its queued work in order, as a HIP stream does, so
// copies queued by tg_hip_batch_stage_async read the caller's pinned arrays later, on another
// thread, and events report completion only once the worker has passed them. A run writes a
// height per cell that names the tile, the cell and the number of craters and boulders in its
// bin (standin_height), which server_driver.cpp recomputes to check each request got its own tile.

#include "tg_hip.h"

#include <condition_variable>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace {

void SetError(char* err, size_t len, const char* what) {
    if (err != nullptr && len > 0) {
        std::strncpy(err, what, len - 1);
        err[len - 1] = '\0';
    }
}

// In-order work on its own thread.
class Stream {
public:
    Stream() : worker_([this] { Run(); }) {}
    ~Stream() {
        {
            std::lock_guard<std::mutex> lock(mutex_);
            stop_ = true;
        }
        wake_.notify_all();
        worker_.join();
    }
    void Enqueue(std::function<void()> task) {
        {
            std::lock_guard<std::mutex> lock(mutex_);
            tasks_.push_back(std::move(task));
        }
        wake_.notify_all();
    }
    void Synchronize() {
        std::unique_lock<std::mutex> lock(mutex_);
        idle_.wait(lock, [&] { return tasks_.empty() && !busy_; });
    }

private:
    void Run() {
        std::unique_lock<std::mutex> lock(mutex_);
        for (;;) {
            wake_.wait(lock, [&] { return stop_ || !tasks_.empty(); });
            if (tasks_.empty()) {
                return;
            }
            std::function<void()> task = std::move(tasks_.front());
            tasks_.pop_front();
            busy_ = true;
            lock.unlock();
            // A little latency, so the server's threads overlap the work as they do on a GPU.
            std::this_thread::sleep_for(std::chrono::microseconds(50));
            task();
            lock.lock();
            busy_ = false;
            if (tasks_.empty()) {
                idle_.notify_all();
            }
        }
    }

    std::mutex mutex_;
    std::condition_variable wake_, idle_;
    std::deque<std::function<void()>> tasks_;
    bool busy_ = false;
    bool stop_ = false;
    std::thread worker_;
};

struct Event {
    std::mutex mutex;
    bool done = true;
};

}  // namespace

struct tg_dev {
    int ordinal = 0;
};

struct tg_batch {
    size_t maxTiles = 0, maxBins = 0, maxEntries = 0;
    std::vector<tg_tile_desc> tiles;
    std::vector<tg_bin> bins;
    std::vector<uint32_t> entries;
    std::vector<tg_bin> boulderBins;
    Stream stream;
};

extern "C" {

// The height the stand-in's run writes at (col, row) of tile `t` whose bin lists `craters` and
// `boulders` entries: server_driver.cpp computes the same FP32 expression.
float standin_height(const tg_tile_desc* t, uint32_t col, uint32_t row, uint32_t craters,
                     uint32_t boulders, float elevation) {
    return elevation + float(t->origin_x) * 0.001f + float(t->origin_y) * 0.002f + float(col) +
           512.0f * float(row % 512u) + 300000.0f * float(craters) + 7000000.0f * float(boulders);
}

int tg_hip_open(int ordinal, tg_dev** out, char* err, size_t errLen) {
    if (out == nullptr) {
        SetError(err, errLen, "null argument");
        return TG_ERR_ARGUMENT;
    }
    static tg_dev dev;
    dev.ordinal = ordinal;
    *out = &dev;
    return TG_OK;
}

const char* tg_hip_device_name(const tg_dev*) {
    return "CPU stand-in";
}

int tg_hip_can_run(const tg_dev*, int) {
    return 0;
}

int tg_hip_run(tg_dev*, int, const void*, size_t, float*, const float*, const float*,
               const tg_kernel_params*, const tg_buf*, size_t, char* err, size_t errLen) {
    SetError(err, errLen, "the CPU stand-in runs no kernels");
    return TG_ERR_ARGUMENT;
}

int tg_hip_batch_create(tg_dev* dev, size_t maxTiles, size_t maxBins, size_t maxEntries,
                        tg_batch** out, char* err, size_t errLen) {
    if (dev == nullptr || out == nullptr) {
        SetError(err, errLen, "null argument");
        return TG_ERR_ARGUMENT;
    }
    auto* b = new tg_batch;
    b->maxTiles = maxTiles;
    b->maxBins = maxBins;
    b->maxEntries = maxEntries;
    *out = b;
    return TG_OK;
}

int tg_hip_batch_free(tg_batch* batch) {
    delete batch;
    return TG_OK;
}

int tg_hip_batch_stage(tg_batch*, const tg_tile_desc*, size_t, const tg_bin*, size_t,
                       const uint32_t*, size_t, char* err, size_t errLen) {
    SetError(err, errLen, "the server stages asynchronously");
    return TG_ERR_ARGUMENT;
}

int tg_hip_batch_stage_async(tg_batch* b, const tg_tile_desc* tiles, size_t tileCount,
                             const tg_bin* bins, size_t binCount, const uint32_t* entries,
                             size_t entryCount, char* err, size_t errLen) {
    if (b == nullptr || tileCount > b->maxTiles || binCount > b->maxBins ||
        entryCount > b->maxEntries) {
        SetError(err, errLen, "batch capacity exceeded");
        return TG_ERR_ARGUMENT;
    }
    // The copies read the caller's arrays when the stream reaches them, as hipMemcpyAsync does.
    b->stream.Enqueue([b, tiles, tileCount, bins, binCount, entries, entryCount] {
        b->tiles.assign(tiles, tiles + tileCount);
        b->bins.assign(bins, bins + binCount);
        b->entries.assign(entries, entries + entryCount);
    });
    return TG_OK;
}

void* tg_hip_batch_stream(tg_batch* b) {
    return b != nullptr ? &b->stream : nullptr;
}

int tg_hip_batch_surface(tg_batch* b, double, char* err, size_t errLen) {
    if (b == nullptr) {
        SetError(err, errLen, "null batch");
        return TG_ERR_ARGUMENT;
    }
    return TG_OK;
}

int tg_hip_batch_stage_boulders(tg_batch* b, const tg_boulder_catalogue* catalogue,
                                const tg_bin* bins, size_t binCount, const uint32_t*, size_t,
                                int async, char* err, size_t errLen) {
    if (b == nullptr || catalogue == nullptr || async == 0) {
        SetError(err, errLen, "the server stages boulders asynchronously");
        return TG_ERR_ARGUMENT;
    }
    b->stream.Enqueue([b, bins, binCount] { b->boulderBins.assign(bins, bins + binCount); });
    return TG_OK;
}

int tg_hip_batch_run(tg_batch* b, const tg_noise_plan*, float elevation, const tg_catalogue*,
                     float* out, size_t capacity, int, char* err, size_t errLen) {
    if (b == nullptr || out == nullptr) {
        SetError(err, errLen, "null argument");
        return TG_ERR_ARGUMENT;
    }
    b->stream.Enqueue([b, elevation, out, capacity] {
        for (const tg_tile_desc& t : b->tiles) {
            if (t.out_offset + uint64_t(t.width) * t.height > capacity) {
                std::abort();  // the server's offsets must fit the region it passes
            }
            for (uint32_t row = 0; row < t.height; ++row) {
                for (uint32_t col = 0; col < t.width; ++col) {
                    const size_t bin = t.bin_first + (row / 16u) * t.bins_across + col / 16u;
                    const uint32_t rocks =
                        b->boulderBins.empty() ? 0u : b->boulderBins.at(bin).count;
                    out[t.out_offset + size_t(row) * t.width + col] =
                        standin_height(&t, col, row, b->bins.at(bin).count, rocks, elevation);
                }
            }
        }
    });
    return TG_OK;
}

int tg_hip_malloc(tg_dev*, size_t bytes, void** out, char* err, size_t errLen) {
    *out = std::malloc(bytes > 0 ? bytes : 1);
    if (*out == nullptr) {
        SetError(err, errLen, "out of memory");
        return TG_ERR_DEVICE;
    }
    return TG_OK;
}

int tg_hip_free(tg_dev*, void* p) {
    std::free(p);
    return TG_OK;
}

int tg_hip_host_alloc(tg_dev* dev, size_t bytes, void** out, char* err, size_t errLen) {
    return tg_hip_malloc(dev, bytes, out, err, errLen);
}

int tg_hip_host_free(tg_dev* dev, void* p) {
    return tg_hip_free(dev, p);
}

int tg_hip_memcpy_d2h_async(tg_dev*, void* dst, const void* src, size_t bytes, void* stream,
                            char* err, size_t errLen) {
    if (stream == nullptr) {
        SetError(err, errLen, "null stream");
        return TG_ERR_ARGUMENT;
    }
    static_cast<Stream*>(stream)->Enqueue([dst, src, bytes] { std::memcpy(dst, src, bytes); });
    return TG_OK;
}

int tg_hip_event_create(tg_dev*, void** out, char*, size_t) {
    *out = new Event;
    return TG_OK;
}

int tg_hip_event_destroy(tg_dev*, void* event) {
    delete static_cast<Event*>(event);
    return TG_OK;
}

int tg_hip_event_record(tg_dev*, void* event, void* stream, char* err, size_t errLen) {
    if (event == nullptr || stream == nullptr) {
        SetError(err, errLen, "null argument");
        return TG_ERR_ARGUMENT;
    }
    auto* e = static_cast<Event*>(event);
    {
        std::lock_guard<std::mutex> lock(e->mutex);
        e->done = false;
    }
    static_cast<Stream*>(stream)->Enqueue([e] {
        std::lock_guard<std::mutex> lock(e->mutex);
        e->done = true;
    });
    return TG_OK;
}

int tg_hip_event_query(tg_dev*, void* event, char*, size_t) {
    auto* e = static_cast<Event*>(event);
    std::lock_guard<std::mutex> lock(e->mutex);
    return e->done ? TG_OK : TG_NOT_READY;
}

int tg_hip_stream_synchronize(tg_dev*, void* stream, char*, size_t) {
    static_cast<Stream*>(stream)->Synchronize();
    return TG_OK;
}

}  // extern "C"
