// SPDX-License-Identifier: BSD-3-Clause
//
// The tile server: continuous batching of tile requests onto the batched kernel.
//
// Requests (a tile and the host memory its heights go to) queue as they arrive. A dispatcher
// thread owns a few slots, each a staged batch with its own stream, event and pinned staging
// memory. Whenever a slot is free it admits the waiting requests in arrival order, as many as fit
// the batch limits and the HBM output pool (the admission budget), bins them against the kept
// catalogue index into the slot's pinned buffers, queues their copies, one persistent kernel for
// the whole batch and the copy of the heights back, all on the slot's stream, and moves on. Batches
// in different slots overlap on the device, and requests that arrive while others run are admitted
// as slots and pool memory free up. A finished slot goes to copier threads, which copy its heights
// into each request's memory and return the slot and its pool range, so the dispatcher never waits
// on a copy. Buffers outgrown by a batch are kept to the end rather than freed, since freeing
// device memory waits for the whole device. The crater catalogue, its index and the noise
// constants stay resident throughout.
//
// Each tile's heights depend only on the tile and the resident inputs (its bins are its own and
// every cell is independent), so a tile gives the same heights alone and in any mix.

#include "batch.h"
#include "common.h"
#include "hip_api.h"

#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>

namespace tgh {

namespace {

using Clock = std::chrono::steady_clock;

double Now() {
    return std::chrono::duration<double>(Clock::now().time_since_epoch()).count();
}

constexpr std::uint64_t kAlign = 256;

std::uint64_t AlignUp(std::uint64_t bytes) {
    return (bytes + kAlign - 1) / kAlign * kAlign;
}

// First-fit ranges of one device allocation.
class RangePool {
public:
    explicit RangePool(std::uint64_t size) : size_(size) {
        if (size > 0) {
            free_[0] = size;
        }
    }

    bool Take(std::uint64_t bytes, std::uint64_t& offset) {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                offset = it->first;
                const std::uint64_t rest = it->second - bytes;
                free_.erase(it);
                if (rest > 0) {
                    free_[offset + bytes] = rest;
                }
                inUse_ += bytes;
                peak_ = std::max(peak_, inUse_);
                return true;
            }
        }
        return false;
    }

    void Give(std::uint64_t offset, std::uint64_t bytes) {
        inUse_ -= bytes;
        auto it = free_.emplace(offset, bytes).first;
        auto next = std::next(it);
        if (next != free_.end() && it->first + it->second == next->first) {
            it->second += next->second;
            free_.erase(next);
        }
        if (it != free_.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free_.erase(it);
            }
        }
    }

    std::uint64_t Peak() const {
        return peak_;
    }
    std::uint64_t Size() const {
        return size_;
    }

private:
    std::uint64_t size_ = 0;
    std::map<std::uint64_t, std::uint64_t> free_;
    std::uint64_t inUse_ = 0;
    std::uint64_t peak_ = 0;
};

struct Request {
    std::uint64_t ticket = 0;
    tgx_tile tile{};
    float* dst = nullptr;
    tgx_ticket_info info{};
    int status = TG_NOT_READY;
    std::string error;
};

struct Slot {
    tg_batch* batch = nullptr;
    void* stream = nullptr;
    void* event = nullptr;
    float* staging = nullptr;  // pinned, max_batch_cells floats
    // Pinned copies of the staged batch, read by the copies queued on the stream.
    tg_tile_desc* hostDescs = nullptr;
    tg_bin* hostBins = nullptr;
    std::uint32_t* hostEntries = nullptr;
    std::size_t maxTiles = 0, maxBins = 0, maxEntries = 0;
    // The same for the batch's boulder lists, when the server draws boulders.
    tg_bin* hostBoulderBins = nullptr;
    std::uint32_t* hostBoulderEntries = nullptr;
    std::size_t maxBoulderBins = 0, maxBoulderEntries = 0;
    bool busy = false;
    bool copying = false;  // finished on the device, its heights with a copier
    std::vector<Request*> requests;  // in batch order
    std::vector<std::uint64_t> offsets;
    std::uint64_t poolOffset = 0, poolBytes = 0;
};

void Raise(int code, const char* err) {
    if (code != TG_OK) {
        throw std::runtime_error(err[0] != 0 ? err : "kernel library call failed");
    }
}

}  // namespace

}  // namespace tgh

struct tgx_server {
    const tgh::HipApi* hip = nullptr;
    tg_dev* dev = nullptr;
    const tg_noise_plan* plan = nullptr;
    float elevation = 0.0f;
    const tg_catalogue* catalogue = nullptr;
    const tgx_index* index = nullptr;
    tgh::IndexPtr emptyIndex;
    // The boulder layer (tgx_server_boulders): its catalogue and the index of its relief boxes.
    const tg_boulder_catalogue* boulders = nullptr;
    const tgx_index* boulderIndex = nullptr;
    tgx_server_config config{};
    float* pool = nullptr;
    std::unique_ptr<tgh::RangePool> ranges;
    std::vector<tgh::Slot> slots;

    std::mutex mutex;
    std::condition_variable wake;  // the dispatcher: new requests or shutdown
    std::condition_variable done;  // waiters: a request finished
    std::deque<tgh::Request*> pending;
    std::unordered_map<std::uint64_t, std::unique_ptr<tgh::Request>> requests;
    std::uint64_t nextTicket = 1;
    std::uint32_t nextBatch = 1;
    bool stop = false;
    tgx_server_stats stats{};
    std::thread dispatcher;
    // Finished slots waiting for a copier, and the copiers.
    std::deque<tgh::Slot*> copyQueue;
    std::condition_variable copyWake;
    bool copiersStop = false;
    std::vector<std::thread> copiers;
    // Batches and pinned buffers outgrown by a slot, freed with the server.
    std::vector<tg_batch*> retiredBatches;
    std::vector<void*> retiredPinned;
};

namespace tgh {

namespace {

void FreeSlot(tgx_server& s, Slot& slot) {
    const HipApi& h = *s.hip;
    if (slot.batch != nullptr) {
        h.batch_free(slot.batch);
        slot.batch = nullptr;
    }
    if (slot.event != nullptr) {
        h.event_destroy(s.dev, slot.event);
        slot.event = nullptr;
    }
    for (void* p : {static_cast<void*>(slot.staging), static_cast<void*>(slot.hostDescs),
                    static_cast<void*>(slot.hostBins), static_cast<void*>(slot.hostEntries),
                    static_cast<void*>(slot.hostBoulderBins),
                    static_cast<void*>(slot.hostBoulderEntries)}) {
        if (p != nullptr) {
            h.host_free(s.dev, p);
        }
    }
    slot.staging = nullptr;
    slot.hostDescs = nullptr;
    slot.hostBins = nullptr;
    slot.hostEntries = nullptr;
    slot.hostBoulderBins = nullptr;
    slot.hostBoulderEntries = nullptr;
}

template <class T> T* PinnedArray(tgx_server& s, std::size_t count) {
    char err[512] = {0};
    void* p = nullptr;
    Raise(s.hip->host_alloc(s.dev, std::max<std::size_t>(count, 1) * sizeof(T), &p, err,
                            sizeof err),
          err);
    return static_cast<T*>(p);
}

// Gives the slot (idle) a batch and pinned buffers of at least the given capacities, keeping
// what it outgrows until the server is freed. The new buffers are made first and the slot takes
// them only once all exist, so a failed allocation leaves the slot as it was.
void EnsureCapacity(tgx_server& s, Slot& slot, std::size_t tiles, std::size_t bins,
                    std::size_t entries) {
    if (slot.batch != nullptr && tiles <= slot.maxTiles && bins <= slot.maxBins &&
        entries <= slot.maxEntries) {
        return;
    }
    const HipApi& h = *s.hip;
    char err[512] = {0};
    // A slot that outgrows its buffers gets twice what the batch needs.
    const bool grow = slot.batch != nullptr;
    const std::size_t t = std::max(tiles, slot.maxTiles);
    const std::size_t b = std::max(grow && bins > slot.maxBins ? bins * 2 : bins, slot.maxBins);
    const std::size_t e = std::max(
        grow && entries > slot.maxEntries ? entries * 2 + 1024 : entries, slot.maxEntries);
    tg_batch* batch = nullptr;
    tg_tile_desc* descs = nullptr;
    tg_bin* binBuf = nullptr;
    std::uint32_t* entryBuf = nullptr;
    try {
        Raise(h.batch_create(s.dev, t, b, e, &batch, err, sizeof err), err);
        if (const double radius = IndexRadius(*s.index); radius > 0.0) {
            // Tiles binned against a sphere's index are longitude-latitude tiles on it.
            Raise(h.batch_surface(batch, radius, err, sizeof err), err);
        }
        descs = PinnedArray<tg_tile_desc>(s, t);
        binBuf = PinnedArray<tg_bin>(s, b);
        entryBuf = PinnedArray<std::uint32_t>(s, e);
    } catch (...) {
        for (void* p : {static_cast<void*>(descs), static_cast<void*>(binBuf),
                        static_cast<void*>(entryBuf)}) {
            if (p != nullptr) {
                h.host_free(s.dev, p);
            }
        }
        if (batch != nullptr) {
            h.batch_free(batch);
        }
        throw;
    }
    if (slot.batch != nullptr) {
        s.retiredBatches.push_back(slot.batch);
        s.retiredPinned.insert(s.retiredPinned.end(),
                               {static_cast<void*>(slot.hostDescs),
                                static_cast<void*>(slot.hostBins),
                                static_cast<void*>(slot.hostEntries)});
    }
    slot.batch = batch;
    slot.stream = h.batch_stream(batch);
    slot.hostDescs = descs;
    slot.hostBins = binBuf;
    slot.hostEntries = entryBuf;
    slot.maxTiles = t;
    slot.maxBins = b;
    slot.maxEntries = e;
}

// Gives the slot (idle) pinned boulder lists of at least the given capacities, as EnsureCapacity
// does for its crater lists: what it outgrows is kept until the server is freed.
void EnsureBoulderCapacity(tgx_server& s, Slot& slot, std::size_t bins, std::size_t entries) {
    if (slot.hostBoulderBins != nullptr && bins <= slot.maxBoulderBins &&
        entries <= slot.maxBoulderEntries) {
        return;
    }
    const HipApi& h = *s.hip;
    const std::size_t b = std::max(bins * 2, slot.maxBoulderBins);
    const std::size_t e = std::max(entries * 2 + 1024, slot.maxBoulderEntries);
    tg_bin* binBuf = PinnedArray<tg_bin>(s, b);
    std::uint32_t* entryBuf = nullptr;
    try {
        entryBuf = PinnedArray<std::uint32_t>(s, e);
    } catch (...) {
        h.host_free(s.dev, binBuf);
        throw;
    }
    if (slot.hostBoulderBins != nullptr) {
        s.retiredPinned.insert(s.retiredPinned.end(),
                               {static_cast<void*>(slot.hostBoulderBins),
                                static_cast<void*>(slot.hostBoulderEntries)});
    }
    slot.hostBoulderBins = binBuf;
    slot.hostBoulderEntries = entryBuf;
    slot.maxBoulderBins = b;
    slot.maxBoulderEntries = e;
}

// Bins, stages and launches one batch in an idle slot, then queues the copy back and the event.
// Runs without the server's lock (the slot is the dispatcher's while it is being launched) and
// returns the launch time, which the dispatcher records under the lock.
double LaunchBatch(tgx_server& s, Slot& slot) {
    const HipApi& h = *s.hip;
    char err[512] = {0};
    std::vector<tgx_tile> tiles(slot.requests.size());
    for (std::size_t i = 0; i < tiles.size(); ++i) {
        tiles[i] = slot.requests[i]->tile;
    }
    std::vector<tg_tile_desc> descs;
    std::vector<tg_bin> bins;
    std::vector<std::uint32_t> entries;
    BinBatch(*s.index, tiles.data(), tiles.size(), s.config.bin_threads, descs, bins, entries);
    EnsureCapacity(s, slot, descs.size(), bins.size(), entries.size());
    std::memcpy(slot.hostDescs, descs.data(), descs.size() * sizeof(tg_tile_desc));
    std::memcpy(slot.hostBins, bins.data(), bins.size() * sizeof(tg_bin));
    std::memcpy(slot.hostEntries, entries.data(), entries.size() * sizeof(std::uint32_t));
    Raise(h.batch_stage_async(slot.batch, slot.hostDescs, descs.size(), slot.hostBins,
                              bins.size(), slot.hostEntries, entries.size(), err, sizeof err),
          err);
    if (s.boulders != nullptr) {
        // The boulder layer's lists of the same tiles, binned against its own index.
        std::vector<tg_tile_desc> boulderDescs;
        std::vector<tg_bin> boulderBins;
        std::vector<std::uint32_t> boulderEntries;
        BinBatch(*s.boulderIndex, tiles.data(), tiles.size(), s.config.bin_threads, boulderDescs,
                 boulderBins, boulderEntries);
        EnsureBoulderCapacity(s, slot, boulderBins.size(), boulderEntries.size());
        std::memcpy(slot.hostBoulderBins, boulderBins.data(), boulderBins.size() * sizeof(tg_bin));
        std::memcpy(slot.hostBoulderEntries, boulderEntries.data(),
                    boulderEntries.size() * sizeof(std::uint32_t));
        Raise(h.batch_stage_boulders(slot.batch, s.boulders, slot.hostBoulderBins,
                                     boulderBins.size(), slot.hostBoulderEntries,
                                     boulderEntries.size(), 1, err, sizeof err),
              err);
    }
    std::uint64_t cells = 0;
    slot.offsets.resize(descs.size());
    for (std::size_t i = 0; i < descs.size(); ++i) {
        slot.offsets[i] = descs[i].out_offset;
        cells += std::uint64_t(descs[i].width) * descs[i].height;
    }
    float* region = s.pool + slot.poolOffset / sizeof(float);
    const double launched = Now();
    Raise(h.batch_run(slot.batch, s.plan, s.elevation, s.catalogue, region, cells, 0, err,
                      sizeof err),
          err);
    Raise(h.memcpy_d2h_async(s.dev, slot.staging, region, cells * sizeof(float), slot.stream, err,
                             sizeof err),
          err);
    Raise(h.event_record(s.dev, slot.event, slot.stream, err, sizeof err), err);
    return launched;
}

void FinishSlot(tgx_server& s, Slot& slot, int status, const std::string& error) {
    const double now = Now();
    for (Request* r : slot.requests) {
        r->status = status;
        r->error = error;
        r->info.done = now;
    }
    if (status == TG_OK) {
        s.stats.tiles += slot.requests.size();
    } else {
        s.stats.failed += slot.requests.size();
    }
    slot.requests.clear();
    s.ranges->Give(slot.poolOffset, slot.poolBytes);
    slot.busy = false;
    slot.copying = false;
}

// Takes waiting requests, in arrival order, for an idle slot: as many as fit the batch limits and
// the pool. False when none can be admitted now.
bool Admit(tgx_server& s, Slot& slot) {
    std::vector<Request*> take;
    std::uint64_t cells = 0;
    while (!s.pending.empty() && take.size() < s.config.max_batch_tiles) {
        Request* r = s.pending.front();
        const std::uint64_t c = std::uint64_t(r->tile.width) * r->tile.height;
        if (cells + c > s.config.max_batch_cells) {
            break;
        }
        take.push_back(r);
        cells += c;
        s.pending.pop_front();
    }
    std::uint64_t offset = 0;
    while (!take.empty() && !s.ranges->Take(AlignUp(cells * sizeof(float)), offset)) {
        Request* r = take.back();  // over the budget: the newest waits for memory
        take.pop_back();
        cells -= std::uint64_t(r->tile.width) * r->tile.height;
        s.pending.push_front(r);
    }
    if (take.empty()) {
        return false;
    }
    const double now = Now();
    const std::uint32_t id = s.nextBatch++;
    for (Request* r : take) {
        r->info.admitted = now;
        r->info.batch = id;
        r->info.batch_tiles = std::uint32_t(take.size());
    }
    slot.requests = std::move(take);
    slot.poolOffset = offset;
    slot.poolBytes = AlignUp(cells * sizeof(float));
    slot.busy = true;
    s.stats.batches += 1;
    s.stats.cells += cells;
    return true;
}

void Dispatch(tgx_server& s) {
    const HipApi& h = *s.hip;
    std::unique_lock<std::mutex> lock(s.mutex);
    for (;;) {
        bool progressed = false;
        // Hand finished batches to the copiers.
        for (Slot& slot : s.slots) {
            if (!slot.busy || slot.copying) {
                continue;
            }
            char err[512] = {0};
            const int q = h.event_query(s.dev, slot.event, err, sizeof err);
            if (q == TG_NOT_READY) {
                continue;
            }
            if (q == TG_OK) {
                slot.copying = true;
                s.copyQueue.push_back(&slot);
                s.copyWake.notify_one();
            } else {
                FinishSlot(s, slot, TG_ERR_DEVICE, err);
                s.done.notify_all();
            }
            progressed = true;
        }
        std::uint32_t inFlight = 0;
        for (const Slot& slot : s.slots) {
            inFlight += slot.busy ? 1u : 0u;
        }
        if (s.stop && s.pending.empty() && inFlight == 0) {
            return;
        }
        // Admit waiting requests into free slots.
        for (Slot& slot : s.slots) {
            if (slot.busy || s.pending.empty()) {
                continue;
            }
            if (!Admit(s, slot)) {
                break;
            }
            // Counted from the slots themselves: a copier may have returned a slot while the lock
            // was down for the previous launch.
            inFlight = 0;
            for (const Slot& other : s.slots) {
                inFlight += other.busy ? 1u : 0u;
            }
            s.stats.max_in_flight = std::max(s.stats.max_in_flight, inFlight);
            lock.unlock();
            std::string failure;
            double launched = 0.0;
            try {
                launched = LaunchBatch(s, slot);
            } catch (const std::exception& e) {
                failure = e.what();
            }
            lock.lock();
            // The slot's event is queried only by this thread, below, so its requests are still
            // the slot's here.
            for (Request* r : slot.requests) {
                r->info.launched = launched;
            }
            if (!failure.empty()) {
                // Nothing of the batch may still be queued: wait for the slot's stream first.
                char err[512] = {0};
                if (slot.stream != nullptr) {
                    (void)h.stream_synchronize(s.dev, slot.stream, err, sizeof err);
                }
                FinishSlot(s, slot, TG_ERR_DEVICE, failure);
                s.done.notify_all();
            }
            progressed = true;
        }
        if (progressed) {
            continue;
        }
        if (inFlight > 0) {
            s.wake.wait_for(lock, std::chrono::microseconds(s.config.poll_microseconds));
        } else {
            s.wake.wait(lock, [&] { return s.stop || !s.pending.empty(); });
        }
    }
}

// A copier: copies each finished slot's heights into its requests' memory, then returns the slot.
void Copy(tgx_server& s) {
    std::unique_lock<std::mutex> lock(s.mutex);
    for (;;) {
        s.copyWake.wait(lock, [&] { return s.copiersStop || !s.copyQueue.empty(); });
        if (s.copyQueue.empty()) {
            return;  // stopping, and nothing left
        }
        Slot& slot = *s.copyQueue.front();
        s.copyQueue.pop_front();
        lock.unlock();  // the slot and its requests' memory belong to this copier now
        for (std::size_t i = 0; i < slot.requests.size(); ++i) {
            const Request* r = slot.requests[i];
            std::memcpy(r->dst, slot.staging + slot.offsets[i],
                        std::size_t(r->tile.width) * r->tile.height * sizeof(float));
        }
        lock.lock();
        FinishSlot(s, slot, TG_OK, "");
        s.done.notify_all();
        s.wake.notify_all();  // a slot is free for the dispatcher
    }
}

void Fill(const Request& r, tgx_ticket_info* info) {
    if (info != nullptr) {
        *info = r.info;
    }
}

}  // namespace

}  // namespace tgh

extern "C" {

int tgx_server_create(tg_dev* dev, const tg_noise_plan* plan, float elevation,
                      const tg_catalogue* catalogue, const tgx_index* index,
                      const tgx_server_config* config, tgx_server** out, char* err,
                      size_t errLen) {
    if (out != nullptr) {
        *out = nullptr;
    }
    auto s = std::make_unique<tgx_server>();
    const int code = tgh::Guard(err, errLen, [&] {
        if (dev == nullptr || config == nullptr || out == nullptr) {
            throw std::invalid_argument("tgx_server_create: null argument");
        }
        const tgx_server_config& c = *config;
        // A full batch must always fit the empty pool, so admission can always make progress.
        if (c.slots == 0 || c.slots > 16 || c.max_batch_tiles == 0 || c.max_batch_cells == 0 ||
            c.max_batch_cells > (std::uint64_t(1) << 40) ||
            tgh::AlignUp(c.max_batch_cells * sizeof(float)) >
                c.pool_bytes / tgh::kAlign * tgh::kAlign) {
            throw std::invalid_argument(
                "tgx_server_create: need 1-16 slots, batch limits above zero and a pool that "
                "holds one full batch");
        }
        if (index != nullptr && catalogue == nullptr && tgx_index_count(index) > 0) {
            throw std::invalid_argument("tgx_server_create: an index needs its catalogue");
        }
        s->hip = &tgh::Hip();
        s->dev = dev;
        s->plan = plan;
        s->elevation = elevation;
        s->catalogue = catalogue;
        s->config = c;
        if (s->config.poll_microseconds == 0) {
            s->config.poll_microseconds = 20;
        }
        if (index == nullptr) {
            s->emptyIndex = tgh::MakeIndex(nullptr, nullptr, nullptr, nullptr, 0);
            index = s->emptyIndex.get();
        }
        s->index = index;
        char e[512] = {0};
        void* pool = nullptr;
        tgh::Raise(s->hip->malloc(dev, c.pool_bytes, &pool, e, sizeof e), e);
        s->pool = static_cast<float*>(pool);
        s->ranges = std::make_unique<tgh::RangePool>(c.pool_bytes / tgh::kAlign * tgh::kAlign);
        s->stats.pool_bytes = c.pool_bytes;
        s->slots.resize(c.slots);
        for (tgh::Slot& slot : s->slots) {
            tgh::Raise(s->hip->event_create(dev, &slot.event, e, sizeof e), e);
            void* staging = nullptr;
            tgh::Raise(s->hip->host_alloc(dev, c.max_batch_cells * sizeof(float), &staging, e,
                                          sizeof e),
                       e);
            slot.staging = static_cast<float*>(staging);
            // Room for a full batch of cells in 16 x 16 bins with ragged edges, and a few
            // craters a bin, so that growing (which keeps the old buffers) is rare.
            const std::size_t bins = std::size_t(c.max_batch_cells / 256 +
                                                 c.max_batch_cells / 1024 +
                                                 std::uint64_t(c.max_batch_tiles) * 64);
            tgh::EnsureCapacity(*s, slot, c.max_batch_tiles, bins, bins * 8);
        }
        s->dispatcher = std::thread([srv = s.get()] { tgh::Dispatch(*srv); });
        for (int i = 0; i < 2; ++i) {
            s->copiers.emplace_back([srv = s.get()] { tgh::Copy(*srv); });
        }
    });
    if (code != TG_OK) {
        if (s->hip != nullptr) {
            for (tgh::Slot& slot : s->slots) {
                tgh::FreeSlot(*s, slot);
            }
            if (s->pool != nullptr) {
                s->hip->free(dev, s->pool);
            }
        }
        return code;
    }
    *out = s.release();
    return TG_OK;
}

int tgx_server_boulders(tgx_server* s, const tg_boulder_catalogue* catalogue,
                        const tgx_index* index, char* err, size_t errLen) {
    return tgh::Guard(err, errLen, [&] {
        if (s == nullptr || catalogue == nullptr || index == nullptr) {
            throw std::invalid_argument("tgx_server_boulders: null argument");
        }
        if (tgh::IndexRadius(*index) != tgh::IndexRadius(*s->index)) {
            throw std::invalid_argument(
                "tgx_server_boulders: the boulders' index is on another surface than the craters'");
        }
        std::lock_guard<std::mutex> lock(s->mutex);
        if (s->nextTicket != 1 || s->boulders != nullptr) {
            throw std::invalid_argument(
                "tgx_server_boulders: set once, before the first request is submitted");
        }
        s->boulders = catalogue;
        s->boulderIndex = index;
    });
}

int tgx_server_submit(tgx_server* s, const tgx_tile* tile, float* dst, uint64_t* ticket,
                      char* err, size_t errLen) {
    return tgh::Guard(err, errLen, [&] {
        if (s == nullptr || tile == nullptr || dst == nullptr || ticket == nullptr) {
            throw std::invalid_argument("tgx_server_submit: null argument");
        }
        tgh::CheckTile(*tile, 0);
        const std::uint64_t cells = std::uint64_t(tile->width) * tile->height;
        if (cells > s->config.max_batch_cells) {
            throw std::invalid_argument("tgx_server_submit: the tile has more cells than a batch "
                                        "holds");
        }
        auto r = std::make_unique<tgh::Request>();
        r->tile = *tile;
        r->dst = dst;
        r->info.submitted = tgh::Now();
        std::lock_guard<std::mutex> lock(s->mutex);
        if (s->stop) {
            throw std::invalid_argument("tgx_server_submit: the server is shutting down");
        }
        r->ticket = s->nextTicket++;
        *ticket = r->ticket;
        s->pending.push_back(r.get());
        s->requests.emplace(r->ticket, std::move(r));
        s->wake.notify_one();
    });
}

int tgx_server_wait(tgx_server* s, uint64_t ticket, double timeout_seconds,
                    tgx_ticket_info* info, char* err, size_t errLen) {
    if (s == nullptr) {
        tgh::SetErr(err, errLen, "tgx_server_wait: null server");
        return TG_ERR_ARGUMENT;
    }
    std::unique_lock<std::mutex> lock(s->mutex);
    auto it = s->requests.find(ticket);
    if (it == s->requests.end()) {
        tgh::SetErr(err, errLen, "tgx_server_wait: unknown ticket");
        return TG_ERR_ARGUMENT;
    }
    tgh::Request* r = it->second.get();
    auto finished = [&] { return r->status != TG_NOT_READY; };
    if (timeout_seconds < 0.0) {
        s->done.wait(lock, finished);
    } else if (!s->done.wait_for(lock, std::chrono::duration<double>(timeout_seconds),
                                 finished)) {
        return TG_NOT_READY;
    }
    tgh::Fill(*r, info);
    const int status = r->status;
    if (status != TG_OK) {
        tgh::SetErr(err, errLen, r->error);
    }
    s->requests.erase(it);
    return status;
}

int tgx_server_stats_get(tgx_server* s, tgx_server_stats* out) {
    if (s == nullptr || out == nullptr) {
        return TG_ERR_ARGUMENT;
    }
    std::lock_guard<std::mutex> lock(s->mutex);
    *out = s->stats;
    out->pool_peak_bytes = s->ranges->Peak();
    out->pending = s->pending.size();
    return TG_OK;
}

int tgx_server_free(tgx_server* s) {
    if (s == nullptr) {
        return TG_OK;
    }
    {
        std::lock_guard<std::mutex> lock(s->mutex);
        s->stop = true;
        s->wake.notify_all();
    }
    if (s->dispatcher.joinable()) {
        s->dispatcher.join();
    }
    {
        std::lock_guard<std::mutex> lock(s->mutex);
        s->copiersStop = true;
        s->copyWake.notify_all();
    }
    for (std::thread& t : s->copiers) {
        t.join();
    }
    for (tgh::Slot& slot : s->slots) {
        tgh::FreeSlot(*s, slot);
    }
    for (tg_batch* b : s->retiredBatches) {
        s->hip->batch_free(b);
    }
    for (void* p : s->retiredPinned) {
        s->hip->host_free(s->dev, p);
    }
    if (s->pool != nullptr) {
        s->hip->free(s->dev, s->pool);
    }
    delete s;
    return TG_OK;
}

const char* tgx_hip_unavailable_reason(void) {
    static thread_local std::string reason;
    reason = tgh::HipUnavailableReason();
    return reason.c_str();
}

}  // extern "C"
