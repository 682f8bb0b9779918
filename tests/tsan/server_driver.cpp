// SPDX-License-Identifier: BSD-3-Clause
//
// TEST INFRASTRUCTURE ONLY. Drives the tile server (terraingen_amd/csrc/host/server.cpp) over
// the CPU stand-in of the kernel library (hip_standin.cpp), built with ThreadSanitizer: submitter
// threads send tiles of many sizes, wait for them in shuffled order (some first with a zero
// timeout), and check every cell against the stand-in's height for that tile, cell and bin, with
// the bins taken from tgx_index_bin of the tile alone. A boulder layer is set, so the boulder
// lists travel through the server too. Exits 0 when every tile is its own and the statistics add
// up; ThreadSanitizer reports any data race it sees on the way.

#include "terraingen_amd.h"

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <thread>
#include <vector>

extern "C" float standin_height(const tg_tile_desc* t, uint32_t col, uint32_t row, uint32_t craters,
                                uint32_t boulders, float elevation);

namespace {

constexpr float kElevation = 3.5f;

tgx_index* MakeIndex(std::mt19937_64& rng, size_t n, double side, double minReach,
                     double maxReach) {
    std::uniform_real_distribution<double> at(0.0, side), reach(minReach, maxReach);
    std::vector<double> cx(n), cy(n), r(n), d(n);
    for (size_t i = 0; i < n; ++i) {
        cx[i] = at(rng);
        cy[i] = at(rng);
        r[i] = reach(rng);
        d[i] = r[i] / 1.5;
    }
    char err[512] = {0};
    tgx_index* index = nullptr;
    if (tgx_index_create(cx.data(), cy.data(), r.data(), d.data(), n, &index, err, sizeof err) !=
        TG_OK) {
        std::fprintf(stderr, "index: %s\n", err);
        std::exit(2);
    }
    return index;
}

// The bin counts of `tile` alone, as the server's batch bins it.
struct Alone {
    std::vector<tg_bin> bins;
    tg_tile_desc desc{};
};

Alone BinAlone(const tgx_index* index, const tgx_tile& tile) {
    char err[512] = {0};
    tg_tile_desc* descs = nullptr;
    tg_bin* bins = nullptr;
    uint32_t* entries = nullptr;
    size_t nb = 0, ne = 0;
    if (tgx_index_bin(index, &tile, 1, 1, &descs, &bins, &nb, &entries, &ne, err, sizeof err) !=
        TG_OK) {
        std::fprintf(stderr, "bin: %s\n", err);
        std::exit(2);
    }
    Alone out;
    out.desc = descs[0];
    out.bins.assign(bins, bins + nb);
    tgx_free(descs);
    tgx_free(bins);
    tgx_free(entries);
    return out;
}

struct Job {
    tgx_tile tile{};
    std::vector<float> heights;
    uint64_t ticket = 0;
};

}  // namespace

int main() {
    char err[512] = {0};
    tg_dev* dev = nullptr;
    if (tg_hip_open(0, &dev, err, sizeof err) != TG_OK) {
        std::fprintf(stderr, "open: %s\n", err);
        return 2;
    }
    std::mt19937_64 rng(7);
    const double side = 2000.0;
    tgx_index* craters = MakeIndex(rng, 3000, side, 1.0, 40.0);
    tgx_index* boulders = MakeIndex(rng, 5000, side, 0.2, 3.0);
    // The stand-in never reads the catalogues: any non-null handles do.
    static int catalogueTag = 0, boulderTag = 0;
    auto* catalogue = reinterpret_cast<const tg_catalogue*>(&catalogueTag);
    auto* rocks = reinterpret_cast<const tg_boulder_catalogue*>(&boulderTag);

    tgx_server_config config{};
    config.slots = 3;
    config.max_batch_tiles = 6;
    config.max_batch_cells = 1u << 18;
    config.pool_bytes = 8ull << 20;
    config.bin_threads = 2;
    config.poll_microseconds = 20;
    tgx_server* server = nullptr;
    if (tgx_server_create(dev, nullptr, kElevation, catalogue, craters, &config, &server, err,
                          sizeof err) != TG_OK) {
        std::fprintf(stderr, "server: %s\n", err);
        return 2;
    }
    if (tgx_server_boulders(server, rocks, boulders, err, sizeof err) != TG_OK) {
        std::fprintf(stderr, "boulders: %s\n", err);
        return 2;
    }

    constexpr int kThreads = 4;
    constexpr int kPerThread = 60;
    std::atomic<int> bad{0};
    std::vector<std::thread> threads;
    for (int t = 0; t < kThreads; ++t) {
        threads.emplace_back([&, t] {
            std::mt19937_64 local(100 + t);
            std::uniform_real_distribution<double> at(0.0, side * 0.9), pixel(0.05, 2.0);
            std::uniform_int_distribution<uint32_t> edge(9, 300);
            std::vector<Job> jobs(kPerThread);
            for (Job& j : jobs) {
                j.tile = tgx_tile{at(local), side - at(local), pixel(local), edge(local),
                                  edge(local)};
                j.heights.assign(size_t(j.tile.width) * j.tile.height, -1.0f);
                char e[512] = {0};
                if (tgx_server_submit(server, &j.tile, j.heights.data(), &j.ticket, e,
                                      sizeof e) != TG_OK) {
                    std::fprintf(stderr, "submit: %s\n", e);
                    bad++;
                }
            }
            std::shuffle(jobs.begin(), jobs.end(), local);
            for (size_t k = 0; k < jobs.size(); ++k) {
                Job& j = jobs[k];
                char e[512] = {0};
                tgx_ticket_info info{};
                if (k % 3 == 0) {
                    const int early = tgx_server_wait(server, j.ticket, 0.0, &info, e, sizeof e);
                    if (early == TG_OK) {
                        goto check;
                    }
                    if (early != TG_NOT_READY) {
                        std::fprintf(stderr, "wait(0): %s\n", e);
                        bad++;
                        continue;
                    }
                }
                if (tgx_server_wait(server, j.ticket, -1.0, &info, e, sizeof e) != TG_OK) {
                    std::fprintf(stderr, "wait: %s\n", e);
                    bad++;
                    continue;
                }
            check:
                if (!(info.submitted <= info.admitted && info.admitted <= info.launched &&
                      info.launched <= info.done)) {
                    std::fprintf(stderr, "ticket %llu: times out of order\n",
                                 static_cast<unsigned long long>(j.ticket));
                    bad++;
                }
                const Alone crater = BinAlone(craters, j.tile);
                const Alone boulder = BinAlone(boulders, j.tile);
                for (uint32_t row = 0; row < j.tile.height; ++row) {
                    for (uint32_t col = 0; col < j.tile.width; ++col) {
                        const size_t bin = (row / 16u) * crater.desc.bins_across + col / 16u;
                        const float want =
                            standin_height(&crater.desc, col, row, crater.bins[bin].count,
                                           boulder.bins[bin].count, kElevation);
                        const float got = j.heights[size_t(row) * j.tile.width + col];
                        if (got != want) {
                            std::fprintf(stderr, "ticket %llu cell (%u, %u): %g, want %g\n",
                                         static_cast<unsigned long long>(j.ticket), col, row,
                                         double(got), double(want));
                            bad++;
                            row = j.tile.height;
                            break;
                        }
                    }
                }
            }
        });
    }
    for (std::thread& t : threads) {
        t.join();
    }
    tgx_server_stats stats{};
    tgx_server_stats_get(server, &stats);
    if (stats.tiles != uint64_t(kThreads) * kPerThread || stats.failed != 0 ||
        stats.max_in_flight > config.slots || stats.max_in_flight == 0) {
        std::fprintf(stderr, "stats: tiles %llu failed %llu max_in_flight %u\n",
                     static_cast<unsigned long long>(stats.tiles),
                     static_cast<unsigned long long>(stats.failed), stats.max_in_flight);
        bad++;
    }
    tgx_server_free(server);
    tgx_index_free(craters);
    tgx_index_free(boulders);
    std::printf("tiles %llu batches %llu max_in_flight %u bad %d\n",
                static_cast<unsigned long long>(stats.tiles),
                static_cast<unsigned long long>(stats.batches), stats.max_in_flight, bad.load());
    return bad.load() == 0 ? 0 : 1;
}
