/* SPDX-License-Identifier: BSD-3-Clause */
/*
 * libterraingen_amd.so — the C++ host of the hip backend, as a C API for hosts in other languages
 * (Python tests and measurements use it through ctypes). It covers what sits above the kernel
 * boundary of include/tg_hip.h on the hot path: binning craters into tiles, crater and boulder
 * placement, crater preparation, and the batched tile server. The C++ classes behind it mirror the
 * reference's names (Backend, ComputeDevice, RunKernel, ...).
 *
 * Every call returns 0 on success or a nonzero code with a message in `err` (at most `err_len`
 * bytes) when `err` is not null. Arrays this library returns are freed with tgx_free.
 */
#ifndef TERRAINGEN_AMD_H
#define TERRAINGEN_AMD_H

#include <stddef.h>
#include <stdint.h>

#include "tg_layouts.h"

#ifdef __cplusplus
extern "C" {
#endif

/* A north-up grid: `width` x `height` cells of `pixel` meters, its west edge at `origin_x` and its
 * north edge at `origin_y` (GeoTransform::NorthUp, include/terraingen/raster/geo_transform.h). */
typedef struct tgx_grid {
    uint32_t width;
    uint32_t height;
    double origin_x;
    double origin_y;
    double pixel;
} tgx_grid;

void tgx_free(void* p);

/* How far a crater's outline can reach over its round radius, `CraterOutline::MaxScale`
 * (src/layers/craters/shaping/irregularity.cpp:229-238), from its record and data. */
double tgx_outline_max_scale(const tg_crater_record* record, const float* data);

/*
 * Bins craters into the tiles of `side` x `side` cells their footprints cover, in crater order:
 * the footprint is the box `reach[n]` meters around (cx[n], cy[n]) (apply.cpp:28-32), its cells
 * the grid's pixel window (src/raster/cell_loops.cpp:20-27), the tiles those of
 * ForEachTileIn. Outputs the `tiles` block the crater kernel reads ([tiles across, side, 0, then
 * each tile's first entry and count], craters.h:1341-1344) and the entries (never empty), as
 * TileLists and FlattenKernelLists lay them out (apply.cpp:119-140, crater_kernel.cpp:608-620).
 */
int tgx_bin_plane(const tgx_grid* grid, const double* cx, const double* cy, const double* reach,
                  size_t n, uint32_t side, uint32_t** tiles, size_t* tile_words,
                  uint32_t** entries, size_t* entry_count, char* err, size_t err_len);

#ifdef __cplusplus
}
#endif

#endif /* TERRAINGEN_AMD_H */
