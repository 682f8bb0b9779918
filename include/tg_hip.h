/* SPDX-License-Identifier: BSD-3-Clause */
/*
 * libtg_hip.so — the C ABI of the MI355X (gfx950) hip backend.
 *
 * The reference reaches a GPU through one interface, `ComputeDevice` (src/gpu/device.h:49-64),
 * opened lazily per backend (src/gpu/device.h:66-124) and driven by the strip dispatcher
 * `RunKernel` (src/gpu/kernel.h:20-47, src/gpu/kernel.cpp:95-153). This header is that boundary
 * as plain C: no exceptions and no C++ or torch types cross it. Every call returns 0 on success,
 * or a nonzero TG_ERR_* code with a message written into `err` (at most `err_len` bytes,
 * NUL-terminated) when `err` is not null.
 *
 * Binary layouts of the kernel structs are in tg_layouts.h.
 */
#ifndef TG_HIP_H
#define TG_HIP_H

#include <stddef.h>
#include <stdint.h>

#include "tg_layouts.h"

#ifdef __cplusplus
extern "C" {
#endif

enum {
    TG_OK = 0,
    TG_ERR_ARGUMENT = 1, /* a null pointer, a bad size or an unknown kernel id */
    TG_ERR_DEVICE = 2,   /* the HIP runtime reported an error; the message is the runtime's */
    TG_ERR_NO_DEVICE = 3, /* no GPU at the ordinal */
    TG_NOT_READY = 4      /* tg_hip_event_query: the work before the event is still running */
};

/* The kernels of the reference the backend runs; one per kernel struct, as the reference keeps one
 * `KernelProgram` per struct (include/terraingen/gpu/program.h:14-23). */
enum {
    TG_KERNEL_NOISE_LAYER = 1, /* NoiseLayerKernel, src/layers/kernels/noise_layer.slang */
    TG_KERNEL_CRATERS = 2,     /* CraterShapeKernel, src/layers/kernels/craters.slang */
    TG_KERNEL_BOULDERS = 3,    /* tg_boulder_kernel: the dome relief BoulderLayer bakes on the
                                * CPU (boulder_relief.cpp:426-440); extras: boulders, tiles,
                                * entries, lows */
    TG_KERNEL_NOISE_BODY = 4,  /* NoiseLayerKernel on a body: heights += EvaluateFractals3 at each
                                * cell's body-fixed position (src/noise/surface_noise.cpp:9-20),
                                * which the reference runs on the CPU only (noise_layer.cpp:113);
                                * one extra, per cell: x, y, z as doubles. Every basis, fractal
                                * type, blend and 3D domain warp (fractals.h:221-269). */
    TG_KERNEL_CRATER_DATUMS = 5, /* tg_crater_datum_kernel, src/layers/kernels/crater_datums.slang:
                                  * a row per fitted crater, a column per lattice point of its
                                  * datum footprint, each adding KernelDatumRelief (craters.h:
                                  * 1266-1305); extras: records, data, band floats, band cells,
                                  * cells, entries, targets (datum_fit.cpp:306-314) */
    TG_KERNEL_CRATER_BODY_DATUMS = 6 /* tg_crater_body_datum_kernel, crater_body_datums.slang:
                                      * KernelBodyDatumRelief (craters.h:1313-1326) at each
                                      * lattice point on a body, no-data points left alone;
                                      * extras: records, data, frames, bounds, lists (first and
                                      * count per row), entries, positions (per cell, split
                                      * points) (datum_fit.cpp:561-565) */
};

typedef struct tg_dev tg_dev;

/* A buffer bound after the kernel struct, as `KernelBuffer` (src/gpu/kernel.h:18-33): host bytes,
 * the element count, and `per_cell` elements per grid cell for a buffer sliced per strip like the
 * heights (0 for a buffer bound whole). */
typedef struct tg_buf {
    const void* data;
    size_t bytes;
    size_t count;
    size_t per_cell;
} tg_buf;

/* Opens GPU `ordinal` once per process; later calls return the same device. Replaces
 * `Device::Open` behind `LazyDevice::Get` (src/gpu/device.h:79-90; CUDA: src/gpu/cuda.cpp).
 * The device is never destroyed, as the reference's never is. */
int tg_hip_open(int ordinal, tg_dev** out, char* err, size_t err_len);

/* The device name the runtime reports, as `ComputeDevice::DeviceName` (src/gpu/device.h:54). */
const char* tg_hip_device_name(const tg_dev* dev);

/* Whether this library has a form of kernel `kernel_id`, as `ComputeDevice::CanRun`
 * (src/gpu/device.h:57). */
int tg_hip_can_run(const tg_dev* dev, int kernel_id);

/* A no-op kept for symmetry: the device lives until the process ends (src/gpu/device.h:66-68). */
int tg_hip_close(tg_dev* dev);

/*
 * Runs kernel `kernel_id` over one strip of grid rows, synchronously, with host memory in and out:
 * `ComputeDevice::Run` for the binding `RunKernel` makes (src/gpu/kernel.cpp:95-126):
 *   heights  params->rows x params->width floats, updated in place (binding 0),
 *   xs       params->width column offsets from the origin (binding 1),
 *   ys       params->rows row offsets from the origin (binding 2),
 *   kernel   the kernel struct, `kernel_bytes` long (binding 3),
 *   extra    the further buffers in binding order (bindings 4...), for the crater kernel:
 *            records, data, tiles, entries, frames, points (per cell), lows
 *            (src/layers/craters/shaping/apply.cpp:181-188).
 * Device memory is allocated and freed per call, as the reference's devices do
 * (src/gpu/cuda.cpp:123-135, 297-314): this is the plain per-strip path. Runs are serialized
 * per device.
 */
int tg_hip_run(tg_dev* dev, int kernel_id, const void* kernel, size_t kernel_bytes, float* heights,
               const float* xs, const float* ys, const tg_kernel_params* params,
               const tg_buf* extra, size_t n_extra, char* err, size_t err_len);

/* ---- Device-resident form ------------------------------------------------------------------
 * The same kernels with every buffer already in device memory and launched on a caller's HIP
 * stream (0 for the device's default stream), without copies or synchronisation: the form the
 * batched path and the measurements use. A `tg_noise_plan` is a NoiseLayerKernel with its
 * cell-invariant octave quantities worked out (seeds, offsets, frequencies, turns, weights),
 * resident in device memory until freed. No counterpart in the reference, which only has the
 * synchronous form above. */
typedef struct tg_noise_plan tg_noise_plan;

int tg_hip_noise_plan_create(tg_dev* dev, const tg_noise_layer_kernel* kernel,
                             tg_noise_plan** out, char* err, size_t err_len);
int tg_hip_noise_plan_free(tg_noise_plan* plan);

/* heights[r * width + c] += stack(xs[c], ys[r]) for r < params->rows, skipping no-data cells:
 * noise_layer.slang's main() with device pointers. */
int tg_hip_noise_launch(tg_dev* dev, const tg_noise_plan* plan, float* heights_dev,
                        const float* xs_dev, const float* ys_dev, const tg_kernel_params* params,
                        void* stream, char* err, size_t err_len);

/* SyntheticSource::NoiseAt (src/sources/synthetic_source.cpp:76-80) on the device:
 * heights[r * width + c] = elevation + stack(xs[c], ys[r]), writing every cell of the strip. The
 * reference evaluates this source on the CPU only. */
int tg_hip_noise_fill(tg_dev* dev, const tg_noise_plan* plan, float elevation, float* heights_dev,
                      const float* xs_dev, const float* ys_dev, const tg_kernel_params* params,
                      void* stream, char* err, size_t err_len);

/* A crater batch resident in device memory: the kernel struct, records, data, tile heads and
 * entries, frames and per-column/per-row remainders, bound as craters.slang binds them
 * (src/layers/craters/shaping/apply.cpp:181-188). `lows` holds `width` column remainders then one
 * per grid row. Frames may be null on a plane. */
typedef struct tg_crater_set tg_crater_set;
typedef struct tg_boulder_set tg_boulder_set;

int tg_hip_crater_set_create(tg_dev* dev, const tg_crater_shape* kernel,
                             const tg_crater_record* records, size_t record_count,
                             const float* data, size_t data_count, const uint32_t* tiles,
                             size_t tile_words, const uint32_t* entries, size_t entry_count,
                             const tg_crater_frame* frames, size_t frame_count, const float* lows,
                             size_t low_count, tg_crater_set** out, char* err, size_t err_len);
int tg_hip_crater_set_free(tg_crater_set* set);

/* craters.slang's main() on a plane strip with device pointers: heights_dev (params->rows x
 * params->width) updated in place, xs_dev per column, ys_dev per strip row. */
int tg_hip_crater_launch(tg_dev* dev, const tg_crater_set* set, float* heights_dev,
                         const float* xs_dev, const float* ys_dev, const tg_kernel_params* params,
                         void* stream, char* err, size_t err_len);
/* tg_hip_crater_launch with a boulder set of the same grid drawn in the same pass (no counterpart
 * in the reference, which bakes boulders on the CPU after its crater pass): each cell takes its
 * craters, then the boulders of its 16 x 16 tile over the craters' height as tg_hip_boulder_launch
 * draws them, read and written once. The same heights as the two launches one after the other. */
int tg_hip_crater_boulder_launch(tg_dev* dev, const tg_crater_set* set,
                                 const tg_boulder_set* boulders, float* heights_dev,
                                 const float* xs_dev, const float* ys_dev,
                                 const tg_kernel_params* params, void* stream, char* err,
                                 size_t err_len);

/* ---- Sources with posts ---------------------------------------------------------------------
 * Upsampler::Fill of a KernelUpsampler (src/sources/upsampling.cpp:297-335), which the reference
 * runs on the CPU for every source that holds posts (SyntheticSource with a pixel size,
 * src/sources/synthetic_source.cpp:36-53,87-95; GeoTiffSource; RasterSource): fills `heights_dev`
 * (spec->width x spec->height floats, row by row) from the posts in device memory
 * (spec->window_width x spec->window_height floats, NaN where a post has no data), on `stream`.
 * Each cell is the kernel's FP64 sum over its taps narrowed to FP32, NaN where a tap of non-zero
 * weight has no data; with spec->at_post_spacing a cell centred on a post takes the post. The
 * heights are the reference's bit for bit for the nearest, bilinear, cubic and B-spline kernels;
 * Lanczos's weights take the device's sine, within an ulp of the C library's. */
int tg_hip_upsample(tg_dev* dev, const tg_upsample* spec, const float* posts_dev,
                    float* heights_dev, void* stream, char* err, size_t err_len);

/* ---- Batched tiles ---------------------------------------------------------------------------
 * Many tiles of uneven size and pixel size evaluated in one persistent launch: the source fill
 * (SyntheticSource noise, or a constant) and crater shaping fused per 16 x 16 bin of cells, bins
 * taken from a work queue by whichever workgroup is free. No counterpart in the reference, which
 * evaluates one grid at a time through RunKernel. */

/* A crater catalogue resident on the device: the kernel struct, one record per crater in apply
 * order, the data, and each crater's FP64 diameter (its ResolutionWeight is worked out per tile
 * from the tile's pixel size, src/layers/fade.cpp:48-52). */
/* Boulders resident on the device for one grid spacing (no counterpart in the reference, which
 * bakes boulders on the CPU): the parameter block, the boulders (tgx_prepare_boulders), their
 * 16 x 16 tile lists (tgx_bin_plane over their relief boxes) and the grid's lows, as
 * tg_hip_run(TG_KERNEL_BOULDERS) binds them. tg_hip_boulder_launch bakes them into heights on the
 * device (each cell's height on entry is its base), on `stream`. */
int tg_hip_boulder_set_create(tg_dev* dev, const tg_boulder_kernel* kernel,
                              const tg_boulder* boulders, size_t count, const uint32_t* tiles,
                              size_t tile_words, const uint32_t* entries, size_t entry_count,
                              const float* lows, size_t low_count, tg_boulder_set** out,
                              char* err, size_t err_len);
int tg_hip_boulder_set_free(tg_boulder_set* set);
int tg_hip_boulder_launch(tg_dev* dev, const tg_boulder_set* set, float* heights_dev,
                          const float* xs_dev, const float* ys_dev, const tg_kernel_params* params,
                          void* stream, char* err, size_t err_len);

typedef struct tg_catalogue tg_catalogue;

int tg_hip_catalogue_create(tg_dev* dev, const tg_crater_shape* kernel,
                            const tg_crater_record* records, size_t record_count,
                            const float* data, size_t data_count, const double* diameters,
                            tg_catalogue** out, char* err, size_t err_len);
int tg_hip_catalogue_free(tg_catalogue* catalogue);
/* The crater frames (MakeCraterFrame, crater_kernel.cpp:285-301; tgx_sphere_frames makes them)
 * of a catalogue for tiles on a sphere: one per crater, in catalogue order, set once. */
int tg_hip_catalogue_frames(tg_catalogue* catalogue, const tg_crater_frame* frames, size_t count,
                            char* err, size_t err_len);

/* One tile: a north-up grid `width` x `height` of `pixel` meters, west and north edges in FP64
 * (cell centres are worked out in FP64 on the device and narrowed as RunKernel narrows them,
 * src/gpu/kernel.cpp:55-70), its heights at `out_offset` floats into the output, row by row, its
 * bins from `bin_first`, `bins_across` per row of bins. */
typedef struct tg_tile_desc {
    double origin_x;
    double origin_y;
    double pixel;
    uint32_t width;
    uint32_t height;
    uint64_t out_offset;
    uint32_t bin_first;
    uint32_t bins_across;
} tg_tile_desc;

/* A bin of 16 x 16 cells of tile `tile`: the catalogue craters entries[first .. first+count), in
 * catalogue (apply) order. */
typedef struct tg_bin {
    uint32_t tile;
    uint32_t first;
    uint32_t count;
    uint32_t pad;
} tg_bin;

typedef struct tg_batch tg_batch;

/* Device buffers for batches of at most `max_tiles` tiles, `max_bins` bins, `max_entries`
 * entries, with their own stream. */
int tg_hip_batch_create(tg_dev* dev, size_t max_tiles, size_t max_bins, size_t max_entries,
                        tg_batch** out, char* err, size_t err_len);
int tg_hip_batch_free(tg_batch* batch);

/* Copies a batch's tiles, bins and entries into its buffers, checking every index against the
 * others and the capacities. */
int tg_hip_batch_stage(tg_batch* batch, const tg_tile_desc* tiles, size_t tile_count,
                       const tg_bin* bins, size_t bin_count, const uint32_t* entries,
                       size_t entry_count, char* err, size_t err_len);
/* tg_hip_batch_stage with the copies queued on the batch's stream instead of waited for (no
 * counterpart in the reference): for a serving loop that bins the next batch while others run.
 * The host arrays must stay unchanged until the batch's next run has completed, and should be
 * pinned (tg_hip_host_alloc) for the copies to be asynchronous. The batch must be idle: a batch
 * whose stream still has work queued is refused. */
int tg_hip_batch_stage_async(tg_batch* batch, const tg_tile_desc* tiles, size_t tile_count,
                             const tg_bin* bins, size_t bin_count, const uint32_t* entries,
                             size_t entry_count, char* err, size_t err_len);

/* Stages a batch binned on the device (no counterpart in the reference, which bins one grid on
 * the CPU): the tiles (descriptors for bins of `bin_side` cells, 8, 16 or 32, as
 * tgx_index_windows makes them) and each tile's crater windows ((first, count) per tile in
 * `tile_windows`, the windows in catalogue order). One wave per bin tests its tile's windows 64 at
 * a time against the bin's cells and ballots the hits; the counts' exclusive sum places each bin's
 * list and a second pass writes it, in window order. With 16-cell bins the bins and entries are
 * those tgx_index_bin gives. Every window is checked on the host; the entry count is read back
 * to check the batch's capacity and returned in `entry_count` (may be null). Synchronous. */
int tg_hip_batch_stage_windows(tg_batch* batch, const tg_tile_desc* tiles, size_t tile_count,
                               uint32_t bin_side, const uint32_t* tile_windows,
                               const tg_window* windows, size_t window_count, size_t* entry_count,
                               char* err, size_t err_len);
/* The staged batch's bin and entry counts and, given host arrays at least that long, its bins and
 * entries (no counterpart in the reference). */
int tg_hip_batch_read_bins(tg_batch* batch, tg_bin* bins, size_t bin_capacity, uint32_t* entries,
                           size_t entry_capacity, size_t* bin_count, size_t* entry_count,
                           char* err, size_t err_len);
/* Entries the culled crater walk tests per pass: 32, 64 (the default) or 128 (each lane testing
 * two before the ballots). A launch choice for the autotuner, as tg_hip_batch_tune's: every depth
 * gives the same heights. */
int tg_hip_batch_cull_depth(tg_batch* batch, uint32_t depth, char* err, size_t err_len);

/* Evaluates the staged batch into `out_dev` (at least `out_capacity` floats): each cell's height
 * is `elevation` plus the plan's noise (none when `plan` is null) at its centre, then the
 * catalogue's craters of its bin (none when `catalogue` is null). With `use_graph` the launch is
 * captured once as a HIP graph for these arguments and replayed. Enqueued on the batch's stream;
 * tg_hip_batch_wait waits for it. */
int tg_hip_batch_run(tg_batch* batch, const tg_noise_plan* plan, float elevation,
                     const tg_catalogue* catalogue, float* out_dev, size_t out_capacity,
                     int use_graph, char* err, size_t err_len);
/* Launch choices of a batch (no counterpart in the reference), for the autotuner. The crater
 * walk: TG_WALK_TOGETHER, where a wave applies a listed crater whenever any of its cells is in
 * reach; TG_WALK_PER_LANE, where each cell walks the list to its own next crater in reach; or
 * TG_WALK_CULLED (the default), where a wave first tests 64 entries at once against the box of
 * its cells and steps only through those that can reach one. Workgroups per CU of the persistent
 * launch: 1-64 (default 8). `hoisted` (default 1) uses the per-crater constants of the MPD profile
 * path where a crater allows it; 0 evaluates every crater in full. `chunk`: bins a wave takes
 * from its queue per atomic add, 1-256 (default 8; fewer on batches too small to give every wave
 * four takes). Every choice gives the same heights. */
enum { TG_WALK_TOGETHER = 0, TG_WALK_PER_LANE = 1, TG_WALK_CULLED = 2 };
int tg_hip_batch_tune(tg_batch* batch, uint32_t crater_walk, uint32_t groups_per_cu,
                      uint32_t hoisted, uint32_t chunk, char* err, size_t err_len);
/* The surface of a batch's tiles: 0 (the default) for a plane, else a sphere of `sphere_radius`
 * meters whose tiles are longitude-latitude rasters (a tile's origin_x, origin_y and pixel are its
 * west and north edges and cell size in degrees). On a sphere each cell's noise is the stack's at
 * its body-fixed position (EvaluateFractals3, src/noise/surface_noise.cpp:9-20), craters
 * are placed by their frames (tg_hip_catalogue_frames) and weighted at the tile's spacing, pixel
 * times radius times pi / 180 (RasterGrid::Spacing, src/raster/grid.cpp:28-33). The bins come
 * from tgx_sphere_bin. The crater walk is always the culled one. */
int tg_hip_batch_surface(tg_batch* batch, double sphere_radius, char* err, size_t err_len);
/* Boulders resident for batched tiles (no counterpart in the reference, which bakes boulders on
 * the CPU one grid at a time): each boulder's record (tgx_prepare_boulders at the finest spacing
 * of the tiles it will be drawn into; its centre split hi + lo, radius, exposed height and facets
 * are read, its reach and weight are not), its FP64 diameter, and on a sphere its frame
 * (tgx_sphere_frames of its centre; null on a plane). A tile draws a boulder as BoulderLayer::
 * PrepareRelief (src/layers/boulders/boulder_relief.cpp:373-423) prepares it for that tile's
 * grid: at the resolution weight of the tile's spacing (src/layers/fade.cpp:48-52), its dome
 * averaged over a band one spacing wide and reaching half of (diameter + spacing). */
typedef struct tg_boulder_catalogue tg_boulder_catalogue;
int tg_hip_boulder_catalogue_create(tg_dev* dev, const tg_boulder* boulders, size_t count,
                                    const double* diameters, const tg_crater_frame* frames,
                                    tg_boulder_catalogue** out, char* err, size_t err_len);
int tg_hip_boulder_catalogue_free(tg_boulder_catalogue* catalogue);
/* The boulder layer of the staged batch: per bin of the staged tiles (16 x 16 cells, one per
 * staged bin, in bin order, `tile` as the crater bins'), the catalogue's boulders
 * entries[first .. first+count) that may reach it (tgx_index_bin over the boulders' relief boxes,
 * for instance). The batch's next runs draw them after each cell's craters, as a BoulderLayer
 * after a CraterLayer: every cell keeps its height after the craters as the base and takes the
 * highest of that and base + weight x relief over its boulders (BoulderLayer::AddRelief without a
 * fillet, boulder_relief.cpp:244-262), on a plane or, with frames, on the batch's sphere. Staging
 * the tiles again drops the boulders. With `async` the copies are queued on the batch's stream
 * after what it holds, as tg_hip_batch_stage_async queues them (the host arrays, pinned for the
 * copies to be asynchronous, must stay unchanged until the next run completes; a list longer
 * than any before waits for the stream, to replace its buffer); else they are waited for. Not
 * for tg_hip_batch_run_static. */
int tg_hip_batch_stage_boulders(tg_batch* batch, const tg_boulder_catalogue* catalogue,
                                const tg_bin* bins, size_t bin_count, const uint32_t* entries,
                                size_t entry_count, int async, char* err, size_t err_len);
int tg_hip_batch_wait(tg_batch* batch, char* err, size_t err_len);
/* The batch's stream, for events and copies ordered with its runs. */
void* tg_hip_batch_stream(tg_batch* batch);

/* tg_hip_batch_stage then tg_hip_batch_run, for one batch, without allocating anything: no
 * tg_batch, no hipMalloc or hipFree, no allocating library call. No counterpart in the reference;
 * it is for callers that cannot allocate device memory (tight serving loops, harnesses that
 * forbid it). The tiles, bins and entries (host memory, checked as tg_hip_batch_stage checks them)
 * are copied into the library's static scratch on this device, which also holds the work queue's
 * counters, and the batch is evaluated into `out_dev` on `stream` (0: the null stream), on a
 * sphere of `sphere_radius` meters when it is positive (as tg_hip_batch_surface), with the default
 * launch choices and no graph. The heights are those tg_hip_batch_run gives. Limits: at most
 * TG_STATIC_MAX_TILES tiles, TG_STATIC_MAX_BINS bins and TG_STATIC_MAX_ENTRIES entries; larger
 * batches are refused with an error. The host arrays must stay unchanged until the run has
 * completed. Every call on a device shares that device's scratch, which keeps the last call's
 * contents: calls that overlap on one device (on different streams, or from several threads)
 * race, and must be ordered by the caller. */
#define TG_STATIC_MAX_TILES 4096
#define TG_STATIC_MAX_BINS 262144
#define TG_STATIC_MAX_ENTRIES 2097152
int tg_hip_batch_run_static(tg_dev* dev, const tg_tile_desc* tiles, size_t tile_count,
                            const tg_bin* bins, size_t bin_count, const uint32_t* entries,
                            size_t entry_count, const tg_noise_plan* plan, float elevation,
                            const tg_catalogue* catalogue, double sphere_radius, float* out_dev,
                            size_t out_capacity, void* stream, char* err, size_t err_len);

/* ---- Device memory, streams and timing --------------------------------------------------
 * No counterpart in the reference. For callers without a HIP runtime of their own (the tests,
 * the measurements, hosts written in other languages): one process should hold one HIP runtime,
 * so these go through this library's. Streams are created at the default priority. */
int tg_hip_malloc(tg_dev* dev, size_t bytes, void** out, char* err, size_t err_len);
int tg_hip_free(tg_dev* dev, void* ptr);
int tg_hip_memcpy_h2d(tg_dev* dev, void* dst, const void* src, size_t bytes, char* err,
                      size_t err_len);
int tg_hip_memcpy_d2h(tg_dev* dev, void* dst, const void* src, size_t bytes, char* err,
                      size_t err_len);
int tg_hip_memset(tg_dev* dev, void* dst, int value, size_t bytes, char* err, size_t err_len);
/* dst[i] = value for i < count, on `stream`: FlatSource's fill (src/sources/flat_source.cpp:13-31)
 * of a height buffer on the device, ordered with the stream's other work. */
int tg_hip_fill(tg_dev* dev, float* dst, float value, size_t count, void* stream, char* err,
                size_t err_len);
/* Device-to-device copy on `stream` (the runtime's copy kernel). */
int tg_hip_memcpy_d2d_async(tg_dev* dev, void* dst, const void* src, size_t bytes, void* stream,
                            char* err, size_t err_len);
int tg_hip_stream_create(tg_dev* dev, void** out, char* err, size_t err_len);
int tg_hip_stream_destroy(tg_dev* dev, void* stream);
int tg_hip_stream_synchronize(tg_dev* dev, void* stream, char* err, size_t err_len);
int tg_hip_event_create(tg_dev* dev, void** out, char* err, size_t err_len);
int tg_hip_event_destroy(tg_dev* dev, void* event);
int tg_hip_event_record(tg_dev* dev, void* event, void* stream, char* err, size_t err_len);
/* Milliseconds between two recorded events, waiting for the second. */
int tg_hip_event_elapsed(tg_dev* dev, void* start, void* stop, float* ms, char* err,
                         size_t err_len);
/* The vector-ALU issue ceiling, measured (no counterpart in the reference; a measurement aid):
 * `blocks` workgroups of 256 lanes each run `iters` iterations of independent chains, kind 0 of
 * eight FP32 FMAs, kind 1 of eight 32-bit xors and eight adds. `ms` is the timed launch (after a
 * warm-up one), `lane_ops` the lane operations it issued. */
int tg_hip_valu_probe(tg_dev* dev, int kind, uint32_t iters, uint32_t blocks, float* ms,
                      double* lane_ops, char* err, size_t err_len);
/* The batched kernels' output stores alone, for reading the HBM bytes a store shape writes (no
 * counterpart in the reference; a measurement aid): `tiles` tiles of width x width floats packed
 * into `out` (`out_capacity` floats), 16 x 16 bins in runs of consecutive bins per wave, each bin
 * written as four 8 x 8 squares (`shape` 0, the batched kernels' stores), four 16 x 4 strips (1)
 * or 64-float rows four bins across (2). `ms` is the launch, synchronous. */
int tg_hip_store_probe(tg_dev* dev, int shape, uint32_t width, uint32_t tiles, float* out,
                       size_t out_capacity, float* ms, char* err, size_t err_len);

/* Pinned (page-locked) host memory, for copies that overlap device work. */
int tg_hip_host_alloc(tg_dev* dev, size_t bytes, void** out, char* err, size_t err_len);
int tg_hip_host_free(tg_dev* dev, void* ptr);
/* Asynchronous copies on `stream`: the host memory must be pinned (tg_hip_host_alloc) for the copy
 * to overlap other work, and must stay valid until the stream has passed the copy. */
int tg_hip_memcpy_d2h_async(tg_dev* dev, void* dst, const void* src, size_t bytes, void* stream,
                            char* err, size_t err_len);
int tg_hip_memcpy_h2d_async(tg_dev* dev, void* dst, const void* src, size_t bytes, void* stream,
                            char* err, size_t err_len);
/* TG_OK once the work recorded before `event` has finished, TG_NOT_READY while it runs. */
int tg_hip_event_query(tg_dev* dev, void* event, char* err, size_t err_len);

#ifdef __cplusplus
}
#endif

#endif /* TG_HIP_H */
