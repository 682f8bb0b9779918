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

#include "tg_hip.h"
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

/* A placed crater (the columns of CraterList, include/terraingen/layers/craters/crater.h). */
typedef struct tgx_crater {
    double x;
    double y;
    double diameter;
    double age_gyr;
    uint64_t id;
} tgx_crater;

/* Generation of craters over the region [min_x, max_x) x [min_y, max_y) meters: GenerateCraters
 * (src/layers/craters/generation/population.cpp:414-424) with CraterGenerationOptions'
 * minDiameter, maxDiameter, marginRadii, cellSizeFactor (0 = 16) and minResolvedDiameter.
 * `model` 0: a power-law density N(>=D) = sfd_coefficient D^-sfd_exponent per m^2 and ages uniform
 * over lifetime_gyr (0 = fresh). `model` 1: the density and ages of DefaultCraterModels
 * (crater_layer.cpp:194-203): the lunar equilibrium surface (Neukum production over 3.5 Gyr
 * capped by Hartmann's equilibrium) with steady-state ages; the sfd and lifetime fields are
 * ignored. */
typedef struct tgx_generation {
    double min_x, min_y, max_x, max_y;
    double sfd_coefficient, sfd_exponent;
    double min_diameter, max_diameter;
    double margin_radii;
    double cell_size_factor;
    double min_resolved_diameter;
    double lifetime_gyr;
    uint32_t seed;
    uint32_t model;
    /* 0: a plane (the region in meters); else a sphere of this radius in meters, the region in
     * longitude and latitude degrees, cells the reference's cube-sphere cells. */
    double sphere_radius;
    /* Cells drawn on this many threads (0: every hardware thread); the list is the same on any. */
    uint32_t threads;
    uint32_t pad;
} tgx_generation;

/* The threads host routines (placement, preparation, binning) run on when their own count is 0:
 * 0 (the default) for every hardware thread. Results never depend on it. No counterpart in the
 * reference, whose routines take a thread count each (ParallelFor, src/util/parallel.cpp:126). */
void tgx_set_host_threads(uint32_t threads);

/* Philox4x32-10 of a 4-word counter under a 2-word key (include/terraingen/random/philox.h:79-91),
 * the generator every placement draw comes from. */
void tgx_philox4x32(const uint32_t* counter, const uint32_t* key, uint32_t* out);

/* The craters in band, cell and index order, bit-identical to the reference's. */
int tgx_generate_craters(const tgx_generation* spec, tgx_crater** craters, size_t* count,
                         char* err, size_t err_len);

/* On a sphere, the union of the lists of the spec's region and `region_count` more regions
 * (min_lon, min_lat, max_lon, max_lat each): each cell of a band is drawn once, and a crater is
 * kept when any region's generation keeps it; in band, cell and index order. No single
 * counterpart in the reference; its GenerateCraters (population.cpp:414-424) per region gives the
 * same craters. For catalogues along a path, such as a camera's track, that a box would
 * overcover. */
int tgx_generate_craters_regions(const tgx_generation* spec, const double* more_regions,
                                 size_t region_count, tgx_crater** craters, size_t* count,
                                 char* err, size_t err_len);

/* The crater models a layer shapes with (include/terraingen/layers/craters/crater_layer.h:29-46).
 * `model` 0: `profile` TG_PROFILE_MPD or TG_PROFILE_FASSETT_THOMSON, complex craters
 * (ComplexCraterProfile with the lunar shape) from `transition_diameter` when `complex_` is 1, and
 * FreshnessDegradation with the lunar fresh morphometry and ages uniform over `lifetime_gyr`.
 * `model` 1: DefaultCraterModels (crater_layer.cpp:194-203): the Martin-Parkes-Dunstan profile
 * diffused in time with matched diffusion over the lunar steady-state ages (the other fields are
 * ignored). */
typedef struct tgx_crater_models {
    uint32_t profile;
    uint32_t complex_;
    double transition_diameter;
    double lifetime_gyr;
    uint32_t model;
    uint32_t pad;
} tgx_crater_models;

/* The CraterApplyOptions switches of the hot path (apply_options.h:296-318): irregular outlines,
 * central peaks, terraces, and the optional rim smoothing, noise and texture amplitudes. */
typedef struct tgx_apply {
    uint32_t irregular;
    uint32_t peaks;
    uint32_t terraces;
    double rim_sigma;
    double noise_amplitude;
    double noise_rim_amplitude;
    double wall_streaks;
    double ejecta;
    /* CraterTerraces onset and full diameters (m); both 0 keep the defaults, 20 and 65 km. */
    double terrace_onset;
    double terrace_full;
} tgx_apply;

/* The crater layer's margin in radii: CraterReachRadii x MaxOutlineScale
 * (src/layers/craters/crater_layer.cpp:77-78). */
double tgx_margin_radii(const tgx_crater_models* models, const tgx_apply* apply);

/* PrepareCraters + MakeCraterKernelBatch over `craters` at sample `spacing` meters, the layer's
 * apply seeds derived from `layer_seed` (SeededCraterApply): the kernel struct, the records in
 * apply order (craters too small to resolve dropped), the data, and when `footprint` is not null
 * four doubles per record (centre x, centre y, footprint half-width, diameter). */
int tgx_prepare_craters(const tgx_crater* craters, size_t n, const tgx_crater_models* models,
                        const tgx_apply* apply, uint32_t layer_seed, double spacing,
                        tg_crater_shape* kernel, tg_crater_record** records, size_t* record_count,
                        float** data, size_t* data_count, double** footprint, char* err,
                        size_t err_len);

/* A boulder as the relief draws it: the fields of the reference's Boulder (include/terraingen/
 * layers/boulders/boulder.h:52-104) the dome relief reads. `exposed` is ExposedHeight(); each
 * facet is azimuth (rad from east), angle (rad from horizontal) and reach (m). */
typedef struct tgx_boulder {
    double x, y, diameter, exposed;
    double facets[6];
    uint32_t facet_count;
    uint32_t model; /* 1: placed as a 3D model, not drawn into the heights */
} tgx_boulder;

/* A boulder layer with generated boulders of power-law sizes, N(>=D) = sfd_coefficient
 * D^-sfd_exponent per m^2 for D in [min_d, max_d] (BoulderSizes with a PowerLawSfd), seed `seed`,
 * the pose ranges of BoulderPose (the reference's defaults: height 0.4-0.6 D, burial 0.1-0.4 of
 * the height, lean up to 0.26 rad), and no density or size-slope map, patches, clusters, bedrock
 * exposure, crater ejecta, fragmentation, rolling, settling or ventifacts; over the region
 * [min_x, max_x) x [min_y, max_y) (meters on a plane; longitude and latitude in degrees on a
 * sphere of `sphere_radius` meters when that is positive) for a grid of `spacing` meters (0: every
 * size). Candidates are drawn on `threads` threads (0: all), which does not change the list. */
typedef struct tgx_boulder_layer {
    double min_x, min_y, max_x, max_y;
    double sphere_radius;
    double spacing;
    double sfd_coefficient, sfd_exponent, min_d, max_d;
    double min_height_ratio, max_height_ratio;
    double min_burial, max_burial;
    double max_lean_degrees;
    uint32_t seed;
    uint32_t threads;
} tgx_boulder_layer;

/* A placed boulder (the Boulder fields the layer draws, boulder.h:52-106): centre, diameter,
 * height above its base, buried share of the height, yaw and lean (radians), wear and ID. */
typedef struct tgx_placed_boulder {
    double x, y, diameter, height, burial, yaw, lean, wear;
    uint64_t id;
    uint32_t model;
    uint32_t pad;
} tgx_placed_boulder;

/* BoulderLayer::Boulders over the layer's region for its spacing (src/layers/boulders/
 * boulder_layer.cpp:390-400): the candidates GenerateCraters draws with the layer's candidate
 * sizes (ScatterBoulderCandidates, :263-294), each thinned by its Keep draw (Thin, :338-352) and
 * posed by its Pose and Wear draws (DrawPose, boulder_pose.cpp:41-77), in candidate order,
 * bit-identical to the reference's list. */
int tgx_place_boulders(const tgx_boulder_layer* layer, tgx_placed_boulder** boulders,
                       size_t* count, char* err, size_t err_len);
/* tgx_place_boulders over the union of `box_count` regions (min_x, min_y, max_x, max_y each; the
 * layer's own region is not read), as a server places the boulders of many tiles at once: on a
 * sphere one generation over the union (each cube-sphere cell drawn once, the reference's order
 * for a union of regions); on a plane one per region, each boulder listed once. Every boulder
 * the layer lists over one of the regions is listed, with the same draws. */
int tgx_place_boulders_boxes(const tgx_boulder_layer* layer, const double* boxes,
                             size_t box_count, tgx_placed_boulder** boulders, size_t* count,
                             char* err, size_t err_len);

/* BoulderLayer::PrepareRelief for one grid `spacing` (boulder_relief.cpp:373-423): the boulders
 * drawn into heights (not models, resolution weight above 0) as device records, in list order,
 * with the parameter block (`dome` 1: the dome averaged over a band one sample wide, the default
 * look; 0: sampled at the point) and, when `footprint` is not null, four doubles per record
 * (centre x, centre y, relief box half-width, diameter) for tgx_bin_plane. */
int tgx_prepare_boulders(const tgx_boulder* boulders, size_t n, double spacing, int dome,
                         tg_boulder_kernel* kernel, tg_boulder** records, size_t* count,
                         double** footprint, char* err, size_t err_len);

/* A tile of a batch: a north-up grid `width` x `height` of `pixel` meters, west and north edges
 * in FP64. */
typedef struct tgx_tile {
    double origin_x;
    double origin_y;
    double pixel;
    uint32_t width;
    uint32_t height;
} tgx_tile;

/* Bins a catalogue (crater centres, footprint half-widths and FP64 diameters, in apply order) into
 * the 16 x 16 bins of every tile as ApplyCraters bins one grid (apply.cpp:119-140): craters the
 * tile's pixel size cannot resolve are left out, the rest join the bins their footprint's pixel
 * window meets, in catalogue order. Outputs the tile descriptors (heights packed one tile after
 * another), the bins and the entries for tg_hip_batch_stage, over `threads` threads (0: all). */
int tgx_batch_bin(const double* cx, const double* cy, const double* reach, const double* diameter,
                  size_t n, const tgx_tile* tiles, size_t tile_count, unsigned threads,
                  tg_tile_desc** descs, tg_bin** bins, size_t* bin_count, uint32_t** entries,
                  size_t* entry_count, char* err, size_t err_len);

/* A catalogue's spatial index for binning tiles, built once and kept (craters by diameter level
 * and, within a level, by square cells): what tgx_batch_bin builds and discards on every call. */
typedef struct tgx_index tgx_index;

int tgx_index_create(const double* cx, const double* cy, const double* reach,
                     const double* diameter, size_t n, tgx_index** out, char* err,
                     size_t err_len);
void tgx_index_free(tgx_index* index);
size_t tgx_index_count(const tgx_index* index);

/* tgx_batch_bin against a kept index. Thread-safe: the index is only read. */
int tgx_index_bin(const tgx_index* index, const tgx_tile* tiles, size_t tile_count,
                  unsigned threads, tg_tile_desc** descs, tg_bin** bins, size_t* bin_count,
                  uint32_t** entries, size_t* entry_count, char* err, size_t err_len);

/* One crater's cells in one tile (tg_window, tg_layouts.h), for binning on the device. */
typedef tg_window tgx_window;

/* The first half of tgx_index_bin, for binning on the device (tg_hip_batch_stage_windows): the
 * tile descriptors for bins of `bin_side` cells (8, 16 or 32), and per tile the windows of its
 * resolved candidates, in catalogue order; `tile_windows` holds (first, count) per tile. The bins
 * the windows give (every window's crater in each bin its cells meet) are those tgx_index_bin
 * gives for 16-cell bins. No counterpart in the reference, which bins one grid on the CPU. */
int tgx_index_windows(const tgx_index* index, const tgx_tile* tiles, size_t tile_count,
                      uint32_t bin_side, unsigned threads, tg_tile_desc** descs,
                      uint32_t** tile_windows, tgx_window** windows, size_t* window_count,
                      char* err, size_t err_len);

/* Counts the (cell, crater) pairs of a binned batch (the output of tgx_batch_bin) whose FP64 cell
 * centre lies strictly inside the crater's footprint half-width `reach`: the pairs a crater kernel
 * evaluates in full. `listed` receives every (cell, listed crater) pair, `inside` those within
 * reach and `ops` the sum of `cost[k]` over them. A measurement aid with no counterpart in the
 * reference; it never touches a device. */
int tgx_batch_pair_ops(const tg_tile_desc* descs, size_t tile_count, const tg_bin* bins,
                       size_t bin_count, const uint32_t* entries, size_t entry_count,
                       const double* cx, const double* cy, const double* reach,
                       const double* cost, size_t n, unsigned threads, uint64_t* listed,
                       uint64_t* inside, double* ops, char* err, size_t err_len);

/* ---- Tiles on a sphere ---------------------------------------------------------------------
 * Longitude-latitude rasters on a sphere of `radius` meters: `west` and `north` are a tile's edges
 * and `step` its cell size, in degrees; cells run east then south from the north-west one. */

/* Surface::Position of each cell centre (src/geometry/surface.cpp:45-50,261-267), width x height
 * x 3 doubles. The sine and cosine are the library's own, the same bits as the batched kernel's on
 * the device, within about an ulp of the C library's. */
int tgx_sphere_positions(double west, double north, double step, uint32_t width, uint32_t height,
                         double radius, double* xyz, char* err, size_t err_len);
/* MakeSplitPoint of each cell centre (crater_kernel.cpp:303-310) from those positions. */
int tgx_sphere_split_points(double west, double north, double step, uint32_t width,
                            uint32_t height, double radius, tg_split_point* out, char* err,
                            size_t err_len);
/* MakeCraterFrame (crater_kernel.cpp:285-301) of craters centred at (lon, lat) degrees, in FP64
 * with the C library's sine and cosine as the reference computes them. */
int tgx_sphere_frames(const double* lon, const double* lat, size_t n, double radius,
                      tg_crater_frame* out, char* err, size_t err_len);
/* tgx_index_bin for tiles on a sphere (tgx_tile's origin and pixel in degrees): crater k, centred
 * at (lon[k], lat[k]) with footprint half-width reach[k] meters of arc, joins every 16 x 16 bin of
 * a tile with a cell centre within reach of it, when it is resolved at the tile's spacing (pixel
 * times radius times pi / 180, RasterGrid::Spacing, src/raster/grid.cpp:28-33), in catalogue
 * order. The bounds are conservative: a bin may list a crater none of its cells reaches. */
int tgx_sphere_bin(const double* lon, const double* lat, const double* reach,
                   const double* diameter, size_t n, double radius, const tgx_tile* tiles,
                   size_t tile_count, unsigned threads, tg_tile_desc** descs, tg_bin** bins,
                   size_t* bin_count, uint32_t** entries, size_t* entry_count, char* err,
                   size_t err_len);
/* A kept index of a catalogue on a sphere (the columns of tgx_sphere_bin), for tgx_index_bin and
 * the tile server: tiles binned against it are longitude-latitude tiles, and a server given it
 * runs its batches on that sphere (tg_hip_batch_surface). Freed with tgx_index_free. */
int tgx_sphere_index_create(const double* lon, const double* lat, const double* reach,
                            const double* diameter, size_t n, double radius, tgx_index** out,
                            char* err, size_t err_len);
/* tgx_batch_pair_ops for tiles on a sphere (the output of tgx_sphere_bin): the (cell, listed
 * crater) pairs whose cell centre lies strictly within the crater's reach along the sphere. A
 * measurement aid with no counterpart in the reference; it never touches a device. */
int tgx_sphere_pair_ops(const tg_tile_desc* descs, size_t tile_count, const tg_bin* bins,
                        size_t bin_count, const uint32_t* entries, size_t entry_count,
                        const double* lon, const double* lat, const double* reach,
                        const double* cost, size_t n, double radius, unsigned threads,
                        uint64_t* listed, uint64_t* inside, double* ops, char* err,
                        size_t err_len);

/* ---- Reading a DEM: one band of a GeoTIFF ----------------------------------------------------
 * What GeoTiffSource reads of a file (src/sources/geotiff_source.cpp, src/raster/geotiff/
 * geotiff_read.cpp): the image's layout and georeferencing, and windows of one band decoded to
 * floats, for its posts on the device (tg_hip_upsample). TIFF and BigTIFF in either byte order,
 * strips or tiles, interleaved or planar samples of 8-64 bits (unsigned, signed, float), no
 * compression, LZW, PackBits or Deflate, horizontal and floating-point predictors. */
typedef struct tgx_geotiff_file tgx_geotiff_file;

typedef struct tgx_geotiff_info {
    uint32_t width, height, bands;
    uint32_t bits_per_sample;
    uint32_t sample_format;  /* 1 unsigned, 2 signed, 3 float */
    uint32_t compression;    /* 1 none, 5 LZW, 8 or 32946 Deflate, 32773 PackBits */
    uint32_t predictor;      /* 1 none, 2 horizontal, 3 floating point */
    uint32_t tiled, chunk_width, chunk_height;
    uint32_t big_tiff, big_endian;
    uint32_t georeferenced;  /* the model tags give a transform */
    uint32_t geographic;     /* GTModelTypeGeoKey: longitude and latitude */
    uint32_t pixel_is_point; /* GTRasterTypeGeoKey: PixelIsPoint (the transform is moved back) */
    uint32_t has_no_data;
    double no_data;          /* GDAL's no-data tag, when has_no_data */
    double transform[6];     /* GDAL order: origin x, pixel width, row rotation, origin y, column
                              * rotation, pixel height */
} tgx_geotiff_info;

int tgx_geotiff_open(const char* path, tgx_geotiff_file** file, char* err, size_t err_len);
/* The file's info and band `band`'s scale and offset (GDAL metadata; 1 and 0 without). */
int tgx_geotiff_info_get(const tgx_geotiff_file* file, uint32_t band, tgx_geotiff_info* info,
                         double* scale, double* offset);
/* Band `band`'s cells [col0, col0 + width) x [row0, row0 + height) into `out` row by row: NaN
 * where a sample is NaN or the no-data value, else float(double(sample) * scale + offset) (the
 * sample as is when scale is 1 and offset 0), as GeoTiffSource's posts decode them
 * (geotiff_source.cpp:115-128). Decodes only the chunks the window meets. */
int tgx_geotiff_read(const tgx_geotiff_file* file, uint32_t band, uint64_t col0, uint64_t row0,
                     uint32_t width, uint32_t height, double scale, double offset, float* out,
                     char* err, size_t err_len);
void tgx_geotiff_close(tgx_geotiff_file* file);

/* ---- The tile server: continuous batching onto the batched kernel ---------------------------
 * No counterpart in the reference, which evaluates one grid per call. Requests (a tile and the host
 * memory its heights go to) queue as they arrive; a dispatcher thread admits them, in arrival order,
 * into whichever of `slots` batches is free, as many as fit `max_batch_tiles`, `max_batch_cells`
 * and the HBM output pool of `pool_bytes` (the admission budget), bins them against the kept index,
 * launches one persistent kernel per batch on the slot's stream and copies the heights back
 * through pinned memory. Batches in different slots overlap on the device. The device handles
 * (tg_dev, tg_noise_plan, tg_catalogue) come from libtg_hip.so, which this library opens on first
 * use. */
typedef struct tgx_server tgx_server;

typedef struct tgx_server_config {
    uint32_t slots;             /* batches in flight, 1-16 */
    uint32_t max_batch_tiles;   /* tiles per batch */
    uint64_t max_batch_cells;   /* cells per batch; also the pinned staging per slot (x4 bytes) */
    uint64_t pool_bytes;        /* HBM for outputs of admitted batches */
    uint32_t bin_threads;       /* host threads binning one batch (0: all) */
    uint32_t poll_microseconds; /* how often the dispatcher looks at running batches (0: 20) */
} tgx_server_config;

typedef struct tgx_ticket_info {
    double submitted; /* seconds on the steady clock */
    double admitted;  /* into a batch */
    double launched;  /* the batch's kernel queued */
    double done;      /* heights in the request's memory */
    uint32_t batch;   /* the batch it ran in, from 1 */
    uint32_t batch_tiles;
} tgx_ticket_info;

typedef struct tgx_server_stats {
    uint64_t batches;
    uint64_t tiles;
    uint64_t cells;
    uint64_t failed;
    uint32_t max_in_flight; /* most batches on the device at once */
    uint32_t pad;
    uint64_t pool_bytes;
    uint64_t pool_peak_bytes;
    uint64_t pending;
} tgx_server_stats;

/* `index` may be null (no craters); otherwise `catalogue` holds the same craters in the same
 * order. The plan, catalogue and index must outlive the server. */
int tgx_server_create(tg_dev* dev, const tg_noise_plan* plan, float elevation,
                      const tg_catalogue* catalogue, const tgx_index* index,
                      const tgx_server_config* config, tgx_server** out, char* err,
                      size_t err_len);
/* A boulder layer drawn after the craters in every served tile (no counterpart in the reference):
 * the device's boulder catalogue (tg_hip_boulder_catalogue_create) and an index of the boulders'
 * relief boxes in the same order (tgx_index_create, or tgx_sphere_index_create on the craters'
 * sphere), each batch's boulder lists binned against it and staged with the craters'
 * (tg_hip_batch_stage_boulders). Set once, before the first submit; both must outlive the
 * server. */
int tgx_server_boulders(tgx_server* server, const tg_boulder_catalogue* catalogue,
                        const tgx_index* index, char* err, size_t err_len);
/* Thread-safe. `dst` receives width x height floats, row-major from the north-west cell, and must
 * stay valid until the ticket is waited for. */
int tgx_server_submit(tgx_server* server, const tgx_tile* tile, float* dst, uint64_t* ticket,
                      char* err, size_t err_len);
/* Blocks until the ticket's heights are in place (timeout < 0: no limit; TG_NOT_READY when the
 * timeout passes first), then releases the ticket. Thread-safe. */
int tgx_server_wait(tgx_server* server, uint64_t ticket, double timeout_seconds,
                    tgx_ticket_info* info, char* err, size_t err_len);
int tgx_server_stats_get(tgx_server* server, tgx_server_stats* out);
/* Finishes every submitted request, then frees the server. No thread may be waiting on it. */
int tgx_server_free(tgx_server* server);

/* ---- The backend registry and RunKernel, with the hip backend ----------------------------------
 * The reference's include/terraingen/gpu/backend.h and src/gpu/kernel.h for hosts that are not
 * C++: backends are numbered as the reference's enum with Hip appended (0 auto, 1 cpu, 2 vulkan,
 * 3 gpu, 4 cuda, 5 metal, 6 hip). The C++ forms are in the host library's backend.h. */

/* BackendNamed (backend.cpp:103-111): ignores case; TG_ERR_ARGUMENT for an unknown name. */
int tgx_backend_named(const char* name, int* backend);
/* BackendName: writes the name ("unknown" out of range) and returns its length. */
size_t tgx_backend_name(int backend, char* out, size_t len);
/* What `terraingen backends` prints (app/cli.cpp:192-205): one line per Backends() entry, CPU then
 * Vulkan, CUDA, Metal and hip, with the device or the reason, then the default. Returns the
 * length of the full text. */
size_t tgx_backends_report(char* out, size_t len);
/* DefaultBackend / SetDefaultBackend (backend.cpp:136-153): TG_ERR_ARGUMENT for an unknown
 * $TERRAINGEN_BACKEND. */
int tgx_default_backend(int* backend, char* err, size_t err_len);
void tgx_set_default_backend(int backend);
/* ResolveBackend (backend.cpp:155-183): TG_ERR_NO_DEVICE with the reason when a requested GPU
 * backend cannot run. */
int tgx_resolve_backend(int requested, int* backend, char* err, size_t err_len);
/* RunKernel (src/gpu/kernel.cpp:130-153): the program is a TG_KERNEL_* id; `values` holds the
 * grid's heights, `no_data` (nullable) its no-data value; extras bind after the kernel struct,
 * `per_cell` ones sliced per strip. `*ran` is 0, leaving `values` untouched, when the backend is
 * not a GPU one, its device is unavailable, the program has no hip form, or the grid is empty or
 * too wide; 1 when the kernel ran. `max_strip_bytes` 0 means 64 MiB, as the reference. */
int tgx_run_kernel(int backend, int program, const void* kernel, size_t kernel_bytes,
                   const tgx_grid* grid, float* values, const float* no_data, double origin_x,
                   double origin_y, const tg_buf* extra, size_t n_extra, size_t max_strip_bytes,
                   int* ran, char* err, size_t err_len);

/* Empty when libtg_hip.so opens, else the loader's reason. */
const char* tgx_hip_unavailable_reason(void);

/* ---- Output files --------------------------------------------------------------------------
 * A heightmap (or several bands of one grid) as the reference's WriteGeoTiff writes it
 * (src/raster/geotiff/geotiff_write.cpp:221-275, GeoTiffWriteOptions in
 * include/terraingen/raster/geotiff/geotiff.h:185-192): tiled little-endian float32, edge tiles
 * padded with the no-data value or 0, compressed tiles through the floating-point predictor,
 * BigTIFF when the uncompressed tiles could pass 4 GiB, the grid's transform (GDAL order:
 * origin x, x per column, x per row, origin y, y per column, y per row) and projection keys
 * (GeoKeyDirectory shorts, GeoDoubleParams, GeoAsciiParams, as ReadGeoTiff returns them) and the
 * no-data value as GDAL_NODATA. With no compression or LZW the file is the reference's byte for
 * byte; Deflate (zlib, level 6) decodes to the same values. `values` holds `bands` planes of
 * width x height floats, row-major from the north-west cell. Tiles are encoded on `threads`
 * threads (0: all), which does not change the file. */
enum { TGX_TIFF_NONE = 0, TGX_TIFF_LZW = 1, TGX_TIFF_DEFLATE = 2 };
typedef struct tgx_geotiff {
    const char* path;
    uint32_t width;
    uint32_t height;
    uint32_t bands;
    uint32_t compression; /* TGX_TIFF_*; the reference's default is LZW */
    uint32_t tile_size;   /* a positive multiple of 16; the reference's default is 256 */
    uint32_t big_tiff;    /* 1: always BigTIFF */
    uint32_t has_nodata;
    float nodata;
    double transform[6];
    const uint16_t* geokeys;
    size_t geokey_count;
    const double* geo_doubles;
    size_t geo_double_count;
    const char* geo_ascii; /* null or empty for none */
    uint32_t threads;
    uint32_t pad;
} tgx_geotiff;
int tgx_write_geotiff(const tgx_geotiff* spec, const float* values, char* err, size_t err_len);

/* CraterAttributes (include/terraingen/layers/craters/shaping/attributes.h:14-21). */
typedef struct tgx_crater_attributes {
    double floor_depth;
    double rim_height;
    double depth_ratio;
    double freshness;
    uint8_t complex_;
    uint8_t pad[7];
} tgx_crater_attributes;

/* A crater list as the reference's WriteCraterList writes CSV (src/io/feature_lists.cpp:124-140,
 * 431-445): columns x, y, diameter, age_gyr, id, manual, replace (each double as "%.17g", flags as
 * 0 or 1), then floor_depth, rim_height, depth_ratio, freshness and complex when `attributes`
 * (one per crater) is given. `manual` and `replace` may be null (all 0). */
int tgx_write_crater_csv(const char* path, const tgx_crater* craters, size_t count,
                         const uint8_t* manual, const uint8_t* replace,
                         const tgx_crater_attributes* attributes, char* err, size_t err_len);

#ifdef __cplusplus
}
#endif

#endif /* TERRAINGEN_AMD_H */
