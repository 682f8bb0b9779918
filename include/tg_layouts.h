/* SPDX-License-Identifier: BSD-3-Clause */
/*
 * Binary layouts shared by the host and the gfx950 kernels of the hip backend.
 *
 * Each struct is the std430 block the reference's kernel headers define, member for member, so a
 * host that builds the reference's kernel structs can hand their bytes straight to this library:
 *   tg_fractal            include/terraingen/noise/fractals.h:29-67       (52 B)
 *   tg_noise_layer_kernel include/terraingen/layers/kernels/noise_layer.h:19-36 (420 B)
 *   tg_crater_profile     include/terraingen/layers/kernels/craters.h:29-49  (76 B)
 *   tg_crater_shape       include/terraingen/layers/kernels/craters.h:52-82  (188 B)
 *   tg_crater_record      include/terraingen/layers/kernels/craters.h:85-122 (144 B)
 *   tg_crater_frame       include/terraingen/layers/kernels/craters.h:132-148 (52 B)
 *   tg_split_point        include/terraingen/layers/kernels/craters.h:188-195 (24 B)
 *   tg_kernel_params      src/gpu/kernel.cpp:34-40, src/gpu/kernel.slang:9-19 (20 B)
 * A C++ `bool` member of the reference occupies one 32-bit word here; its padding bytes must be
 * zero, as the reference's GPU conformance tests require (tests/unit/gpu/noise_test.cpp:65-90).
 */
#ifndef TG_LAYOUTS_H
#define TG_LAYOUTS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { TG_MAX_FRACTALS = 8, TG_MAX_OCTAVES = 32 };

/* NoiseBasis, FractalType and FractalBlend (include/terraingen/noise/fractal.h:20-40,
 * include/terraingen/noise/fractals.h:12-18). */
enum {
    TG_BASIS_PERLIN = 0,
    TG_BASIS_SIMPLEX = 1,
    TG_BASIS_VALUE = 2,
    TG_BASIS_WORLEY_F1 = 3,
    TG_BASIS_WORLEY_F2 = 4,
    TG_BASIS_WORLEY_F2_MINUS_F1 = 5,
    TG_BASIS_VORONOI = 6
};
enum { TG_FRACTAL_FBM = 0, TG_FRACTAL_RIDGED = 1, TG_FRACTAL_BILLOW = 2, TG_FRACTAL_HYBRID = 3 };
enum { TG_BLEND_ADD = 0, TG_BLEND_MAX = 1, TG_BLEND_MIN = 2, TG_BLEND_MASK = 3 };

typedef struct tg_fractal {
    uint32_t enabled; /* the reference's bool, zero-padded to a word */
    uint32_t seed;
    float wavelength;
    float amplitude;
    int32_t octaves;
    float lacunarity;
    float gain;
    uint32_t basis;
    uint32_t type;
    float offset;
    uint32_t blend;
    float warp_amplitude;
    float warp_wavelength;
} tg_fractal;

typedef struct tg_noise_layer_kernel {
    tg_fractal fractals[TG_MAX_FRACTALS];
    uint32_t count;
} tg_noise_layer_kernel;

/* CraterProfileKind (craters.h:16-23). */
enum { TG_PROFILE_FASSETT_THOMSON = 0, TG_PROFILE_MPD = 1, TG_PROFILE_TABULATED = 2 };

typedef struct tg_crater_profile {
    uint32_t simple;
    uint32_t rim;
    uint32_t complex_;
    uint32_t diffused;
    float transition_diameter;
    float floor_coefficient;
    float floor_exponent;
    float peak_height_coefficient;
    float peak_height_exponent;
    float peak_width_coefficient;
    float peak_width_exponent;
    uint32_t peak_from_volume;
    float peak_fit_max_diameter;
    float ejecta_exponent;
    float complex_reach_radii;
    float table_extent;
    uint32_t table_cells;
    uint32_t ages_offset;
    uint32_t age_count;
} tg_crater_profile;

typedef struct tg_crater_shape {
    tg_crater_profile profile;
    float reach_radii;
    float profile_reach_radii;
    float blend_start;
    float blend_end;
    uint32_t blend_shape;
    float interior_replacement;
    float replace_start;
    float replace_end;
    float mantling;
    float ejecta_reach_radii;
    float rim_sigma;
    float rim_inner;
    float rim_outer;
    float rim_strength;
    uint32_t rim_shape;
    float noise_amplitude;
    float noise_wavelength;
    int32_t noise_octaves;
    float noise_gain;
    float noise_rim_amplitude;
    float noise_rim_wavelength;
    float noise_rim_inner;
    float noise_rim_outer;
    float wall_streaks;
    float wall_streak_width;
    float ejecta;
    float ejecta_lineation;
    float min_freshness;
} tg_crater_shape;

typedef struct tg_crater_record {
    float x;
    float y;
    float diameter;
    float resolution;
    float fresh_depth;
    float fresh_rim;
    float current_depth;
    float current_rim;
    float diffusion;
    float ellipticity;
    float heading;
    uint32_t first_harmonic;
    uint32_t harmonic_count;
    uint32_t harmonic_offset;
    uint32_t table_count;
    uint32_t table_offset;
    uint32_t profile_offset;
    uint32_t has_datum;
    float datum_a;
    float datum_b;
    float datum_c;
    uint32_t noise_key;
    uint32_t texture_key;
    uint32_t mark_count;
    uint32_t mark_offset;
    float shift_x;
    float shift_y;
    float upslope_x;
    float upslope_y;
    uint32_t terrace_count;
    uint32_t terrace_offset;
    uint32_t peak_count;
    uint32_t peak_offset;
    uint32_t replace;
    float x_low;
    float y_low;
} tg_crater_record;

typedef struct tg_crater_frame {
    float x, y, z;
    float east_x, east_y, east_z;
    float north_x, north_y, north_z;
    float sphere_radius;
    float x_low, y_low, z_low;
} tg_crater_frame;

typedef struct tg_split_point {
    float x, x_low, y, y_low, z, z_low;
} tg_split_point;

/* One crater's cells in one tile, for binning on the device: the inclusive column and row ranges
 * of the tile's cells its footprint box covers (its pixel window, src/layers/craters/shaping/
 * apply.cpp:119-140; on a sphere the cells within reach with one to spare each side). */
typedef struct tg_window {
    uint32_t crater;
    uint32_t col0, col1;
    uint32_t row0, row1;
} tg_window;

/* CraterDatumKernel (include/terraingen/layers/kernels/craters.h:1214-1221): what fitting crater
 * datums on a plane takes: the craters' shapes, the fit footprint in crater radii, lattice points
 * per side and the size bands of the crater index. */
typedef struct tg_crater_datum_kernel {
    tg_crater_shape shape;
    float fit_radii;
    uint32_t lattice_side;
    uint32_t band_count;
} tg_crater_datum_kernel;

/* CraterBodyDatumKernel (craters.h:1308-1311): the craters' shapes, for datums on a body. */
typedef struct tg_crater_body_datum_kernel {
    tg_crater_shape shape;
} tg_crater_body_datum_kernel;

/* One boulder as the dome relief reads it. The reference bakes boulders on the CPU in FP64
 * (BoulderLayer::AddRelief and DomeHeight, src/layers/boulders/boulder_relief.cpp:44-57,244-275)
 * and has no GPU layout for them; this is the device's, prepared on the host from a boulder list
 * for one grid spacing. */
typedef struct tg_boulder {
    float x, y;          /* centre, FP32 high parts */
    float x_low, y_low;  /* and the remainders (the split of MakeSplitPoint) */
    float radius;        /* 0.5 D */
    float exposed;       /* Boulder::ExposedHeight: height (1 - burial) */
    float reach;         /* 0.5 (D + band width): half the side of its relief box */
    float weight;        /* ResolutionWeight(D, spacing) of the grid */
    float facet[8];      /* per ventifact: cos azimuth, sin azimuth, reach, tan angle */
    uint32_t facet_count;
    uint32_t pad[3];
} tg_boulder;

/* The boulder relief's parameter block: the band the dome is averaged over (the sample spacing
 * for BoulderRelief::Dome, 0 for a point sample) and the number of records. */
typedef struct tg_boulder_kernel {
    float width;
    uint32_t count;
    uint32_t pad[2];
} tg_boulder_kernel;

/* ResampleMethod (include/terraingen/raster/resample.h:12-23), the kernel upsamplers' kernels. */
enum {
    TG_RESAMPLE_NEAREST = 0,
    TG_RESAMPLE_BILINEAR = 1,
    TG_RESAMPLE_CUBIC = 2,
    TG_RESAMPLE_LANCZOS3 = 3,
    TG_RESAMPLE_CUBIC_BSPLINE = 4
};

/* One KernelUpsampler fill of a north-up grid from a lattice of posts (the reference's
 * Upsampler::Fill over a PostLattice, src/sources/upsampling.cpp:297-312; no GPU layout there,
 * which upsamples on the CPU). The lattice is north-up from (lattice_x, lattice_y) with cells
 * `post` wide and `post_height` high: post (col, row) sits at the centre of that cell. The posts held are a window of
 * window_width x window_height from lattice (col0, row0), row by row; reads beyond it repeat its
 * nearest edge post (RasterPosts, detail::PostIndex), after wrapping columns into
 * [0, lattice_columns) when `wrap_columns`. The grid is width x height cells of `pixel` from its
 * north-west corner (origin_x, origin_y), written `out_pitch` floats apart row to row (0: width).
 * With `at_post_spacing` (the sample spacing is no finer than the posts, AtPostSpacing) a cell
 * centred on a post takes the post. With `covered`, a cell whose centre lies outside the lattice
 * cells [cover_col0, cover_col1) x [cover_row0, cover_row1) is NaN (RasterHeights::Covers,
 * src/sources/raster_source.cpp:78-82). */
typedef struct tg_upsample {
    double lattice_x, lattice_y;
    double post;         /* the lattice's cell width */
    double post_height;  /* and height (0: the width) */
    int64_t col0, row0;
    uint32_t window_width, window_height;
    uint32_t wrap_columns;
    uint32_t lattice_columns;
    double origin_x, origin_y;
    double pixel;
    uint32_t width, height;
    uint32_t method;
    uint32_t at_post_spacing;
    uint32_t out_pitch;
    uint32_t covered;
    int64_t cover_col0, cover_row0, cover_col1, cover_row1;
} tg_upsample;

/* No-data modes of tg_kernel_params.no_data (src/gpu/kernel.cpp:22-27). */
enum { TG_NODATA_NONE = 0, TG_NODATA_NAN = 1, TG_NODATA_VALUE = 2 };

typedef struct tg_kernel_params {
    uint32_t width;
    uint32_t rows;
    uint32_t no_data;
    float no_data_value;
    uint32_t first_row;
} tg_kernel_params;

#ifdef __cplusplus
}
static_assert(sizeof(tg_fractal) == 52, "tg_fractal must match Fractal's std430 layout");
static_assert(sizeof(tg_noise_layer_kernel) == 420, "tg_noise_layer_kernel must be 420 bytes");
static_assert(sizeof(tg_crater_profile) == 76, "tg_crater_profile must be 76 bytes");
static_assert(sizeof(tg_crater_shape) == 188, "tg_crater_shape must be 188 bytes");
static_assert(sizeof(tg_crater_record) == 144, "tg_crater_record must be 144 bytes");
static_assert(sizeof(tg_crater_frame) == 52, "tg_crater_frame must be 52 bytes");
static_assert(sizeof(tg_split_point) == 24, "tg_split_point must be 24 bytes");
static_assert(sizeof(tg_window) == 20, "tg_window is 20 bytes");
static_assert(sizeof(tg_crater_datum_kernel) == 200, "tg_crater_datum_kernel must be 200 bytes");
static_assert(sizeof(tg_crater_body_datum_kernel) == 188,
              "tg_crater_body_datum_kernel must be 188 bytes");
static_assert(sizeof(tg_kernel_params) == 20, "tg_kernel_params must be 20 bytes");
static_assert(sizeof(tg_boulder) == 80, "tg_boulder is 80 bytes");
static_assert(sizeof(tg_boulder_kernel) == 16, "tg_boulder_kernel is 16 bytes");
static_assert(sizeof(tg_upsample) == 144, "tg_upsample is 144 bytes");
#endif

#endif /* TG_LAYOUTS_H */
