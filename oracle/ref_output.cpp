// SPDX-License-Identifier: BSD-3-Clause
//
// ORACLE — test infrastructure only (see noise_oracle.cpp). A driver of the reference's output
// writers, built into oracle/_ref/libtgref.so from its sources: WriteGeoTiff
// (src/raster/geotiff/geotiff_write.cpp:278-290) of a raster with its transform, no-data value and
// projection keys, and WriteCraterList (src/io/feature_lists.cpp:431-445) as CSV. The tests write
// the same data through this repository's writers and compare the files. Nothing of the reference
// is copied here: this file calls its functions.

#include "terraingen/io/feature_lists.h"
#include "terraingen/raster/geotiff/geotiff.h"
#include "terraingen/raster/raster.h"

#include <cstdint>
#include <exception>
#include <string>
#include <vector>

using namespace terraingen;

namespace {

thread_local std::string gOutputError;

}  // namespace

extern "C" {

const char* ref_output_error() {
    return gOutputError.c_str();
}

// `bands` planes of width x height floats; `transform` in GDAL order; compression 0 none, 1 LZW,
// 2 Deflate.
int ref_write_geotiff(const char* path, const float* values, std::uint32_t width,
                      std::uint32_t height, std::uint32_t bands, const double* transform,
                      int hasNoData, float noData, const std::uint16_t* geokeys,
                      std::size_t geokeyCount, const double* doubles, std::size_t doubleCount,
                      const char* ascii, int compression, std::uint32_t tileSize, int bigTiff) {
    try {
        const GeoTransform t = GeoTransform::FromGdal(
            {transform[0], transform[1], transform[2], transform[3], transform[4], transform[5]});
        const RasterGrid grid{width, height, t};
        std::vector<Raster<float>> rasters;
        for (std::uint32_t b = 0; b < bands; ++b) {
            Raster<float> r(grid, 0.0f);
            for (std::size_t i = 0; i < std::size_t(width) * height; ++i) {
                r.Values()[i] = values[b * std::size_t(width) * height + i];
            }
            if (hasNoData) {
                r.SetNoData(noData);
            }
            rasters.push_back(std::move(r));
        }
        GeoKeys keys;
        keys.directory.assign(geokeys, geokeys + geokeyCount);
        keys.doubleParams.assign(doubles, doubles + doubleCount);
        keys.asciiParams = ascii != nullptr ? ascii : "";
        GeoTiffWriteOptions options;
        options.compression = compression == 0   ? TiffCompression::None
                              : compression == 1 ? TiffCompression::Lzw
                                                 : TiffCompression::Deflate;
        options.tileSize = tileSize;
        options.bigTiff = bigTiff != 0;
        if (bands == 1) {
            WriteGeoTiff(path, rasters.front(), keys, options);
        } else {
            WriteGeoTiff(path, rasters, keys, options);
        }
        return 0;
    } catch (const std::exception& e) {
        gOutputError = e.what();
        return 1;
    }
}

// The craters (x, y, diameter, age, id; manual and replace flags may be null) as CSV, with
// attributes (floor depth, rim height, depth ratio, freshness per crater, and complex flags) when
// `attributes` is not null.
int ref_write_crater_csv(const char* path, const double* x, const double* y, const double* d,
                         const double* age, const std::uint64_t* id, const std::uint8_t* manual,
                         const std::uint8_t* replace, std::size_t count,
                         const double* attributes, const std::uint8_t* complexFlags) {
    try {
        CraterList list;
        for (std::size_t i = 0; i < count; ++i) {
            Crater c;
            c.x = x[i];
            c.y = y[i];
            c.diameter = d[i];
            c.ageGyr = age[i];
            c.id = id[i];
            c.manual = manual != nullptr && manual[i] != 0;
            c.replace = replace != nullptr && replace[i] != 0;
            list.Append(c);
        }
        std::vector<CraterAttributes> attrs;
        if (attributes != nullptr) {
            for (std::size_t i = 0; i < count; ++i) {
                CraterAttributes a;
                a.floorDepth = attributes[4 * i];
                a.rimHeight = attributes[4 * i + 1];
                a.depthRatio = attributes[4 * i + 2];
                a.freshness = attributes[4 * i + 3];
                a.complex = complexFlags[i] != 0;
                attrs.push_back(a);
            }
        }
        WriteCraterList(list, path, ListFormat::Csv, attrs);
        return 0;
    } catch (const std::exception& e) {
        gOutputError = e.what();
        return 1;
    }
}

}  // extern "C"
