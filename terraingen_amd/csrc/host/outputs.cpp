// SPDX-License-Identifier: BSD-3-Clause
//
// Output files of a generated terrain (tgx_write_geotiff, tgx_write_crater_csv): the heightmap as
// the reference's WriteGeoTiff writes it (src/raster/geotiff/geotiff_write.cpp) and the crater
// list as its WriteCraterList writes CSV (src/io/feature_lists.cpp:124-160,431-445).
//
// The GeoTIFF follows the TIFF 6.0 and GeoTIFF 1.1 formats with the reference's choices, which
// fix every byte of the file: little-endian, tiled float32 (square tiles, edge tiles padded with
// the no-data value or 0), bands as separate planes, compressed tiles through the floating-point
// predictor (TIFF Technical Note 3: each row's bytes split into planes from the most significant
// byte, then differenced), tiles at even offsets from the end of the header, then the IFD with its
// entries in tag order and its out-of-line values after it at even offsets, and BigTIFF when the
// uncompressed tiles could pass 4 GiB. LZW is TIFF's variant: MSB-first codes of 9 to 12 bits,
// a Clear code first, greedy matching, widths raised one code early, a Clear once 4094 codes are
// in use. Tiles are encoded on threads and written in order, so the file does not depend on the
// thread count. Deflate is zlib's (level 6, 32 KiB window): the file decodes to the same values
// as the reference's, whose compressed bytes come from another deflate encoder.

#include "common.h"
#include "terraingen_amd.h"

#include <zlib.h>

#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace tgh {
namespace {

// TIFF tag numbers and field types (TIFF 6.0, BigTIFF, GeoTIFF 1.1, GDAL's no-data tag).
enum : std::uint16_t {
    kTagWidth = 256,
    kTagLength = 257,
    kTagBitsPerSample = 258,
    kTagCompression = 259,
    kTagPhotometric = 262,
    kTagSamplesPerPixel = 277,
    kTagPlanar = 284,
    kTagPredictor = 317,
    kTagTileWidth = 322,
    kTagTileLength = 323,
    kTagTileOffsets = 324,
    kTagTileByteCounts = 325,
    kTagExtraSamples = 338,
    kTagSampleFormat = 339,
    kTagPixelScale = 33550,
    kTagTiepoint = 33922,
    kTagTransformation = 34264,
    kTagGeoKeys = 34735,
    kTagGeoDoubles = 34736,
    kTagGeoAscii = 34737,
    kTagNoData = 42113,
};
enum : std::uint16_t { kTypeAscii = 2, kTypeShort = 3, kTypeLong = 4, kTypeDouble = 12, kTypeLong8 = 16 };

// An IFD under construction: each tag's type, count and value bytes.
class Ifd {
  public:
    explicit Ifd(bool big) : m_word(big ? 8 : 4), m_countBytes(big ? 8 : 2) {}

    template <class T> void Put(std::uint16_t tag, std::uint16_t type, const std::vector<T>& v) {
        Field& f = m_fields[tag];
        f.type = type;
        f.count = v.size();
        f.bytes.resize(v.size() * sizeof(T));
        if (!v.empty()) {
            std::memcpy(f.bytes.data(), v.data(), f.bytes.size());
        }
    }

    void PutText(std::uint16_t tag, const std::string& text) {
        Field& f = m_fields[tag];
        f.type = kTypeAscii;
        f.bytes.assign(text.begin(), text.end());
        f.bytes.push_back(0);
        f.count = f.bytes.size();
    }

    // The entry table (count, entries in tag order, a zero next-IFD offset) followed by the
    // values that do not fit an entry, as they lie from file offset `at`.
    std::vector<std::uint8_t> Bytes(std::uint64_t at) const {
        const std::size_t entry = 4 + 2 * m_word;
        const std::size_t table = m_countBytes + m_fields.size() * entry + m_word;
        std::vector<std::uint8_t> out(table, 0);
        std::vector<std::uint8_t> values;
        const auto store = [&](std::size_t offset, std::uint64_t value, std::size_t bytes) {
            std::memcpy(out.data() + offset, &value, bytes);  // little-endian host
        };
        store(0, m_fields.size(), m_countBytes);
        std::size_t e = m_countBytes;
        for (const auto& [tag, f] : m_fields) {
            store(e, tag, 2);
            store(e + 2, f.type, 2);
            store(e + 4, f.count, m_word);
            if (f.bytes.size() <= m_word) {
                std::memcpy(out.data() + e + 4 + m_word, f.bytes.data(), f.bytes.size());
            } else {
                values.resize(values.size() + (values.size() & 1u), 0);
                store(e + 4 + m_word, at + table + values.size(), m_word);
                values.insert(values.end(), f.bytes.begin(), f.bytes.end());
            }
            e += entry;
        }
        out.insert(out.end(), values.begin(), values.end());
        return out;
    }

  private:
    struct Field {
        std::uint16_t type = 0;
        std::uint64_t count = 0;
        std::vector<std::uint8_t> bytes;
    };
    std::size_t m_word, m_countBytes;
    std::map<std::uint16_t, Field> m_fields;
};

// MSB-first codes of 9 to 12 bits.
class CodeStream {
  public:
    void Put(std::uint32_t code, int bits) {
        m_acc = (m_acc << bits) | code;
        m_held += bits;
        while (m_held >= 8) {
            m_held -= 8;
            m_out.push_back(static_cast<std::uint8_t>(m_acc >> m_held));
        }
        m_acc &= (std::uint64_t{1} << m_held) - 1u;
    }
    std::vector<std::uint8_t> Close() {
        if (m_held > 0) {
            m_out.push_back(static_cast<std::uint8_t>(m_acc << (8 - m_held)));
            m_held = 0;
        }
        return std::move(m_out);
    }

  private:
    std::vector<std::uint8_t> m_out;
    std::uint64_t m_acc = 0;
    int m_held = 0;
};

// TIFF LZW over `in`: strings are (prefix code, byte) pairs in a table whose slots carry the
// pass that filled them, so a Clear forgets them at once.
std::vector<std::uint8_t> Lzw(const std::vector<std::uint8_t>& in) {
    constexpr std::uint32_t kClearCode = 256, kEnd = 257, kFirst = 258, kLimit = 4094;
    constexpr std::uint32_t kSlots = 1u << 13;
    struct Slot {
        std::uint32_t key = 0, pass = 0;
        std::uint16_t code = 0;
    };
    std::vector<Slot> table(kSlots);
    std::uint32_t pass = 1;
    const auto slotOf = [](std::uint32_t key) {
        std::uint32_t h = key ^ (key >> 7);
        h *= 0x9E3779B1u;
        return (h >> 19) & (kSlots - 1u);
    };
    CodeStream out;
    int bits = 9;
    std::uint32_t next = kFirst;
    out.Put(kClearCode, bits);
    const auto emit = [&](std::uint32_t code) {
        out.Put(code, bits);
        ++next;
        if (next >= (1u << bits) && bits < 12) {
            ++bits;  // one code early, as TIFF readers expect
        }
    };
    if (!in.empty()) {
        std::uint32_t w = in[0];
        for (std::size_t i = 1; i < in.size(); ++i) {
            const std::uint32_t key = (w << 8) | in[i];
            std::uint32_t s = slotOf(key);
            while (table[s].pass == pass && table[s].key != key) {
                s = (s + 1u) & (kSlots - 1u);
            }
            if (table[s].pass == pass) {
                w = table[s].code;
                continue;
            }
            table[s] = {key, pass, static_cast<std::uint16_t>(next)};
            emit(w);
            if (next >= kLimit) {
                out.Put(kClearCode, bits);
                bits = 9;
                next = kFirst;
                ++pass;
            }
            w = in[i];
        }
        emit(w);
    }
    out.Put(kEnd, bits);
    return out.Close();
}

// The floating-point predictor of one tile of `side` x `side` float32 samples, in place.
void FloatPredictor(std::vector<std::uint8_t>& tile, std::size_t side) {
    const std::size_t rowBytes = side * 4;
    std::vector<std::uint8_t> row(rowBytes);
    for (std::size_t r = 0; r < side; ++r) {
        std::uint8_t* p = tile.data() + r * rowBytes;
        for (std::size_t s = 0; s < side; ++s) {
            for (std::size_t b = 0; b < 4; ++b) {
                row[(3 - b) * side + s] = p[s * 4 + b];  // plane 0: the most significant bytes
            }
        }
        p[0] = row[0];
        for (std::size_t i = 1; i < rowBytes; ++i) {
            p[i] = static_cast<std::uint8_t>(row[i] - row[i - 1]);
        }
    }
}

std::vector<std::uint8_t> Deflate(const std::vector<std::uint8_t>& in) {
    z_stream z{};
    if (deflateInit2(&z, 6, Z_DEFLATED, 15, 8, Z_DEFAULT_STRATEGY) != Z_OK) {
        throw std::runtime_error("zlib deflateInit2 failed");
    }
    std::vector<std::uint8_t> out(deflateBound(&z, static_cast<uLong>(in.size())));
    z.next_in = const_cast<Bytef*>(in.data());
    z.avail_in = static_cast<uInt>(in.size());
    z.next_out = out.data();
    z.avail_out = static_cast<uInt>(out.size());
    const int done = deflate(&z, Z_FINISH);
    out.resize(z.total_out);
    deflateEnd(&z);
    if (done != Z_STREAM_END) {
        throw std::runtime_error("zlib deflate failed");
    }
    return out;
}

// "%.<digits>g" of a value, as the reference formats numbers in files (util/text.cpp:31-35).
std::string Number(double value, int digits) {
    char text[40];
    std::snprintf(text, sizeof text, "%.*g", digits, value);
    return text;
}

// The short value of GeoKey `key` stored in the directory itself, if any.
bool ShortKey(const std::uint16_t* dir, std::size_t n, std::uint16_t key, std::uint16_t& value) {
    if (dir == nullptr || n < 4) {
        return false;
    }
    for (std::size_t k = 0; k < dir[3] && 4 + 4 * k + 3 < n; ++k) {
        const std::uint16_t* e = dir + 4 + 4 * k;
        if (e[0] == key && e[1] == 0) {
            value = e[3];
            return true;
        }
    }
    return false;
}

void WriteGeoTiff(const tgx_geotiff& g, const float* values) {
    const std::string path = g.path != nullptr ? g.path : "";
    const auto fail = [&](const std::string& why) {
        throw std::invalid_argument(path + ": " + why);
    };
    if (g.width == 0 || g.height == 0 || g.bands == 0 || values == nullptr) {
        fail("cannot write an empty raster");
    }
    if (g.compression > TGX_TIFF_DEFLATE) {
        fail("unknown compression; use none, LZW or Deflate");
    }
    if (g.tile_size == 0 || g.tile_size % 16 != 0) {
        fail("the tile size must be a positive multiple of 16");
    }
    const std::size_t side = g.tile_size;
    const std::size_t across = (g.width + side - 1) / side;
    const std::size_t down = (g.height + side - 1) / side;
    const std::size_t tiles = across * down * g.bands;
    const std::uint64_t raw = std::uint64_t(tiles) * side * side * 4 + (1u << 20);
    const bool big = g.big_tiff != 0 || raw >= (std::uint64_t{1} << 32);
    const std::size_t word = big ? 8 : 4;
    const float fill = g.has_nodata ? g.nodata : 0.0f;

    // Each tile's bytes, encoded on threads.
    std::vector<std::vector<std::uint8_t>> encoded(tiles);
    const std::size_t plane = std::size_t(g.width) * g.height;
    ForEachIndex(tiles, g.threads, [&](std::size_t t) {
        const std::size_t band = t / (across * down);
        const std::size_t ty = (t % (across * down)) / across;
        const std::size_t tx = t % across;
        std::vector<std::uint8_t> bytes(side * side * 4);
        for (std::size_t y = 0; y < side; ++y) {
            const std::size_t row = ty * side + y;
            for (std::size_t x = 0; x < side; ++x) {
                const std::size_t col = tx * side + x;
                const float v = row < g.height && col < g.width
                                    ? values[band * plane + row * g.width + col]
                                    : fill;
                std::memcpy(bytes.data() + (y * side + x) * 4, &v, 4);
            }
        }
        if (g.compression != TGX_TIFF_NONE) {
            FloatPredictor(bytes, side);
            bytes = g.compression == TGX_TIFF_LZW ? Lzw(bytes) : Deflate(bytes);
        }
        encoded[t] = std::move(bytes);
    });

    std::ofstream file(path, std::ios::binary);
    if (!file) {
        fail("cannot open the file for writing");
    }
    std::vector<std::uint8_t> header{'I', 'I', static_cast<std::uint8_t>(big ? 43 : 42), 0};
    if (big) {
        header.insert(header.end(), {8, 0, 0, 0});
    }
    const std::size_t ifdAt = header.size();
    header.resize(header.size() + word, 0);
    file.write(reinterpret_cast<const char*>(header.data()), std::streamsize(header.size()));
    std::vector<std::uint64_t> offsets(tiles), counts(tiles);
    std::uint64_t end = header.size();
    for (std::size_t t = 0; t < tiles; ++t) {
        offsets[t] = end;
        counts[t] = encoded[t].size();
        file.write(reinterpret_cast<const char*>(encoded[t].data()),
                   std::streamsize(encoded[t].size()));
        end += encoded[t].size();
        if (end & 1u) {
            file.put(0);
            ++end;
        }
        std::vector<std::uint8_t>().swap(encoded[t]);
    }
    if (!big && end >= (std::uint64_t{1} << 32) - (1u << 20)) {
        fail("data too large for classic TIFF; set big_tiff");
    }

    Ifd ifd(big);
    const auto bands = static_cast<std::uint16_t>(g.bands);
    ifd.Put<std::uint32_t>(kTagWidth, kTypeLong, {g.width});
    ifd.Put<std::uint32_t>(kTagLength, kTypeLong, {g.height});
    ifd.Put<std::uint16_t>(kTagBitsPerSample, kTypeShort, std::vector<std::uint16_t>(bands, 32));
    const std::uint16_t codes[] = {1, 5, 8};
    ifd.Put<std::uint16_t>(kTagCompression, kTypeShort, {codes[g.compression]});
    ifd.Put<std::uint16_t>(kTagPhotometric, kTypeShort, {1});
    ifd.Put<std::uint16_t>(kTagSamplesPerPixel, kTypeShort, {bands});
    ifd.Put<std::uint16_t>(kTagPlanar, kTypeShort, {static_cast<std::uint16_t>(bands == 1 ? 1 : 2)});
    if (g.compression != TGX_TIFF_NONE) {
        ifd.Put<std::uint16_t>(kTagPredictor, kTypeShort, {3});
    }
    ifd.Put<std::uint32_t>(kTagTileWidth, kTypeLong, {g.tile_size});
    ifd.Put<std::uint32_t>(kTagTileLength, kTypeLong, {g.tile_size});
    if (big) {
        ifd.Put<std::uint64_t>(kTagTileOffsets, kTypeLong8, offsets);
        ifd.Put<std::uint64_t>(kTagTileByteCounts, kTypeLong8, counts);
    } else {
        ifd.Put<std::uint32_t>(kTagTileOffsets, kTypeLong,
                               std::vector<std::uint32_t>(offsets.begin(), offsets.end()));
        ifd.Put<std::uint32_t>(kTagTileByteCounts, kTypeLong,
                               std::vector<std::uint32_t>(counts.begin(), counts.end()));
    }
    if (bands > 1) {
        ifd.Put<std::uint16_t>(kTagExtraSamples, kTypeShort,
                               std::vector<std::uint16_t>(bands - 1u, 0));
    }
    ifd.Put<std::uint16_t>(kTagSampleFormat, kTypeShort, std::vector<std::uint16_t>(bands, 3));

    // The pixel-to-world transform (GDAL order: x0, dx/col, dx/row, y0, dy/col, dy/row), stored
    // unless it is the identity and there are no keys; a raster whose keys say PixelIsPoint is
    // stored half a pixel on, at its first cell's centre.
    const double* m = g.transform;
    const bool keys = g.geokey_count > 0;
    const bool identity = m[0] == 0.0 && m[1] == 1.0 && m[2] == 0.0 && m[3] == 0.0 &&
                          m[4] == 0.0 && m[5] == 1.0;
    if (keys || !identity) {
        std::uint16_t raster = 0;
        const bool point = ShortKey(g.geokeys, g.geokey_count, 1025, raster) && raster == 2;
        double x0 = m[0], y0 = m[3];
        if (point) {
            x0 += 0.5 * (m[1] + m[2]);
            y0 += 0.5 * (m[4] + m[5]);
        }
        if (m[2] == 0.0 && m[4] == 0.0) {
            ifd.Put<double>(kTagPixelScale, kTypeDouble, {m[1], -m[5], 0.0});
            ifd.Put<double>(kTagTiepoint, kTypeDouble, {0.0, 0.0, 0.0, x0, y0, 0.0});
        } else {
            ifd.Put<double>(kTagTransformation, kTypeDouble,
                            {m[1], m[2], 0.0, x0, m[4], m[5], 0.0, y0, 0.0, 0.0, 0.0, 0.0, 0.0,
                             0.0, 0.0, 1.0});
        }
    }
    if (keys) {
        ifd.Put<std::uint16_t>(kTagGeoKeys, kTypeShort,
                               std::vector<std::uint16_t>(g.geokeys, g.geokeys + g.geokey_count));
        if (g.geo_double_count > 0) {
            ifd.Put<double>(kTagGeoDoubles, kTypeDouble,
                            std::vector<double>(g.geo_doubles,
                                                g.geo_doubles + g.geo_double_count));
        }
        if (g.geo_ascii != nullptr && g.geo_ascii[0] != '\0') {
            ifd.PutText(kTagGeoAscii, g.geo_ascii);
        }
    }
    if (g.has_nodata) {
        ifd.PutText(kTagNoData, Number(static_cast<double>(g.nodata), 9));
    }
    const std::vector<std::uint8_t> table = ifd.Bytes(end);
    file.write(reinterpret_cast<const char*>(table.data()), std::streamsize(table.size()));
    file.seekp(std::streamoff(ifdAt));
    file.write(reinterpret_cast<const char*>(&end), std::streamsize(word));
    if (!file) {
        fail("write error");
    }
}

void WriteCraterCsv(const char* path, const tgx_crater* craters, std::size_t n,
                    const std::uint8_t* manual, const std::uint8_t* replace,
                    const tgx_crater_attributes* attributes) {
    std::ostringstream text;
    text << "x,y,diameter,age_gyr,id,manual,replace";
    if (attributes != nullptr) {
        text << ",floor_depth,rim_height,depth_ratio,freshness,complex";
    }
    text << '\n';
    for (std::size_t i = 0; i < n; ++i) {
        const tgx_crater& c = craters[i];
        text << Number(c.x, 17) << ',' << Number(c.y, 17) << ',' << Number(c.diameter, 17) << ','
             << Number(c.age_gyr, 17) << ',' << c.id << ','
             << (manual != nullptr && manual[i] ? '1' : '0') << ','
             << (replace != nullptr && replace[i] ? '1' : '0');
        if (attributes != nullptr) {
            const tgx_crater_attributes& a = attributes[i];
            text << ',' << Number(a.floor_depth, 17) << ',' << Number(a.rim_height, 17) << ','
                 << Number(a.depth_ratio, 17) << ',' << Number(a.freshness, 17) << ','
                 << (a.complex_ ? '1' : '0');
        }
        text << '\n';
    }
    std::ofstream file(path, std::ios::binary);
    const std::string bytes = text.str();
    file.write(bytes.data(), std::streamsize(bytes.size()));
    if (!file) {
        throw std::runtime_error(std::string("Cannot write ") + path);
    }
}

}  // namespace
}  // namespace tgh

extern "C" {

int tgx_write_geotiff(const tgx_geotiff* spec, const float* values, char* err, size_t errLen) {
    return tgh::Guard(err, errLen, [&] {
        if (spec == nullptr) {
            throw std::invalid_argument("tgx_write_geotiff: null spec");
        }
        tgh::WriteGeoTiff(*spec, values);
    });
}

int tgx_write_crater_csv(const char* path, const tgx_crater* craters, size_t count,
                         const uint8_t* manual, const uint8_t* replace,
                         const tgx_crater_attributes* attributes, char* err, size_t errLen) {
    return tgh::Guard(err, errLen, [&] {
        if (path == nullptr || (count > 0 && craters == nullptr)) {
            throw std::invalid_argument("tgx_write_crater_csv: null argument");
        }
        tgh::WriteCraterCsv(path, craters, count, manual, replace, attributes);
    });
}

}  // extern "C"
