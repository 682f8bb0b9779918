// SPDX-License-Identifier: BSD-3-Clause
//
// Reading a band of a GeoTIFF into floats (tgx_geotiff_open, tgx_geotiff_read): the DEM a
// GeoTiffSource serves (src/sources/geotiff_source.cpp), read as the reference's reader reads it
// (src/raster/geotiff/geotiff_read.cpp, tiff_codecs.cpp, tiff_format.cpp) so that its posts are
// the same floats:
//   * TIFF 6.0 and BigTIFF, either byte order; the first image only;
//   * strips or tiles, samples interleaved or in planes, 8 to 64-bit unsigned, signed or float
//     samples, each converted to float as a C++ cast converts it;
//   * no compression, LZW (MSB-first codes of 9-12 bits, widths raised one code early), PackBits
//     and Deflate; the horizontal predictor (after the samples are brought to this host's byte
//     order) and the floating-point one (TIFF Technical Note 3);
//   * the transform from ModelTransformation, else ModelPixelScale with one ModelTiepoint, moved
//     half a pixel for PixelIsPoint rasters; GTModelTypeGeoKey for a geographic CRS;
//   * GDAL's no-data tag (42113, text read with strtod) and band scale and offset from GDAL's
//     metadata tag (42112: <Item name=.. sample="n" role="scale|offset">).
// A window read decodes only the chunks it meets and applies GeoTiffPosts::Decode's rule: a
// sample that is NaN or equals the no-data value becomes NaN, any other is scaled as
// float(double(v) * scale + offset) unless the scaling is the identity.

#include "common.h"
#include "terraingen_amd.h"

#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

struct tgx_geotiff_file {
    std::string path;
    tgx_geotiff_info info{};
    std::vector<std::uint64_t> offsets, counts;
    bool planar = false;
    std::vector<double> scales, shifts;  // per band (GDAL metadata)
};

namespace tgh {
namespace {

[[noreturn]] void Bad(const std::string& why) {
    throw std::runtime_error("GeoTIFF: " + why);
}

enum : std::uint16_t {
    kWidth = 256, kLength = 257, kBits = 258, kCompression = 259, kStripOffsets = 273,
    kSamples = 277, kRowsPerStrip = 278, kStripCounts = 279, kPlanar = 284, kPredictor = 317,
    kTileWidth = 322, kTileLength = 323, kTileOffsets = 324, kTileCounts = 325,
    kSampleFormat = 339, kPixelScale = 33550, kTiepoint = 33922, kTransformation = 34264,
    kGeoKeys = 34735, kGdalMetadata = 42112, kGdalNoData = 42113,
};

// The bytes of a file with its byte order: every multi-byte read goes through Get.
class Bytes {
  public:
    explicit Bytes(const std::string& path) {
        std::ifstream in(path, std::ios::binary);
        if (!in) {
            Bad("cannot open " + path);
        }
        m_data.assign(std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>());
        if (m_data.size() < 8) {
            Bad("file too short for a TIFF header");
        }
        if (m_data[0] == 'I' && m_data[1] == 'I') {
            m_big = false;
        } else if (m_data[0] == 'M' && m_data[1] == 'M') {
            m_big = true;
        } else {
            Bad("not a TIFF file");
        }
    }

    bool BigEndian() const { return m_big; }
    std::size_t Size() const { return m_data.size(); }

    std::uint64_t Get(std::uint64_t at, unsigned n) const {
        if (at + n > m_data.size()) {
            Bad("read past the end of the file");
        }
        std::uint64_t v = 0;
        for (unsigned i = 0; i < n; ++i) {
            const std::uint64_t b = static_cast<unsigned char>(m_data[at + i]);
            v |= m_big ? b << (8 * (n - 1 - i)) : b << (8 * i);
        }
        return v;
    }

    std::vector<std::uint8_t> Slice(std::uint64_t at, std::uint64_t n) const {
        if (at + n > m_data.size() || at + n < at) {
            Bad("a tile or strip runs past the end of the file");
        }
        return std::vector<std::uint8_t>(m_data.begin() + at, m_data.begin() + at + n);
    }

  private:
    std::string m_data;
    bool m_big = false;
};

// One IFD entry: its type, count and where its values lie.
struct Entry {
    std::uint16_t type = 0;
    std::uint64_t count = 0;
    std::uint64_t at = 0;  // file offset of the values
};

unsigned TypeBytes(std::uint16_t type) {
    switch (type) {
    case 1: case 2: case 6: case 7: return 1;
    case 3: case 8: return 2;
    case 4: case 9: case 11: case 13: return 4;
    case 5: case 10: case 12: case 16: case 17: case 18: return 8;
    default: return 0;
    }
}

class Directory {
  public:
    Directory(const Bytes& file, bool bigTiff, std::uint64_t at) : m_file(file) {
        const unsigned word = bigTiff ? 8 : 4;
        const std::uint64_t n = file.Get(at, bigTiff ? 8 : 2);
        std::uint64_t e = at + (bigTiff ? 8 : 2);
        for (std::uint64_t i = 0; i < n; ++i, e += 4 + 2 * word) {
            Entry entry;
            const auto tag = static_cast<std::uint16_t>(file.Get(e, 2));
            entry.type = static_cast<std::uint16_t>(file.Get(e + 2, 2));
            entry.count = file.Get(e + 4, word);
            const unsigned size = TypeBytes(entry.type);
            if (size == 0) {
                continue;  // a type this reader does not know: the tag is skipped
            }
            entry.at = entry.count * size <= word ? e + 4 + word : file.Get(e + 4 + word, word);
            m_entries[tag] = entry;
        }
    }

    const Entry* Find(std::uint16_t tag) const {
        const auto it = m_entries.find(tag);
        return it == m_entries.end() ? nullptr : &it->second;
    }

    std::vector<std::uint64_t> Integers(const Entry& e) const {
        std::vector<std::uint64_t> out(e.count);
        const unsigned size = TypeBytes(e.type);
        for (std::uint64_t i = 0; i < e.count; ++i) {
            out[i] = m_file.Get(e.at + i * size, size);
        }
        return out;
    }

    std::vector<double> Doubles(const Entry& e) const {
        std::vector<double> out(e.count);
        for (std::uint64_t i = 0; i < e.count; ++i) {
            if (e.type == 12) {
                const std::uint64_t bits = m_file.Get(e.at + 8 * i, 8);
                std::memcpy(&out[i], &bits, 8);
            } else if (e.type == 11) {
                const auto bits = static_cast<std::uint32_t>(m_file.Get(e.at + 4 * i, 4));
                float f;
                std::memcpy(&f, &bits, 4);
                out[i] = f;
            } else {
                out[i] = static_cast<double>(m_file.Get(e.at + i * TypeBytes(e.type),
                                                        TypeBytes(e.type)));
            }
        }
        return out;
    }

    std::string Text(const Entry& e) const {
        std::string s;
        for (std::uint64_t i = 0; i < e.count; ++i) {
            const char c = static_cast<char>(m_file.Get(e.at + i, 1));
            if (c == '\0') {
                break;
            }
            s.push_back(c);
        }
        return s;
    }

    std::uint64_t Single(std::uint16_t tag, std::uint64_t fallback) const {
        const Entry* e = Find(tag);
        if (e == nullptr || e->count == 0) {
            return fallback;
        }
        return Integers(*e)[0];
    }

  private:
    const Bytes& m_file;
    std::map<std::uint16_t, Entry> m_entries;
};

// The value of `name`="..." in an XML start tag, if present.
bool Attribute(const std::string& tag, const std::string& name, std::string& value) {
    const std::string key = name + "=\"";
    const std::size_t at = tag.find(key);
    if (at == std::string::npos) {
        return false;
    }
    const std::size_t end = tag.find('"', at + key.size());
    if (end == std::string::npos) {
        return false;
    }
    value = tag.substr(at + key.size(), end - at - key.size());
    return true;
}

// GDAL's per-band scale and offset items.
void BandScaling(const std::string& xml, std::size_t bands, std::vector<double>& scales,
                 std::vector<double>& shifts) {
    scales.assign(bands, 1.0);
    shifts.assign(bands, 0.0);
    for (std::size_t at = 0; (at = xml.find("<Item", at)) != std::string::npos;) {
        const std::size_t close = xml.find('>', at);
        const std::size_t end = close == std::string::npos ? close : xml.find("</Item>", close);
        if (end == std::string::npos) {
            break;
        }
        const std::string tag = xml.substr(at, close - at);
        const std::string text = xml.substr(close + 1, end - close - 1);
        at = end;
        std::string sample, role;
        if (!Attribute(tag, "sample", sample) || !Attribute(tag, "role", role) ||
            (role != "scale" && role != "offset")) {
            continue;
        }
        char* stop = nullptr;
        const unsigned long long band = std::strtoull(sample.c_str(), &stop, 10);
        if (stop == sample.c_str() || band >= bands) {
            continue;
        }
        const double v = std::strtod(text.c_str(), &stop);
        if (stop == text.c_str() || !std::isfinite(v)) {
            continue;
        }
        (role == "scale" ? scales : shifts)[band] = v;
    }
}

// The short value of GeoKey `key`, or `fallback`.
std::uint64_t GeoKey(const std::vector<std::uint64_t>& dir, std::uint64_t key,
                     std::uint64_t fallback) {
    for (std::size_t i = 4; i + 3 < dir.size(); i += 4) {
        if (dir[i] == key && dir[i + 1] == 0) {
            return dir[i + 3];
        }
    }
    return fallback;
}

void Open(tgx_geotiff_file& f) {
    const Bytes file(f.path);
    const std::uint64_t magic = file.Get(2, 2);
    const bool bigTiff = magic == 43;
    if (magic != 42 && !bigTiff) {
        Bad("not a TIFF file");
    }
    const std::uint64_t first = bigTiff ? file.Get(8, 8) : file.Get(4, 4);
    const Directory ifd(file, bigTiff, first);
    tgx_geotiff_info& info = f.info;
    info = tgx_geotiff_info{};
    info.big_tiff = bigTiff;
    info.big_endian = file.BigEndian();
    const Entry* w = ifd.Find(kWidth);
    const Entry* h = ifd.Find(kLength);
    if (w == nullptr || h == nullptr) {
        Bad("missing image width or length");
    }
    info.width = static_cast<std::uint32_t>(ifd.Integers(*w)[0]);
    info.height = static_cast<std::uint32_t>(ifd.Integers(*h)[0]);
    if (info.width == 0 || info.height == 0) {
        Bad("empty image");
    }
    info.bands = static_cast<std::uint32_t>(ifd.Single(kSamples, 1));
    info.bits_per_sample = static_cast<std::uint32_t>(ifd.Single(kBits, 1));
    info.sample_format = static_cast<std::uint32_t>(ifd.Single(kSampleFormat, 1));
    const std::uint32_t bits = info.bits_per_sample;
    if (info.sample_format < 1 || info.sample_format > 3 ||
        (bits != 8 && bits != 16 && bits != 32 && bits != 64) ||
        (info.sample_format == 3 && bits != 32 && bits != 64)) {
        Bad("samples of " + std::to_string(bits) + " bits in format " +
            std::to_string(info.sample_format) + " are not supported");
    }
    info.compression = static_cast<std::uint32_t>(ifd.Single(kCompression, 1));
    if (info.compression != 1 && info.compression != 5 && info.compression != 8 &&
        info.compression != 32946 && info.compression != 32773) {
        Bad("compression " + std::to_string(info.compression) + " is not supported");
    }
    info.predictor = static_cast<std::uint32_t>(ifd.Single(kPredictor, 1));
    if (info.predictor < 1 || info.predictor > 3 ||
        (info.predictor == 3 && info.sample_format != 3)) {
        Bad("unsupported predictor " + std::to_string(info.predictor));
    }
    f.planar = ifd.Single(kPlanar, 1) == 2;
    info.tiled = ifd.Find(kTileWidth) != nullptr;
    const Entry* offsets = ifd.Find(info.tiled ? kTileOffsets : kStripOffsets);
    const Entry* counts = ifd.Find(info.tiled ? kTileCounts : kStripCounts);
    if (offsets == nullptr || counts == nullptr) {
        Bad("missing tile or strip offsets or byte counts");
    }
    if (info.tiled) {
        info.chunk_width = static_cast<std::uint32_t>(ifd.Single(kTileWidth, 0));
        info.chunk_height = static_cast<std::uint32_t>(ifd.Single(kTileLength, 0));
    } else {
        info.chunk_width = info.width;
        info.chunk_height = static_cast<std::uint32_t>(
            std::min<std::uint64_t>(ifd.Single(kRowsPerStrip, info.height), info.height));
    }
    if (info.chunk_width == 0 || info.chunk_height == 0) {
        Bad("zero tile or strip size");
    }
    f.offsets = ifd.Integers(*offsets);
    f.counts = ifd.Integers(*counts);
    const std::uint64_t across = (info.width + info.chunk_width - 1) / info.chunk_width;
    const std::uint64_t down = (info.height + info.chunk_height - 1) / info.chunk_height;
    if (f.offsets.size() < across * down * (f.planar ? info.bands : 1) ||
        f.counts.size() < f.offsets.size()) {
        Bad("fewer tile or strip offsets than the image needs");
    }
    // The transform: ModelTransformation, else scale and one tiepoint.
    std::vector<std::uint64_t> keys;
    if (const Entry* k = ifd.Find(kGeoKeys)) {
        keys = ifd.Integers(*k);
    }
    info.geographic = GeoKey(keys, 1024, 0) == 2;      // GTModelTypeGeoKey: geographic
    info.pixel_is_point = GeoKey(keys, 1025, 1) == 2;  // GTRasterTypeGeoKey: PixelIsPoint
    double t[6] = {0.0, 1.0, 0.0, 0.0, 0.0, 1.0};      // GDAL order
    bool referenced = false;
    if (const Entry* m = ifd.Find(kTransformation)) {
        const std::vector<double> v = ifd.Doubles(*m);
        if (v.size() < 16) {
            Bad("ModelTransformation needs 16 values");
        }
        const double g[6] = {v[3], v[0], v[1], v[7], v[4], v[5]};
        std::copy(g, g + 6, t);
        referenced = true;
    } else if (const Entry* s = ifd.Find(kPixelScale); s != nullptr) {
        if (const Entry* p = ifd.Find(kTiepoint)) {
            const std::vector<double> sc = ifd.Doubles(*s);
            const std::vector<double> tp = ifd.Doubles(*p);
            if (sc.size() < 2 || tp.size() < 6) {
                Bad("ModelPixelScale needs 2 and ModelTiepoint 6 values");
            }
            if (tp.size() > 6) {
                Bad("multiple tiepoints are not supported");
            }
            const double g[6] = {tp[3] - tp[0] * sc[0], sc[0], 0.0, tp[4] + tp[1] * sc[1], 0.0,
                                 -sc[1]};
            std::copy(g, g + 6, t);
            referenced = true;
        }
    }
    if (referenced && info.pixel_is_point) {
        // GeoTIFF places a PixelIsPoint raster's transform at its first cell's centre.
        t[0] += -1.0 * 0.5 * (t[1] + t[2]);
        t[3] += -1.0 * 0.5 * (t[4] + t[5]);
    }
    info.georeferenced = referenced;
    std::copy(t, t + 6, info.transform);
    if (const Entry* nd = ifd.Find(kGdalNoData)) {
        const std::string text = ifd.Text(*nd);
        char* stop = nullptr;
        const double v = std::strtod(text.c_str(), &stop);
        if (stop != text.c_str()) {
            info.has_no_data = 1;
            info.no_data = v;
        }
    }
    const Entry* meta = ifd.Find(kGdalMetadata);
    BandScaling(meta != nullptr ? ifd.Text(*meta) : std::string(), info.bands, f.scales, f.shifts);
}

// TIFF LZW: MSB-first codes, 256 Clear, 257 end of information, widths raised one code early.
std::vector<std::uint8_t> Unlzw(const std::vector<std::uint8_t>& in, std::size_t size) {
    if (in.size() >= 2 && in[0] == 0 && (in[1] & 1u) != 0) {
        Bad("old-style LZW is not supported");
    }
    std::vector<std::uint8_t> out;
    out.reserve(size);
    // Each code is a prefix code and its last byte; strings are emitted by walking prefixes.
    std::vector<int> prefix(4096, -1);
    std::vector<std::uint8_t> last(4096), firstByte(4096);
    std::vector<std::uint16_t> length(4096, 1);
    for (int c = 0; c < 256; ++c) {
        last[c] = firstByte[c] = static_cast<std::uint8_t>(c);
    }
    int next = 258, width = 9, previous = -1;
    std::uint64_t bit = 0;
    const std::uint64_t bits = in.size() * 8;
    const auto read = [&](int n) -> int {
        if (bit + n > bits) {
            return 257;  // treat a truncated stream as its end
        }
        int v = 0;
        for (int i = 0; i < n; ++i, ++bit) {
            v = (v << 1) | ((in[bit >> 3] >> (7 - (bit & 7))) & 1);
        }
        return v;
    };
    const auto emit = [&](int code) {
        const std::size_t start = out.size();
        out.resize(start + length[code]);
        for (int c = code, k = length[code] - 1; k >= 0; c = prefix[c], --k) {
            out[start + k] = last[c];
        }
    };
    while (out.size() < size) {
        const int code = read(width);
        if (code == 257) {
            break;
        }
        if (code == 256) {
            next = 258;
            width = 9;
            previous = -1;
            continue;
        }
        if (previous < 0) {
            if (code > 255) {
                Bad("corrupt LZW data: the first code after a Clear is not a byte");
            }
            emit(code);
            previous = code;
            continue;
        }
        if (code > next || next >= 4096) {
            Bad("corrupt LZW data: a code out of range");
        }
        std::uint8_t head;
        if (code < next) {
            emit(code);
            head = firstByte[code];
        } else {  // the code being defined: the previous string and its own first byte
            emit(previous);
            head = firstByte[previous];
            out.push_back(head);
        }
        prefix[next] = previous;
        last[next] = head;
        firstByte[next] = firstByte[previous];
        length[next] = static_cast<std::uint16_t>(length[previous] + 1);
        ++next;
        if (next + 1 >= (1 << width) && width < 12) {
            ++width;
        }
        previous = code;
    }
    if (out.size() < size) {
        Bad("LZW data decodes to fewer bytes than the chunk holds");
    }
    out.resize(size);
    return out;
}

std::vector<std::uint8_t> Unpackbits(const std::vector<std::uint8_t>& in, std::size_t size) {
    std::vector<std::uint8_t> out;
    out.reserve(size);
    for (std::size_t i = 0; i < in.size() && out.size() < size;) {
        const auto n = static_cast<std::int8_t>(in[i++]);
        if (n >= 0) {
            for (int k = 0; k <= n && i < in.size(); ++k) {
                out.push_back(in[i++]);
            }
        } else if (n != -128 && i < in.size()) {
            out.insert(out.end(), std::size_t(1 - n), in[i++]);
        }
    }
    if (out.size() < size) {
        Bad("PackBits data decodes to fewer bytes than the chunk holds");
    }
    out.resize(size);
    return out;
}

std::vector<std::uint8_t> Inflate(const std::vector<std::uint8_t>& in, std::size_t size) {
    std::vector<std::uint8_t> out(size);
    uLongf got = static_cast<uLongf>(size);
    const int rc = uncompress(out.data(), &got, in.data(), static_cast<uLong>(in.size()));
    if (rc != Z_OK && !(rc == Z_BUF_ERROR && got == size)) {
        Bad("corrupt Deflate data");
    }
    if (got < size) {
        Bad("Deflate data decodes to fewer bytes than the chunk holds");
    }
    return out;
}

// Chunk `chunk`, `rows` rows of `rowSamples` samples, decompressed, in this host's byte order,
// its predictor undone.
std::vector<std::uint8_t> Decode(const tgx_geotiff_file& f, const Bytes& file, std::size_t chunk,
                                 std::size_t rows, std::size_t rowSamples, std::size_t stride) {
    const tgx_geotiff_info& info = f.info;
    const std::size_t sampleBytes = info.bits_per_sample / 8;
    const std::size_t size = rows * rowSamples * sampleBytes;
    if (f.counts[chunk] == 0) {
        return std::vector<std::uint8_t>(size, 0);
    }
    std::vector<std::uint8_t> raw = file.Slice(f.offsets[chunk], f.counts[chunk]);
    std::vector<std::uint8_t> b;
    switch (info.compression) {
    case 5: b = Unlzw(raw, size); break;
    case 8: case 32946: b = Inflate(raw, size); break;
    case 32773: b = Unpackbits(raw, size); break;
    default:
        if (raw.size() < size) {
            Bad("a chunk holds fewer bytes than its samples");
        }
        raw.resize(size);
        b = std::move(raw);
    }
    if (info.predictor == 3) {
        // Bytes differenced along the row, the planes from the most significant byte.
        std::vector<std::uint8_t> planes(rowSamples * sampleBytes);
        for (std::size_t r = 0; r < rows; ++r) {
            std::uint8_t* row = b.data() + r * planes.size();
            for (std::size_t i = stride; i < planes.size(); ++i) {
                row[i] = static_cast<std::uint8_t>(row[i] + row[i - stride]);
            }
            std::copy(row, row + planes.size(), planes.begin());
            for (std::size_t s = 0; s < rowSamples; ++s) {
                for (std::size_t k = 0; k < sampleBytes; ++k) {
                    row[s * sampleBytes + k] = planes[(sampleBytes - 1 - k) * rowSamples + s];
                }
            }
        }
        return b;
    }
    if (info.big_endian && sampleBytes > 1) {
        for (std::size_t i = 0; i + sampleBytes <= b.size(); i += sampleBytes) {
            std::reverse(b.begin() + i, b.begin() + i + sampleBytes);
        }
    }
    if (info.predictor == 2) {
        for (std::size_t r = 0; r < rows; ++r) {
            std::uint8_t* row = b.data() + r * rowSamples * sampleBytes;
            for (std::size_t i = stride; i < rowSamples; ++i) {
                std::uint64_t a = 0, p = 0;
                std::memcpy(&a, row + i * sampleBytes, sampleBytes);
                std::memcpy(&p, row + (i - stride) * sampleBytes, sampleBytes);
                a += p;  // modulo the sample's width: only its low bytes are stored back
                std::memcpy(row + i * sampleBytes, &a, sampleBytes);
            }
        }
    }
    return b;
}

// Sample `at` of a decoded chunk, cast to float as the stored type casts.
float SampleAt(const tgx_geotiff_info& info, const std::uint8_t* p) {
    const unsigned bits = info.bits_per_sample;
    if (info.sample_format == 3) {
        if (bits == 32) {
            float v;
            std::memcpy(&v, p, 4);
            return v;
        }
        double v;
        std::memcpy(&v, p, 8);
        return static_cast<float>(v);
    }
    if (info.sample_format == 2) {
        switch (bits) {
        case 8: { std::int8_t v; std::memcpy(&v, p, 1); return static_cast<float>(v); }
        case 16: { std::int16_t v; std::memcpy(&v, p, 2); return static_cast<float>(v); }
        case 32: { std::int32_t v; std::memcpy(&v, p, 4); return static_cast<float>(v); }
        default: { std::int64_t v; std::memcpy(&v, p, 8); return static_cast<float>(v); }
        }
    }
    switch (bits) {
    case 8: return static_cast<float>(*p);
    case 16: { std::uint16_t v; std::memcpy(&v, p, 2); return static_cast<float>(v); }
    case 32: { std::uint32_t v; std::memcpy(&v, p, 4); return static_cast<float>(v); }
    default: { std::uint64_t v; std::memcpy(&v, p, 8); return static_cast<float>(v); }
    }
}

void ReadWindow(const tgx_geotiff_file& f, std::uint32_t band, std::uint64_t col0,
                std::uint64_t row0, std::uint32_t width, std::uint32_t height, double scale,
                double offset, float* out) {
    const tgx_geotiff_info& info = f.info;
    if (band >= info.bands) {
        Bad("band " + std::to_string(band) + " requested of " + std::to_string(info.bands));
    }
    if (col0 + width > info.width || row0 + height > info.height) {
        Bad("the window runs past the image");
    }
    const Bytes file(f.path);
    const std::size_t sampleBytes = info.bits_per_sample / 8;
    const std::size_t inChunk = f.planar ? 1 : info.bands;
    const std::size_t across = (info.width + info.chunk_width - 1) / info.chunk_width;
    const std::size_t down = (info.height + info.chunk_height - 1) / info.chunk_height;
    const std::size_t rowSamples = std::size_t(info.chunk_width) * inChunk;
    const bool identity = scale == 1.0 && offset == 0.0;
    const float noData = static_cast<float>(info.no_data);
    const std::size_t cx0 = col0 / info.chunk_width;
    const std::size_t cx1 = (col0 + width - 1) / info.chunk_width;
    const std::size_t cy0 = row0 / info.chunk_height;
    const std::size_t cy1 = (row0 + height - 1) / info.chunk_height;
    for (std::size_t cy = cy0; cy <= cy1; ++cy) {
        const std::size_t rows = info.tiled ? info.chunk_height
                                            : std::min<std::size_t>(info.chunk_height,
                                                                    info.height -
                                                                        cy * info.chunk_height);
        for (std::size_t cx = cx0; cx <= cx1; ++cx) {
            const std::size_t chunk =
                (f.planar ? std::size_t(band) * across * down : 0) + cy * across + cx;
            const std::vector<std::uint8_t> b = Decode(f, file, chunk, rows, rowSamples, inChunk);
            const std::uint64_t x0 = cx * info.chunk_width, y0 = cy * info.chunk_height;
            const std::uint64_t xa = std::max<std::uint64_t>(x0, col0);
            const std::uint64_t xb = std::min<std::uint64_t>(x0 + info.chunk_width, col0 + width);
            const std::uint64_t ya = std::max<std::uint64_t>(y0, row0);
            const std::uint64_t yb = std::min<std::uint64_t>(y0 + rows, row0 + height);
            for (std::uint64_t y = ya; y < yb; ++y) {
                const std::uint8_t* row = b.data() + (y - y0) * rowSamples * sampleBytes;
                for (std::uint64_t x = xa; x < xb; ++x) {
                    const std::size_t sample = (x - x0) * inChunk + (f.planar ? 0 : band);
                    float v = SampleAt(info, row + sample * sampleBytes);
                    if (std::isnan(v) || (info.has_no_data && v == noData)) {
                        v = std::numeric_limits<float>::quiet_NaN();
                    } else if (!identity) {
                        v = static_cast<float>(static_cast<double>(v) * scale + offset);
                    }
                    out[(y - row0) * width + (x - col0)] = v;
                }
            }
        }
    }
}

}  // namespace
}  // namespace tgh

extern "C" {

int tgx_geotiff_open(const char* path, tgx_geotiff_file** out, char* err, size_t errLen) {
    if (out != nullptr) {
        *out = nullptr;
    }
    return tgh::Guard(err, errLen, [&] {
        if (path == nullptr || out == nullptr) {
            throw std::invalid_argument("tgx_geotiff_open: null argument");
        }
        auto f = std::make_unique<tgx_geotiff_file>();
        f->path = path;
        tgh::Open(*f);
        *out = f.release();
    });
}

int tgx_geotiff_info_get(const tgx_geotiff_file* file, uint32_t band, tgx_geotiff_info* info,
                         double* scale, double* offset) {
    if (file == nullptr || info == nullptr || band >= file->info.bands) {
        return TG_ERR_ARGUMENT;
    }
    *info = file->info;
    if (scale != nullptr) {
        *scale = file->scales[band];
    }
    if (offset != nullptr) {
        *offset = file->shifts[band];
    }
    return TG_OK;
}

int tgx_geotiff_read(const tgx_geotiff_file* file, uint32_t band, uint64_t col0, uint64_t row0,
                     uint32_t width, uint32_t height, double scale, double offset, float* out,
                     char* err, size_t errLen) {
    return tgh::Guard(err, errLen, [&] {
        if (file == nullptr || (size_t(width) * height > 0 && out == nullptr)) {
            throw std::invalid_argument("tgx_geotiff_read: null argument");
        }
        if (!std::isfinite(scale) || !std::isfinite(offset)) {
            throw std::invalid_argument("tgx_geotiff_read: scale and offset must be finite");
        }
        if (width == 0 || height == 0) {
            return;
        }
        tgh::ReadWindow(*file, band, col0, row0, width, height, scale, offset, out);
    });
}

void tgx_geotiff_close(tgx_geotiff_file* file) {
    delete file;
}

}  // extern "C"
