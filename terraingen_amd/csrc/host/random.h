// SPDX-License-Identifier: BSD-3-Clause
//
// Counter-based random numbers for host placement: Philox4x32-10 (Salmon et al. 2011, the
// Random123 generator), a stream that hands out its words in order, and the stream tags the
// reference keys each purpose with (include/terraingen/random/{philox,stream,streams}.h). Every
// draw is a pure function of (seed, stream, counter words), so placements are the same on any
// machine, thread count or tiling.
#pragma once

#include <cmath>
#include <cstdint>
#include <stdexcept>

namespace tgh {

struct Words4 {
    std::uint32_t x = 0, y = 0, z = 0, w = 0;
};

// Philox4x32 with 10 rounds: multipliers 0xD2511F53 and 0xCD9E8D57, Weyl key increments
// 0x9E3779B9 and 0xBB67AE85 (Random123's constants).
inline Words4 Philox(Words4 c, std::uint32_t k0, std::uint32_t k1) {
    for (int round = 0; round < 10; ++round) {
        if (round > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const std::uint64_t p0 = std::uint64_t(0xD2511F53u) * c.x;
        const std::uint64_t p1 = std::uint64_t(0xCD9E8D57u) * c.z;
        c = {std::uint32_t(p1 >> 32) ^ c.y ^ k0, std::uint32_t(p1),
             std::uint32_t(p0 >> 32) ^ c.w ^ k1, std::uint32_t(p0)};
    }
    return c;
}

// Stream tags (include/terraingen/random/streams.h:15-31): ASCII words per purpose.
constexpr std::uint32_t kCraterStream = 0x43524154u;
constexpr std::uint32_t kCraterAgeStream = 0x43524147u;
constexpr std::uint32_t kCraterIdStream = 0x43524944u;
constexpr std::uint32_t kCraterIrregularityStream = 0x49525247u;
constexpr std::uint32_t kCraterTextureStream = 0x54455854u;
constexpr std::uint32_t kCraterNoiseStream = 0x434E4F53u;
constexpr std::uint32_t kCraterSlopeStream = 0x534C4F50u;
constexpr std::uint32_t kCraterTerraceStream = 0x54455252u;
constexpr std::uint32_t kCraterPeakStream = 0x5045414Bu;
constexpr std::uint32_t kBoulderCandidateStream = 0x524F434Bu;
constexpr std::uint32_t kBoulderAttributeStream = 0x41545452u;

// The words of Philox((a, b, c, block), (seed, stream)) in order, four per block
// (src/random/stream.cpp:8-27).
class Stream {
  public:
    Stream(std::uint32_t seed, std::uint32_t stream, std::uint32_t a = 0, std::uint32_t b = 0,
           std::uint32_t c = 0)
        : m_seed(seed), m_stream(stream), m_counter{a, b, c, 0u} {}

    std::uint32_t NextU32() {
        if (m_used == 4) {
            m_block = Philox(m_counter, m_seed, m_stream);
            ++m_counter.w;
            m_used = 0;
        }
        const std::uint32_t words[4] = {m_block.x, m_block.y, m_block.z, m_block.w};
        return words[m_used++];
    }

    // 53 random bits: 27 from one word, 26 from the next, as a double in [0, 1).
    double NextDouble() {
        const std::uint64_t high = NextU32() >> 5u;
        const std::uint64_t low = NextU32() >> 6u;
        return static_cast<double>((high << 26u) | low) * 0x1p-53;
    }

  private:
    std::uint32_t m_seed, m_stream;
    Words4 m_counter;
    Words4 m_block{};
    int m_used = 4;
};

// SubstreamSeed (stream.h:60-63): a per-purpose seed derived from a layer seed.
inline std::uint32_t SubstreamSeed(std::uint32_t seed, std::uint32_t stream, std::uint32_t a = 0) {
    return Philox({a, 0u, 0u, 0xFFFFFFFFu}, seed, stream).x;
}

// ReseedWord (stream.h:66-69).
inline std::uint32_t ReseedWord(std::uint64_t seed, std::uint32_t field) {
    return Philox({field, 0u, 0u, 0xFFFFFFFEu}, std::uint32_t(seed), std::uint32_t(seed >> 32)).x;
}

inline std::uint64_t JoinWords(std::uint32_t high, std::uint32_t low) {
    return (std::uint64_t(high) << 32) | low;
}

// The PCG hash and its 64-bit key form (include/terraingen/noise/hash.h:33-37,67-69).
inline std::uint32_t PcgHash(std::uint32_t v) {
    const std::uint32_t state = v * 747796405u + 2891336453u;
    const std::uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

inline std::uint32_t HashWord64(std::uint64_t word, std::uint32_t seed) {
    return PcgHash(std::uint32_t(word) ^ PcgHash(std::uint32_t(word >> 32u) ^ seed));
}

// Poisson counts (src/random/samplers.cpp:16-63): multiplication of uniforms below a mean of
// 10, Hormann's transformed rejection (PTRS, 1993) above.
inline std::uint64_t SamplePoisson(Stream& random, double mean) {
    if (!(std::isfinite(mean) && mean >= 0.0 && mean <= 1e9)) {
        throw std::invalid_argument("Poisson mean must be finite and in [0, 1e9]");
    }
    if (mean == 0.0) {
        return 0;
    }
    if (mean < 10.0) {
        const double limit = std::exp(-mean);
        std::uint64_t count = 0;
        double product = random.NextDouble();
        while (product > limit) {
            ++count;
            product *= random.NextDouble();
        }
        return count;
    }
    const double sqrtMean = std::sqrt(mean);
    const double logMean = std::log(mean);
    const double b = 0.931 + 2.53 * sqrtMean;
    const double a = -0.059 + 0.02483 * b;
    const double inverseAlpha = 1.1239 + 1.1328 / (b - 3.4);
    const double vr = 0.9277 - 3.6224 / (b - 2.0);
    while (true) {
        const double u = random.NextDouble() - 0.5;
        const double v = random.NextDouble();
        const double us = 0.5 - std::fabs(u);
        const double k = std::floor((2.0 * a / us + b) * u + mean + 0.43);
        if (us >= 0.07 && v <= vr) {
            return static_cast<std::uint64_t>(k);
        }
        if (k < 0.0 || (us < 0.013 && v > us)) {
            continue;
        }
        if (std::log(v) + std::log(inverseAlpha) - std::log(a / (us * us) + b) <=
            -mean + k * logMean - std::lgamma(k + 1.0)) {
            return static_cast<std::uint64_t>(k);
        }
    }
}

// Box-Muller from two uniforms (samplers.cpp:65-74).
inline double StandardNormal(double u1, double u2) {
    return std::sqrt(-2.0 * std::log1p(-u1)) * std::cos(2.0 * 3.14159265358979323846 * u2);
}

}  // namespace tgh
