// SPDX-License-Identifier: BSD-3-Clause
//
// Crater (and boulder-candidate) placement on a plane or a sphere, drawn cell by cell in
// power-of-two size bands exactly as the reference draws them
// (src/layers/craters/generation/population.cpp), so the same seed gives the same list, in the
// same order, on any grid, tiling or thread count:
//   bands      [2^k, 2^(k+1)) clipped to [minDiameter, maxDiameter), drawn in cells
//              cellSizeFactor x 2^(k+1) meters across (population.cpp:348-356): squares on a
//              plane, cube-sphere cells on a sphere (cube_cells.h);
//   per cell   a Philox stream keyed (seed, kCraterStream, i, j, k) gives a Poisson count of the
//              band's expected craters, then per crater a diameter by inverse CDF and a uniform
//              centre; a second stream keyed by kCraterAgeStream gives ages; IDs are
//              Philox((i, j, k, index), (seed, kCraterIdStream)) (population.cpp:381-408,24-31);
//   keeping    craters centred in the region grown by marginRadii of their radius, in band, cell
//              and index order (population.cpp:174-259).
// Densities are power laws N(>=D) = c D^-e per m^2 (include/terraingen/distribution/power_law.h).

#include "common.h"
#include "cube_cells.h"
#include "placement.h"
#include "random.h"

#include <atomic>
#include <cstdlib>
#include <exception>
#include <limits>
#include <memory>
#include <mutex>
#include <thread>
#include <utility>

namespace tgh {

double PowerLawDensity::ExpectedCount(double area, double minD, double maxD) const {
    if (!(minD > 0.0 && std::isfinite(minD) && maxD > minD) || std::isnan(maxD)) {
        throw std::invalid_argument("Diameter interval must satisfy 0 < minDiameter < maxDiameter");
    }
    const double upper = std::isinf(maxD) ? 0.0 : Cumulative(maxD);
    return area * (Cumulative(minD) - upper);
}

// SizeFrequencyDistribution::SampleDiameter (src/distribution/size_frequency.cpp:131-140) with
// the power law's closed-form inverse.
double PowerLawDensity::SampleDiameter(Stream& random, double minD, double maxD) const {
    const double count = ExpectedCount(1.0, minD, maxD);
    if (!(count > 0.0)) {
        throw std::invalid_argument("The distribution has no mass in the diameter interval");
    }
    const double top = Cumulative(minD);
    const double target = top - random.NextDouble() * count;
    return std::clamp(Invert(target), minD, maxD);
}

double Ages::Lifetime(double diameter) const {
    if (!(diameter > 0.0 && std::isfinite(diameter))) {
        throw std::domain_error("Diameter must be positive and finite");
    }
    return surface != nullptr ? surface->LifetimeGyr(diameter) : lifetimeGyr;
}

namespace {

// BisectRising (src/util/numerics.h:63-127) with the Illinois bracket (:30-61): the interval of
// `Bisect` on `value(x) >= target` for a non-decreasing value, reached with fewer evaluations.
struct Illinois {
    double a, b, fa, fb;
    int side = 0;
    double Next() const {
        const double c = b - fb * (b - a) / (fb - fa);
        return c > a && c < b ? c : 0.5 * (a + b);
    }
    void Update(double c, double fc) {
        if ((fc < 0.0) == (fa < 0.0)) {
            a = c;
            fa = fc;
            fb *= side == -1 ? 0.5 : 1.0;
            side = -1;
        } else {
            b = c;
            fb = fc;
            fa *= side == 1 ? 0.5 : 1.0;
            side = 1;
        }
    }
};

template <class Value>
std::pair<double, double> BisectRising(double low, double high, int steps, double tolerance,
                                       double target, const Value& value) {
    constexpr int kEstimateSteps = 12;
    const double margin = 1e-13 * (high - low);
    Illinois bracket{low, high, value(low) - target, value(high) - target};
    const bool belowKnown = bracket.fa < 0.0;
    const bool aboveKnown = !(bracket.fb < 0.0);
    double below = low;
    double above = high;
    if (belowKnown && aboveKnown) {
        double c = low;
        for (int i = 0; i < kEstimateSteps; ++i) {
            const double previous = c;
            c = bracket.Next();
            bracket.Update(c, value(c) - target);
            if (i > 0 && std::fabs(c - previous) <= margin) {
                break;
            }
        }
        const double a = bracket.a;
        const double b = bracket.b;
        below = a;
        above = b;
        for (const double spread : {margin, 64.0 * margin}) {
            if (b - a <= 2.0 * spread) {
                break;
            }
            if (c - spread > a && value(c - spread) < target) {
                below = c - spread;
            }
            if (c + spread < b && !(value(c + spread) < target)) {
                above = c + spread;
            }
            if (above - below <= 4.0 * spread) {
                break;
            }
        }
    }
    const double aboveFrom = aboveKnown ? above + margin : std::numeric_limits<double>::infinity();
    const double belowTo = belowKnown ? below - margin : -std::numeric_limits<double>::infinity();
    for (int i = 0; i < steps && high - low > tolerance; ++i) {
        const double middle = 0.5 * (low + high);
        bool isAbove = middle >= aboveFrom;
        if (!isAbove && !(middle <= belowTo)) {
            isAbove = !(value(middle) < target);
        }
        high = isAbove ? middle : high;
        low = isAbove ? low : middle;
    }
    return {low, high};
}

}  // namespace

double Ages::Sample(Stream& random, double diameter) const {
    const double lifetime = Lifetime(diameter);
    double age = 0.0;
    if (surface != nullptr) {
        // SteadyStateAges::SampleAgeAt (steady_state.cpp:53-59): a formation age by bisection on
        // the chronology's N1(T) (SampleFormationAge, :17-24).
        if (lifetime != 0.0) {
            const double target = random.NextDouble() * NeukumChronology(lifetime);
            const auto found = BisectRising(0.0, lifetime, 100, 1e-15 * lifetime, target,
                                            [](double t) { return NeukumChronology(t); });
            age = 0.5 * (found.first + found.second);
        }
    } else if (lifetime != 0.0) {
        age = random.NextDouble() * lifetime;  // UniformAges::SampleAgeAt
    }
    return std::clamp(age, 0.0, lifetime);
}

double LunarDensity::ExpectedCount(double area, double minD, double maxD) const {
    if (!(minD > 0.0 && std::isfinite(minD) && maxD > minD) || std::isnan(maxD)) {
        throw std::invalid_argument("Diameter interval must satisfy 0 < minDiameter < maxDiameter");
    }
    if (!(area >= 0.0 && std::isfinite(area))) {
        throw std::invalid_argument("Area must be finite and not negative");
    }
    if (std::isinf(maxD)) {
        throw std::domain_error("An infinite maxDiameter needs an unbounded valid range");
    }
    const double upper = m_surface.DensityChecked(maxD);
    return area * (m_surface.DensityChecked(minD) - upper);
}

double InvertCumulative(const std::function<double(double)>& cumulative, double target,
                        double minD, double maxD) {
    double low = std::log(minD);
    double high = std::isinf(maxD) ? low + 1.0 : std::log(maxD);
    while (std::isinf(maxD) && cumulative(std::exp(high)) > target) {
        high += 1.0;
    }
    if (target > 0.0) {
        // InvertBySecant (size_frequency.cpp:34-61): Illinois steps on log N over log D.
        const double logTarget = std::log(target);
        const auto excess = [&](double u) { return std::log(cumulative(std::exp(u))) - logTarget; };
        Illinois bracket{low, high, excess(low), excess(high)};
        if (!(bracket.fa > 0.0)) {
            return std::exp(low);
        }
        if (!(bracket.fb < 0.0)) {
            return std::exp(high);
        }
        double u = low;
        bool converged = false;
        for (int i = 0; i < 100; ++i) {
            const double previous = u;
            u = bracket.Next();
            const double g = excess(u);
            if (std::isnan(g)) {
                break;
            }
            if (g == 0.0 || std::fabs(u - previous) <= 1e-14 * std::fmax(1.0, std::fabs(u))) {
                converged = true;
                break;
            }
            bracket.Update(u, g);
        }
        if (converged) {
            return std::exp(u);
        }
        low = bracket.a;
        high = bracket.b;
    }
    // Bisect (numerics.h:22-29) on !(N(exp u) > target).
    const double tolerance = 1e-15 * std::fmax(1.0, std::fabs(high));
    for (int i = 0; i < 200 && high - low > tolerance; ++i) {
        const double middle = 0.5 * (low + high);
        (!(cumulative(std::exp(middle)) > target) ? high : low) = middle;
    }
    return std::exp(0.5 * (low + high));
}

double LunarDensity::Invert(double target, double minD, double maxD) const {
    return InvertCumulative([this](double d) { return m_surface.Density(d); }, target, minD, maxD);
}

double SummedPowerLaws::Cumulative(double d) const {
    if (!(d > 0.0 && std::isfinite(d))) {
        throw std::domain_error("Diameter must be positive and finite");
    }
    double sum = 0.0;
    for (const Term& t : m_terms) {
        sum += t.weight * (t.law.coefficient * std::pow(d, -t.law.exponent));
    }
    return sum;
}

double SummedPowerLaws::Incremental(double d) const {
    if (!(d > 0.0 && std::isfinite(d))) {
        throw std::domain_error("Diameter must be positive and finite");
    }
    double sum = 0.0;
    for (const Term& t : m_terms) {
        sum += t.weight * t.law.Incremental(d);
    }
    return sum;
}

double SummedPowerLaws::ExpectedCount(double area, double minD, double maxD) const {
    if (!(minD > 0.0 && std::isfinite(minD) && maxD > minD) || std::isnan(maxD)) {
        throw std::invalid_argument("Diameter interval must satisfy 0 < minDiameter < maxDiameter");
    }
    if (!(area >= 0.0 && std::isfinite(area))) {
        throw std::invalid_argument("Area must be finite and not negative");
    }
    const double upper = std::isinf(maxD) ? 0.0 : Cumulative(maxD);
    return area * (Cumulative(minD) - upper);
}

double SummedPowerLaws::SampleDiameter(Stream& random, double minD, double maxD) const {
    const double count = ExpectedCount(1.0, minD, maxD);
    if (!(count > 0.0)) {
        throw std::invalid_argument("The distribution has no mass in the diameter interval");
    }
    const double top = Cumulative(minD);
    const double target = top - random.NextDouble() * count;
    const auto cumulative = [this](double d) {
        double sum = 0.0;
        for (const Term& t : m_terms) {
            sum += t.weight * (t.law.coefficient * std::pow(d, -t.law.exponent));
        }
        return sum;
    };
    return std::clamp(InvertCumulative(cumulative, target, minD, maxD), minD, maxD);
}

double LunarDensity::SampleDiameter(Stream& random, double minD, double maxD) const {
    const double count = ExpectedCount(1.0, minD, maxD);
    if (!(count > 0.0)) {
        throw std::invalid_argument("The distribution has no mass in the diameter interval");
    }
    const double top = m_surface.DensityChecked(minD);
    const double target = top - random.NextDouble() * count;
    return std::clamp(Invert(target, minD, maxD), minD, maxD);
}

std::vector<Band> CraterBands(double minD, double maxD, double cellSizeFactor) {
    std::vector<Band> bands;
    for (int k = std::ilogb(minD); std::ldexp(1.0, k) < maxD; ++k) {
        const double low = std::fmax(minD, std::ldexp(1.0, k));
        const double high = std::fmin(maxD, std::ldexp(1.0, k + 1));
        bands.push_back({k, low, high, cellSizeFactor * std::ldexp(1.0, k + 1)});
    }
    return bands;
}

namespace {

std::int32_t CellCoordinate(double value) {
    if (!(std::fabs(value) < 2147483647.0)) {
        throw std::invalid_argument("Region is too far from the origin for its cell size");
    }
    return static_cast<std::int32_t>(value);
}

// Square cells on a plane (PlanarCells, src/geometry/planar_cells.cpp): cell (i, j) of side L
// covers [i L, (i + 1) L) x [j L, (j + 1) L); a region's cells in row-major order.
struct PlaneCells {
    std::vector<CellIndex> Covering(const Box& r, double size) const {
        const std::int32_t iMin = CellCoordinate(std::floor(r.minX / size));
        const std::int32_t jMin = CellCoordinate(std::floor(r.minY / size));
        const std::int32_t iMax = CellCoordinate(std::ceil(r.maxX / size) - 1.0);
        const std::int32_t jMax = CellCoordinate(std::ceil(r.maxY / size) - 1.0);
        std::vector<CellIndex> cells;
        for (std::int32_t j = jMin; j <= jMax; ++j) {
            for (std::int32_t i = iMin; i <= iMax; ++i) {
                cells.push_back({i, j});
            }
        }
        return cells;
    }
    double Area(const CellIndex&, double size) const {
        return size * size;
    }
    void Sample(double u, double v, const CellIndex& cell, double size, double& x,
                double& y) const {
        x = (cell.i + u) * size;
        y = (cell.j + v) * size;
    }
    Box Expand(const Box& b, double m) const {
        return b.Expanded(m);
    }
    bool Contains(const Box& b, double x, double y) const {
        return b.Contains(x, y);
    }
    // The order Covering lists cells in: row-major.
    static bool Before(const CellIndex& a, const CellIndex& b) {
        return a.j != b.j ? a.j < b.j : a.i < b.i;
    }
};

// Cube-sphere cells on a sphere, with longitude-latitude regions (cube_cells.h). On a sphere every
// point of a cell weighs the same, so no drawn point is thinned (CubeSphereCells::CellSampling's
// bound is 0 there).
struct SphereCells {
    CubeCells cells;
    std::vector<CellIndex> Covering(const Box& r, double size) const {
        return cells.Covering(r, size);
    }
    double Area(const CellIndex& cell, double size) const {
        return cells.Area(cell, size);
    }
    void Sample(double u, double v, const CellIndex& cell, double size, double& x,
                double& y) const {
        cells.Sample(u, v, cell, size, x, y);
    }
    Box Expand(const Box& b, double m) const {
        return cells.Regions().Expand(b, m);
    }
    bool Contains(const Box& b, double x, double y) const {
        return cells.Regions().Contains(b, x, y);
    }
    // The order Covering lists cells in: by i, then j.
    static bool Before(const CellIndex& a, const CellIndex& b) {
        return a < b;
    }
};

// Generation::Run (population.cpp:89-108,145-200): every band's cells over the region grown by the
// band's largest margin, each cell's craters drawn and kept, cells in band then cell order. Cells
// are drawn on threads and their kept craters joined in that order, as the reference's
// DrawPlannedCells joins its ParallelFor's lists, so the list is the same on any thread count.
// Over several regions, a band's cells are those any region covers, each drawn once, and a crater
// is kept when any region keeps it: the union of the regions' own lists.
template <class Cells> class Generator {
  public:
    Generator(const Cells& cells, std::vector<Box> regions, const Density& density,
              const Ages& ages, std::uint32_t seed, const GenerationOptions& o)
        : m_cells(cells), m_regions(std::move(regions)), m_density(density), m_ages(ages),
          m_seed(seed), m_o(o) {
        if (!(o.maxDiameter > o.minDiameter) || !std::isfinite(o.maxDiameter)) {
            throw std::invalid_argument("maxDiameter must be finite and exceed minDiameter");
        }
        if (!(o.minDiameter > 0.0)) {
            throw std::invalid_argument("minDiameter must be positive");
        }
        if (m_regions.empty()) {
            throw std::invalid_argument("Generation needs a region");
        }
    }

    std::vector<Crater> Run() const {
        struct Planned {
            Band band;
            std::vector<CellIndex> cells;
            // The regions whose coverings hold cells[n]: regionOf[firstRegion[n] ...
            // firstRegion[n + 1]).
            std::vector<std::uint32_t> firstRegion, regionOf;
            std::vector<Box> inner, outer;  // per region
        };
        std::vector<Planned> plan;
        for (const Band& band : CraterBands(m_o.minDiameter, m_o.maxDiameter, m_o.cellSizeFactor)) {
            if (band.high > m_o.minResolvedDiameter) {
                plan.push_back({band, {}, {}, {}, {}, {}});
            }
        }
        // Each (band, region)'s covering on threads, then each band's union of them.
        const std::size_t regions = m_regions.size();
        std::vector<std::vector<CellIndex>> coverings(plan.size() * regions);
        ForEachIndex(coverings.size(), m_o.threads, [&](std::size_t n) {
            const Band& band = plan[n / regions].band;
            coverings[n] =
                m_cells.Covering(Grown(m_regions[n % regions], Margin(band.high)), band.cellSize);
        });
        ForEachIndex(plan.size(), m_o.threads, [&](std::size_t b) {
            Planned& p = plan[b];
            std::vector<std::pair<CellIndex, std::uint32_t>> covered;
            for (std::size_t r = 0; r < regions; ++r) {
                for (const CellIndex& cell : coverings[b * regions + r]) {
                    covered.emplace_back(cell, static_cast<std::uint32_t>(r));
                }
                p.inner.push_back(Grown(m_regions[r], Margin(p.band.low)));
                p.outer.push_back(Grown(m_regions[r], Margin(p.band.high)));
            }
            if (regions > 1) {
                std::stable_sort(covered.begin(), covered.end(),
                                 [](const auto& a, const auto& c) {
                                     return Cells::Before(a.first, c.first);
                                 });
            }
            for (std::size_t n = 0; n < covered.size(); ++n) {
                if (n == 0 || !(covered[n].first == covered[n - 1].first)) {
                    p.cells.push_back(covered[n].first);
                    p.firstRegion.push_back(static_cast<std::uint32_t>(p.regionOf.size()));
                }
                p.regionOf.push_back(covered[n].second);
            }
            p.firstRegion.push_back(static_cast<std::uint32_t>(p.regionOf.size()));
        });
        std::vector<std::pair<const Planned*, std::size_t>> work;
        for (const Planned& p : plan) {
            for (std::size_t n = 0; n < p.cells.size(); ++n) {
                work.emplace_back(&p, n);
            }
        }
        // The expected count: each band's sum over its cells in order, then the bands' sum, as
        // the reference adds them.
        std::vector<double> areas(work.size());
        ForEachIndex(work.size(), m_o.threads, [&](std::size_t w) {
            const Planned& p = *work[w].first;
            areas[w] = m_cells.Area(p.cells[work[w].second], p.band.cellSize);
        });
        double expected = 0.0;
        for (std::size_t w = 0; w < work.size();) {
            const Planned& p = *work[w].first;
            const double perArea = m_density.ExpectedCount(1.0, p.band.low, p.band.high);
            double band = 0.0;
            for (; w < work.size() && work[w].first == &p; ++w) {
                band += perArea * areas[w];
            }
            expected += band;
        }
        if (expected > m_o.maxExpectedCraters) {
            throw std::invalid_argument("Expected crater count " + std::to_string(expected) +
                                        " exceeds maxExpectedCraters");
        }
        std::vector<std::vector<Crater>> lists(work.size());
        ForEachIndex(work.size(), m_o.threads, [&](std::size_t w) {
            const Planned& p = *work[w].first;
            const std::size_t n = work[w].second;
            const std::vector<Crater> drawn = DrawCell(p.band, p.cells[n]);
            double largest = 0.0;
            for (const Crater& d : drawn) {
                largest = std::fmax(largest, d.diameter);
            }
            for (const Crater& c : drawn) {
                for (std::uint32_t q = p.firstRegion[n]; q < p.firstRegion[n + 1]; ++q) {
                    const std::uint32_t r = p.regionOf[q];
                    if (Keeps(r, p.band, p.inner[r], p.outer[r], largest, c)) {
                        lists[w].push_back(c);
                        break;
                    }
                }
            }
        });
        std::vector<Crater> kept;
        kept.reserve(static_cast<std::size_t>(std::min(expected, 1e8)));
        for (const std::vector<Crater>& some : lists) {
            kept.insert(kept.end(), some.begin(), some.end());
        }
        return kept;
    }

  private:
    double Margin(double diameter) const {
        return 0.5 * m_o.marginRadii * diameter;
    }

    Box Grown(const Box& region, double margin) const {
        return m_cells.Expand(region, margin);
    }

    bool InRegion(std::size_t r, const Crater& c, double margin) const {
        return m_cells.Contains(Grown(m_regions[r], margin), c.x, c.y);
    }

    // GenerateCellCraters (population.cpp:381-408): a Philox stream keyed (seed, kCraterStream,
    // i, j, k) gives a Poisson count of the band's expected craters over the cell's area, then per
    // crater its diameter by inverse CDF and a point of the cell from two uniforms; a second stream
    // gives ages; IDs are Philox((i, j, k, index), (seed, kCraterIdStream)).
    std::vector<Crater> DrawCell(const Band& band, const CellIndex& cell) const {
        const auto k = static_cast<std::uint32_t>(band.exponent);
        const auto i = static_cast<std::uint32_t>(cell.i);
        const auto j = static_cast<std::uint32_t>(cell.j);
        Stream geometry(m_seed, kCraterStream, i, j, k);
        Stream ageStream(m_seed, kCraterAgeStream, i, j, k);
        const double perArea = m_density.ExpectedCount(1.0, band.low, band.high);
        const double area = m_cells.Area(cell, band.cellSize);
        const std::uint64_t count = SamplePoisson(geometry, perArea * area);
        std::vector<Crater> out;
        out.reserve(count);
        for (std::uint64_t index = 0; index < count; ++index) {
            const double d = m_density.SampleDiameter(geometry, band.low, band.high);
            const double u = geometry.NextDouble();
            const double v = geometry.NextDouble();
            double x, y;
            m_cells.Sample(u, v, cell, band.cellSize, x, y);
            const double age = m_ages.Sample(ageStream, d);
            const Words4 id = Philox({i, j, k, static_cast<std::uint32_t>(index)}, m_seed,
                                     kCraterIdStream);
            out.push_back({x, y, d, age, JoinWords(id.x, id.y)});
        }
        return out;
    }

    // Generation::DrawKept, OutsideBand and InBandRegion (population.cpp:202-238) for region r:
    // a drawn crater in the rows its cell's band can reach (`largest`: the cell's largest drawn
    // diameter), then inside the region grown by its margin.
    bool Keeps(std::size_t r, const Band& band, const Box& inner, const Box& outer,
               double largest, const Crater& c) const {
        const Box reach = largest > band.high ? Grown(m_regions[r], Margin(largest)) : outer;
        if (!(c.y >= reach.minY && c.y <= reach.maxY)) {
            return false;  // CellCraters::Rows keeps y in [minY, maxY]
        }
        const bool inBand = c.diameter >= band.low && c.diameter <= band.high;
        if (inBand && !m_cells.Contains(outer, c.x, c.y)) {
            return false;
        }
        if (!inBand) {
            return InRegion(r, c, Margin(c.diameter));
        }
        return m_cells.Contains(inner, c.x, c.y) || InRegion(r, c, Margin(c.diameter));
    }

    const Cells& m_cells;
    std::vector<Box> m_regions;
    const Density& m_density;
    const Ages& m_ages;
    std::uint32_t m_seed;
    GenerationOptions m_o;
};

}  // namespace

std::vector<Crater> GenerateCraters(double minX, double minY, double maxX, double maxY,
                                    const Density& density, const Ages& ages,
                                    std::uint32_t seed, const GenerationOptions& options) {
    if (!(maxX > minX && maxY > minY) || !std::isfinite(minX + minY + maxX + maxY)) {
        throw std::invalid_argument("Region must be finite and not empty");
    }
    const PlaneCells cells;
    return Generator<PlaneCells>(cells, {{minX, minY, maxX, maxY}}, density, ages, seed, options)
        .Run();
}

std::vector<Crater> GenerateCratersOnSphere(double radius, const std::vector<Box>& regions,
                                            const Density& density, const Ages& ages,
                                            std::uint32_t seed, const GenerationOptions& options) {
    for (const Box& r : regions) {
        if (!(r.maxX >= r.minX && r.maxY >= r.minY) ||
            !std::isfinite(r.minX + r.minY + r.maxX + r.maxY)) {
            throw std::invalid_argument("Region must be finite and not inverted");
        }
    }
    const SphereCells cells{CubeCells(radius)};
    return Generator<SphereCells>(cells, regions, density, ages, seed, options).Run();
}

}  // namespace tgh

extern "C" {

void tgx_set_host_threads(uint32_t threads) {
    tgh::HostThreads() = threads;
}

void tgx_philox4x32(const uint32_t* counter, const uint32_t* key, uint32_t* out) {
    const tgh::Words4 r = tgh::Philox({counter[0], counter[1], counter[2], counter[3]}, key[0],
                                      key[1]);
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
    out[3] = r.w;
}

namespace {

int Generate(const tgx_generation* g, const double* more, std::size_t region_count,
             tgx_crater** out, size_t* count, char* err, size_t errLen) {
    return tgh::Guard(err, errLen, [&] {
        if (g == nullptr || out == nullptr || count == nullptr) {
            throw std::invalid_argument("tgx_generate_craters: null argument");
        }
        // model 1: DefaultCraterModels' lunar equilibrium surface and steady-state ages.
        if (g->model > 1u) {
            throw std::invalid_argument("tgx_generate_craters: unknown crater model");
        }
        const tgh::PowerLawDensity powerLaw{g->sfd_coefficient, g->sfd_exponent};
        const tgh::LunarDensity lunar;
        const tgh::Density& density =
            g->model == 1u ? static_cast<const tgh::Density&>(lunar) : powerLaw;
        tgh::Ages ages{g->lifetime_gyr, g->model == 1u ? &lunar.Surface() : nullptr};
        tgh::GenerationOptions o;
        o.minDiameter = g->min_diameter;
        o.maxDiameter = g->max_diameter;
        o.marginRadii = g->margin_radii;
        o.cellSizeFactor = g->cell_size_factor > 0.0 ? g->cell_size_factor : 16.0;
        o.minResolvedDiameter = g->min_resolved_diameter;
        o.threads = g->threads;
        std::vector<tgh::Box> regions{{g->min_x, g->min_y, g->max_x, g->max_y}};
        if (region_count > 0) {
            if (g->sphere_radius <= 0.0 || more == nullptr) {
                throw std::invalid_argument("more regions are taken on a sphere only");
            }
            for (std::size_t r = 0; r < region_count; ++r) {
                regions.push_back({more[4 * r], more[4 * r + 1], more[4 * r + 2],
                                   more[4 * r + 3]});
            }
        }
        const std::vector<tgh::Crater> craters =
            g->sphere_radius > 0.0
                ? tgh::GenerateCratersOnSphere(g->sphere_radius, regions, density, ages, g->seed,
                                               o)
                : tgh::GenerateCraters(g->min_x, g->min_y, g->max_x, g->max_y, density, ages,
                                       g->seed, o);
        std::vector<tgx_crater> flat(craters.size());
        for (std::size_t n = 0; n < craters.size(); ++n) {
            flat[n] = {craters[n].x, craters[n].y, craters[n].diameter, craters[n].ageGyr,
                       craters[n].id};
        }
        *out = tgh::Export(flat);
        *count = flat.size();
    });
}

}  // namespace

int tgx_generate_craters(const tgx_generation* g, tgx_crater** out, size_t* count, char* err,
                         size_t errLen) {
    return Generate(g, nullptr, 0, out, count, err, errLen);
}

int tgx_generate_craters_regions(const tgx_generation* g, const double* more,
                                 size_t region_count, tgx_crater** out, size_t* count,
                                 char* err, size_t errLen) {
    return Generate(g, more, region_count, out, count, err, errLen);
}

}  // extern "C"
