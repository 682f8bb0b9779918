// SPDX-License-Identifier: BSD-3-Clause
//
// Placement cells on a sphere: the faces of a cube through the body, each split into n x n cells of
// equal angle (the reference's CubeSphereCells on Surface::Sphere,
// src/geometry/cube_sphere_cells.cpp, include/terraingen/geometry/cells.h:113-140), and the
// sphere's longitude-latitude regions (Surface::Expand / Contains / Coordinates,
// src/geometry/surface.cpp:45-50,286-330, include/terraingen/geometry/surface.h:172-182).
//
// A crater's ID and every draw of its cell's stream depend on the cell's index, its solid angle and
// the mapping of uniform numbers to points in it, so these follow the reference's definitions and
// floating-point steps: cell (i, j) is column i mod n and row j of face i / n (faces +x, -x, +y,
// -y, +z, -z with their u, v axes), n = ceil(R (pi / 2) / L) for cells of side L, edges at
// tan(-pi/4 + (pi/2) k / n), areas from the gnomonic rectangle's solid angle, and sample points
// uniform in solid angle (a bisection along u, a closed form along v).
#pragma once

#include <cstdint>
#include <vector>

namespace tgh {

struct CellIndex {
    std::int32_t i, j;
    bool operator<(const CellIndex& o) const {
        return i != o.i ? i < o.i : j < o.j;
    }
    bool operator==(const CellIndex& o) const {
        return i == o.i && j == o.j;
    }
};

// A half-open box of x, y: meters on a plane, longitude and latitude in degrees on a sphere.
struct Box {
    double minX, minY, maxX, maxY;
    Box Expanded(double m) const {
        return {minX - m, minY - m, maxX + m, maxY + m};
    }
    bool Contains(double x, double y) const {
        return x >= minX && x < maxX && y >= minY && y < maxY;
    }
};

struct Vec3 {
    double x, y, z;
};

// Longitude-latitude regions of a sphere of `radius` meters.
class SphereRegions {
  public:
    explicit SphereRegions(double radius);

    double Radius() const {
        return m_radius;
    }
    // Meters of arc per degree.
    double UnitLength() const {
        return m_unitLength;
    }
    // A box holding every point within `meters` of `region`: latitudes clamped to [-90, 90], and
    // every longitude once the box reaches a pole or grows past a turn; longitudes not wrapped.
    Box Expand(const Box& region, double meters) const;
    // Whether `region` holds (lon, lat), a longitude also counting a turn either way and a box
    // reaching 90 degrees holding the north pole.
    bool Contains(const Box& region, double lon, double lat) const;
    // Longitude and latitude (degrees) of a nonzero direction; -180 reads as 180.
    void Coordinates(const Vec3& d, double& lon, double& lat) const;

  private:
    double m_radius, m_unitLength;
};

// The unit vector at (lon, lat) degrees.
Vec3 UnitVector(double lonDeg, double latDeg);

class CubeCells {
  public:
    explicit CubeCells(double radius);

    const SphereRegions& Regions() const {
        return m_regions;
    }
    // Cells per face edge for cells of side `size` meters.
    std::int32_t PerEdge(double size) const;
    // The area of a cell in m^2, from its solid angle.
    double Area(const CellIndex& cell, double size) const;
    // The point at uniform (u, v) in [0, 1)^2 of a cell, uniform in solid angle over it.
    void Sample(double u, double v, const CellIndex& cell, double size, double& lon,
                double& lat) const;
    // Every cell reaching into `region` (degrees), plus some next to them, sorted by i then j.
    std::vector<CellIndex> Covering(const Box& region, double size) const;

  private:
    struct Face {
        Vec3 normal, u, v;
    };
    struct Rect {
        double a1, a2, b1, b2;  // gnomonic bounds along the face's u and v
    };
    static const Face kFaces[6];

    Rect Bounds(const CellIndex& cell, std::int32_t n) const;
    CellIndex Home(double lonDeg, double latDeg, double size) const;
    void Ring(const CellIndex& cell, double size, std::vector<CellIndex>& out) const;
    void AcrossEdge(const Face& face, std::int32_t col, std::int32_t row, std::int32_t n,
                    double size, std::vector<CellIndex>& out) const;
    std::vector<CellIndex> Near(const Box& box, double size) const;
    std::vector<CellIndex> Seeds(const Box& box, double size) const;
    std::vector<CellIndex> Reaching(const Box& box, const std::vector<CellIndex>& seeds,
                                    const std::vector<CellIndex>& cells, double size) const;

    SphereRegions m_regions;
};

}  // namespace tgh
