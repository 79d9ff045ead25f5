// Point geometry of OP-PIC's opp_cluster.h (opp_lib/include/opp_cluster.h:42-126, 254-266) that
// application mesh loaders use without MPI: a 3-D point record, centroids, distances, the signed
// planar triangle area and a point-in-triangle test, and a tetrahedron volume. The header's k-means
// and block clustering serve only the MPI rank partitioning, which this single-GPU library does not
// build. FEM-PIC's fempic_defs.h includes this header unconditionally
// (apps_codegen/app_fempic_cg/fempic_defs.h:42).
#pragma once

#include <cmath>
#include <vector>

namespace opp {

// Three doubles in x, y, z order: mesh loaders reinterpret a node's `double pos[3]` as a Point3D.
struct Point3D {
    double x = 0.0;
    double y = 0.0;
    double z = 0.0;
    bool operator==(const Point3D& o) const { return x == o.x && y == o.y && z == o.z; }
};

inline bool isEqual(const Point3D& a, const Point3D& b) { return a == b; }

inline double euclideanDistance3D(const Point3D& a, const Point3D& b) {
    const double dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    return std::sqrt(dx * dx + dy * dy + dz * dz);
}

inline double euclideanDistancePlaneXY(const Point3D& a, const Point3D& b) {
    const double dx = a.x - b.x, dy = a.y - b.y;
    return std::sqrt(dx * dx + dy * dy);
}

// Arithmetic means, summed in argument order and divided once.
inline Point3D getTriangleCentroid3D(const Point3D& a, const Point3D& b, const Point3D& c) {
    Point3D m;
    m.x = (a.x + b.x + c.x) / 3.0;
    m.y = (a.y + b.y + c.y) / 3.0;
    m.z = (a.z + b.z + c.z) / 3.0;
    return m;
}

inline Point3D getTetraCentroid3D(const Point3D& a, const Point3D& b, const Point3D& c, const Point3D& d) {
    Point3D m;
    m.x = (a.x + b.x + c.x + d.x) / 4.0;
    m.y = (a.y + b.y + c.y + d.y) / 4.0;
    m.z = (a.z + b.z + c.z + d.z) / 4.0;
    return m;
}

// Signed area of the triangle's projection on the xy plane (positive when counter-clockwise).
inline double calculateTriangleArea(const Point3D& a, const Point3D& b, const Point3D& c) {
    return 0.5 * ((b.x - a.x) * (c.y - a.y) - (c.x - a.x) * (b.y - a.y));
}

// True when p's xy projection lies in the triangle's (edges included): all three sub-areas share the
// sign of the whole.
inline bool isPointInTriangle(const Point3D& p, const Point3D& a, const Point3D& b, const Point3D& c) {
    const double whole = calculateTriangleArea(a, b, c);
    const double u = calculateTriangleArea(p, b, c) / whole;
    const double v = calculateTriangleArea(a, p, c) / whole;
    const double w = calculateTriangleArea(a, b, p) / whole;
    return u >= 0 && v >= 0 && w >= 0;
}

// Unsigned volume |(b-a) . ((c-a) x (d-a))| / 6.
inline double calculateTetraVolume(const Point3D& a, const Point3D& b, const Point3D& c, const Point3D& d) {
    const double ux = b.x - a.x, uy = b.y - a.y, uz = b.z - a.z;
    const double vx = c.x - a.x, vy = c.y - a.y, vz = c.z - a.z;
    const double wx = d.x - a.x, wy = d.y - a.y, wz = d.z - a.z;
    const double triple = ux * (vy * wz - vz * wy) - uy * (vx * wz - vz * wx) + uz * (vx * wy - vy * wx);
    return std::fabs(triple) / 6.0;
}

}  // namespace opp
