// Double-precision 3-vector used by the host side (scene loading, camera set-up,
// tree builds). Operation order follows lib/nmath/vector.inl:280-500 so that
// host-side values (camera basis, mesh modifiers) are bit-identical to the
// reference when this file is compiled with -ffp-contract=off.
#pragma once
#include <cmath>

namespace xt {

struct dvec3 {
    double x = 0, y = 0, z = 0;
    dvec3() = default;
    dvec3(double a, double b, double c) : x(a), y(b), z(c) {}
    double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    double &operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};

inline dvec3 operator+(const dvec3 &a, const dvec3 &b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline dvec3 operator-(const dvec3 &a, const dvec3 &b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline dvec3 operator-(const dvec3 &a) { return {-a.x, -a.y, -a.z}; }
inline dvec3 operator*(const dvec3 &a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline dvec3 operator*(double s, const dvec3 &a) { return {a.x * s, a.y * s, a.z * s}; }
inline dvec3 operator/(const dvec3 &a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline double dot(const dvec3 &a, const dvec3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline dvec3 cross(const dvec3 &a, const dvec3 &b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double length(const dvec3 &a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
// Zero vectors are returned unchanged (lib/nmath/vector.inl:444-448).
inline dvec3 normalized(const dvec3 &a) {
    double l = length(a);
    return l != 0 ? dvec3(a.x / l, a.y / l, a.z / l) : a;
}

// Row-major 3x3 matrix, rows applied as in Vector3f::transformed
// (lib/nmath/vector.inl:504-510).
struct dmat3 {
    double m[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
};
inline dmat3 mul(const dmat3 &a, const dmat3 &b) {
    dmat3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            r.m[i][j] = a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j] + a.m[i][2] * b.m[2][j];
    return r;
}
inline dvec3 apply(const dmat3 &a, const dvec3 &v) {
    return {a.m[0][0] * v.x + a.m[0][1] * v.y + a.m[0][2] * v.z,
            a.m[1][0] * v.x + a.m[1][1] * v.y + a.m[1][2] * v.z,
            a.m[2][0] * v.x + a.m[2][1] * v.y + a.m[2][2] * v.z};
}

}  // namespace xt
