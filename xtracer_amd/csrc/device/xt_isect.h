// Primitive intersection on the device.
//
// FP32 tests (used by every traversal) decide robust cases and raise a
// "re-test" mark for a ray whenever a decision might differ from the
// reference's FP64 predicate: a barycentric coordinate within XT_BAND of an
// edge, a grazing angle, a near-tangent sphere, a tiny t, or two candidates
// within a relative 2e-6 of each other. Marked rays are traced again with the
// FP64 predicates below, which restate src/xtcore/math/{triangle,sphere,
// plane}.cc operation for operation (compiled with FP contraction off so the
// roundings are the reference's).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../xt_layout.h"

#define XT_BAND 2e-5f        // barycentric band around edges
#define XT_GRAZE 1e-5f       // |cos| below which the FP32 test defers
#define XT_TSMALL 1e-5f      // relative t below which the FP32 test defers
#define XT_TIE 2e-6f         // relative t difference treated as a possible tie
#define XT_EPS64 1e-8        // nmath EPSILON (lib/nmath/precision.h:87-89)

struct f3 {
    float x, y, z;
};
__device__ __forceinline__ f3 mk(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ f3 sub(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 add(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 scl(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross3(f3 a, f3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ f3 norm3(f3 a) {
    float l = sqrtf(dot3(a, a));
    return l != 0.f ? scl(a, 1.f / l) : a;
}

// Running closest-hit state of one ray.
struct Best {
    float t;
    uint32_t prim;
    uint32_t obj;
    float u, v;
    float mark_t;    // smallest t of a deferred candidate (INF: none)
    uint32_t graze;  // a grazing triangle left to the FP64 graze pass (XT_NONE: none)
};

__device__ __forceinline__ void offer(Best &b, float t, uint32_t prim, uint32_t obj, float u, float v) {
    // a kd-tree references a primitive from every leaf it overlaps: meeting
    // the current best again is not a tie
    if (prim == b.prim) return;
    // ties go to the earlier object, then to the earlier soup index
    if (t < b.t || (t == b.t && (obj < b.obj || (obj == b.obj && prim < b.prim)))) {
        if (b.t - t <= XT_TIE * t) b.mark_t = fminf(b.mark_t, t);  // runner-up within the tie band
        b.t = t;
        b.prim = prim;
        b.obj = obj;
        b.u = u;
        b.v = v;
    } else if (t - b.t <= XT_TIE * b.t) {
        b.mark_t = fminf(b.mark_t, b.t);
    }
}

// FP32 triangle test (Moller-Trumbore) with deferral marks.
// Returns true when the ray grazes the triangle (|n.d| < XT_GRAZE): FP32
// cannot place that hit and the caller decides it in FP64.
__device__ __forceinline__ bool tri32(const f3 o, const f3 d, float tsmall, const float4 a, const float4 b,
                                      const float4 c, uint32_t prim, Best &best) {
    const uint32_t tf = __float_as_uint(a.w);
    if (XT_PRIM_FLAGS(tf) & XT_PF_DEGENERATE) return false;
    const f3 v0 = mk(a.x, a.y, a.z);
    const f3 e1 = sub(mk(b.x, b.y, b.z), v0);
    const f3 e2 = sub(mk(c.x, c.y, c.z), v0);
    const f3 pv = cross3(d, e2);
    const float det = dot3(e1, pv);
    if (fabsf(det) * c.w < XT_GRAZE) return true;  // |n.d| small: the reference threshold is 1e-8
    const float inv = 1.f / det;
    const f3 tv = sub(o, v0);
    const float u = dot3(tv, pv) * inv;
    const f3 qv = cross3(tv, e1);
    const float v = dot3(d, qv) * inv;
    const float t = dot3(e2, qv) * inv;
    const float w = 1.f - u - v;
    const float m = fminf(fminf(u, v), w);
    if (t < -tsmall || t > best.t * (1.f + XT_TIE)) return false;
    if (m > -XT_BAND && m < XT_BAND) best.mark_t = fminf(best.mark_t, t);
    if (m < 0.f) return false;
    if (t < tsmall) {
        best.mark_t = fminf(best.mark_t, fmaxf(t, 0.f));
        return false;
    }
    offer(best, t, prim, __float_as_uint(b.w), u, v);
    return false;
}

// Sphere test. The reference keeps only the smaller root and needs it
// positive (sphere.cc:35-55), so rays starting inside a sphere miss it. The
// arithmetic is FP64 on the FP32 inputs: origin - centre is then exact and
// |oc|^2 - r^2 keeps its accuracy for origins near the surface (spheres are
// few, so this costs little).
__device__ __forceinline__ void sph32(const f3 o, const f3 d, float tsmall, const float4 a, const float4 b,
                                      uint32_t prim, Best &best) {
    const uint32_t tf = __float_as_uint(a.w);
    if (XT_PRIM_FLAGS(tf) & XT_PF_TINY_SPHERE) return;
    const double ox = (double)o.x - (double)a.x, oy = (double)o.y - (double)a.y, oz = (double)o.z - (double)a.z;
    const double bb = ox * (double)d.x + oy * (double)d.y + oz * (double)d.z;
    const double r2 = (double)b.x * (double)b.x;
    const double cc = (ox * ox + oy * oy + oz * oz) - r2;
    const double disc = bb * bb - cc;
    const double scale = fmax(bb * bb, r2);
    if (fabs(disc) <= 1e-6 * scale) {
        best.mark_t = fminf(best.mark_t, fmaxf((float)-bb, 0.f));
        if (disc <= 0.0) return;
    } else if (disc <= 0.0) {
        return;
    }
    if (fabs(cc) <= 1e-6 * r2) {  // origin on the surface
        best.mark_t = fminf(best.mark_t, 0.f);
        return;
    }
    if (cc < 0.0 || bb >= 0.0) return;  // inside, or sphere behind
    const float t = (float)(cc / (sqrt(disc) - bb));
    if (t > best.t * (1.f + XT_TIE)) return;
    if (t < tsmall) {
        best.mark_t = fminf(best.mark_t, t);
        return;
    }
    offer(best, t, prim, __float_as_uint(b.w), 0.f, 0.f);
}

__device__ __forceinline__ void plane32(const f3 o, const f3 d, float tsmall, const XtPlane &pl, uint32_t id,
                                        Best &best) {
    const f3 n = mk(pl.normal[0], pl.normal[1], pl.normal[2]);
    const float ndd = dot3(n, d);
    const float nl = sqrtf(dot3(n, n));
    if (fabsf(ndd) < XT_GRAZE * nl) {
        if (nl > 0.f) best.mark_t = fminf(best.mark_t, 0.f);
        return;
    }
    const float t = dot3(sub(mk(pl.point[0], pl.point[1], pl.point[2]), o), n) / ndd;
    if (t < -tsmall || t > best.t * (1.f + XT_TIE)) return;
    if (t < tsmall) {
        best.mark_t = fminf(best.mark_t, fmaxf(t, 0.f));
        return;
    }
    offer(best, t, id, pl.object, 0.f, 0.f);
}

// ------------------------------------------------------------ FP64 predicates

struct d3 {
    double x, y, z;
};

#pragma clang fp contract(off)

__device__ __forceinline__ d3 dsub(d3 a, d3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ d3 dadd(d3 a, d3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ d3 dscl(d3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ double ddot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ d3 dcross(d3 a, d3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ double dlen(d3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
__device__ __forceinline__ d3 dnorm(d3 a) {
    double l = dlen(a);
    return l != 0.0 ? d3{a.x / l, a.y / l, a.z / l} : a;
}

// Triangle::intersection (triangle.cc:82-123) with calc_normal (:142-147) and
// calc_barycentric (:149-184). Returns t, or -1 for a miss; bary gets bc.
__device__ __noinline__ double tri64(const d3 o, const d3 d, const XtPrim64 &p, double bc[3]) {
    const d3 A = {p.a[0], p.a[1], p.a[2]}, B = {p.b[0], p.b[1], p.b[2]}, C = {p.c[0], p.c[1], p.c[2]};
    const d3 normal = dnorm(dcross(dsub(B, A), dsub(C, A)));
    const double ndd = ddot(normal, d);
    if (fabs(ndd) < XT_EPS64) return -1.0;
    const d3 vo = dsub(o, A);
    const double t = -ddot(normal, vo) / ndd;
    if (t < XT_EPS64) return -1.0;
    const d3 pos = dadd(o, dscl(d, t));
    const d3 v1 = dsub(B, A), v2 = dsub(C, A);
    const d3 x = dcross(v1, v2);
    const d3 nrm = dnorm(x);
    const double area = fabs(ddot(x, nrm)) * 0.5;
    double b0 = 0, b1 = 0, b2 = 0;
    if (!(area < XT_EPS64)) {
        const d3 p0 = dsub(A, pos), p1 = dsub(B, pos), p2 = dsub(C, pos);
        b0 = fabs(ddot(dcross(p1, p2), nrm)) * 0.5 / area;
        b1 = fabs(ddot(dcross(p2, p0), nrm)) * 0.5 / area;
        b2 = fabs(ddot(dcross(p0, p1), nrm)) * 0.5 / area;
    }
    const double sum = b0 + b1 + b2;
    if (sum < 1.0 - XT_EPS64 || sum > 1.0 + XT_EPS64) return -1.0;
    bc[0] = b0;
    bc[1] = b1;
    bc[2] = b2;
    return t;
}

// Sphere::intersection, active branch (sphere.cc:35-61).
__device__ __noinline__ double sph64(const d3 o, const d3 d, const XtPrim64 &p) {
    const d3 c = {p.a[0], p.a[1], p.a[2]};
    const double r = p.b[0];
    const double b = 2 * ddot(dsub(o, c), d);
    const d3 nc = {-c.x, -c.y, -c.z};
    const double cq = ddot(c, c) + ddot(o, o) + 2 * ddot(nc, o) - r * r;
    const double disc = (b * b - 4 * cq);
    if (disc > 0.0) {
        const double s = sqrt(disc);
        const double t1 = (-b + s) / 2.0;
        const double t2 = (-b - s) / 2.0;
        const double t = t1 < t2 ? t1 : t2;
        if (t > XT_EPS64) return t;
    }
    return -1.0;
}

// Plane::intersection (plane.cc:19-61).
__device__ __noinline__ double plane64(const d3 o, const d3 d, const XtPlane &pl) {
    const d3 n = {pl.normal64[0], pl.normal64[1], pl.normal64[2]};
    const double ndd = ddot(n, d);
    if (fabs(ndd) < XT_EPS64) return -1.0;
    const d3 v = {pl.point64[0], pl.point64[1], pl.point64[2]};
    const double t = ddot(dsub(v, o), n) / ndd;
    if (t < XT_EPS64) return -1.0;
    return t;
}

// Plane-equation t of a triangle in FP64 from its FP32 vertices (the
// reference's t = -n.(o - v0) / n.d, triangle.cc:84-94). The FP32 t loses
// about 1e-7 / |n.d| relative at grazing angles; hits with |n.d| < 0.05 get
// this value instead.
__device__ __forceinline__ float tri_t64(const f3 o, const f3 d, const float4 a, const float4 b, const float4 c) {
    const d3 A = {a.x, a.y, a.z};
    const d3 n = dnorm(dcross(dsub(d3{b.x, b.y, b.z}, A), dsub(d3{c.x, c.y, c.z}, A)));
    const d3 D = {d.x, d.y, d.z};
    const double ndd = ddot(n, D);
    return (float)(-ddot(n, dsub(d3{o.x, o.y, o.z}, A)) / ndd);
}

#pragma clang fp contract(on)
