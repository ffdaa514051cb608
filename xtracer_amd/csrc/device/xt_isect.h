// Primitive intersection on the device.
//
// FP32 tests (used by every traversal; triangles with the watertight test)
// decide robust cases and raise a
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
    float mark_t;     // smallest t of a deferred candidate (INF: none)
    uint32_t c0, c1;  // deferred candidates for the FP64 re-check (c1 = XT_CAND_OVERFLOW: more than two)
};

#define XT_CAND_OVERFLOW 0xfffffffdu

// Defers primitive `prim` at about t to the FP64 decision.
__device__ __forceinline__ void defer(Best &b, float t, uint32_t prim) {
    b.mark_t = fminf(b.mark_t, t);
    if (prim == b.c0 || prim == b.c1) return;
    if (b.c0 == 0xffffffffu) b.c0 = prim;
    else if (b.c1 == 0xffffffffu) b.c1 = prim;
    else b.c1 = XT_CAND_OVERFLOW;
}

// A candidate hit at t. The reference breaks exact ties by object, then by
// soup index; the FP32 pass defers every candidate within the tie band,
// exact ties included, to the FP64 re-check, which orders them that way.
__device__ __forceinline__ void offer(Best &b, float t, uint32_t prim) {
    // a kd-tree references a primitive from every leaf it overlaps: meeting
    // the current best again is not a tie
    if (prim == b.prim) return;
    if (t < b.t) {
        if (b.t - t <= XT_TIE * t) defer(b, t, b.prim);  // the displaced best is within the tie band
        b.t = t;
        b.prim = prim;
    } else if (t - b.t <= XT_TIE * b.t) {
        defer(b, b.t, prim);
    }
}

// Per-ray shear of the watertight ray/triangle test (Woop, Benthin, Wald,
// "Watertight Ray/Triangle Intersection", JCGT 2(1), 2013): with kz the
// largest direction component and (kx, ky) the others in winding order, the
// ray becomes the +z axis of a sheared frame in which p - o sits at
// (q[kx] - sx q[kz], q[ky] - sy q[kz], sz q[kz]). Triangles that share an
// edge evaluate its 2D edge function from the same transformed vertices, so
// a ray can cross a mesh between two triangles only by hitting one of them.
struct Shear {
    float sx, sy, sz;
    float dz;      // |d[kz]|: |n.d| = |det| * dz / |e1 x e2|
    uint32_t k;    // kx | ky << 2 | kz << 4
};

__device__ __forceinline__ float comp3(f3 v, uint32_t k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

__device__ __forceinline__ Shear make_shear(f3 d) {
    const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    const uint32_t kz = ax > ay ? (ax > az ? 0u : 2u) : (ay > az ? 1u : 2u);
    uint32_t kx = kz == 2u ? 0u : kz + 1u;
    uint32_t ky = kx == 2u ? 0u : kx + 1u;
    const float dk = comp3(d, kz);
    if (dk < 0.f) {  // keep the winding
        const uint32_t t = kx;
        kx = ky;
        ky = t;
    }
    Shear s;
    s.k = kx | ky << 2 | kz << 4;
    if (dk == 0.f) {  // zero direction: no triangle is hit
        s.sx = s.sy = s.sz = s.dz = 0.f;
        return s;
    }
    s.sz = 1.f / dk;
    s.sx = comp3(d, kx) * s.sz;
    s.sy = comp3(d, ky) * s.sz;
    s.dz = fabsf(dk);
    return s;
}

// A point relative to the origin in the sheared frame: (x, y) and the
// unscaled depth p[kz].
__device__ __forceinline__ void shear3(const Shear &sh, f3 p, float &x, float &y, float &z) {
    z = comp3(p, sh.k >> 4);
    x = comp3(p, sh.k & 3u) - sh.sx * z;
    y = comp3(p, (sh.k >> 2) & 3u) - sh.sy * z;
}

#pragma clang fp contract(off)
// 2D edge function p*q - r*s with both products rounded: the triangle on the
// other side of an edge computes the exact negation.
__device__ __forceinline__ float edge2(float p, float q, float r, float s) { return p * q - r * s; }
#pragma clang fp contract(on)

// FP32 watertight test of one triangle. Returns true when the ray grazes it
// (|n.d| < XT_GRAZE): FP32 cannot place that hit and the caller decides it
// in FP64.
__device__ __forceinline__ bool tri32(const f3 o, const Shear &sh, float tsmall, const float4 a, const float4 b,
                                      const float4 c, uint32_t prim, Best &best) {
    const uint32_t tf = __float_as_uint(a.w);
    if ((XT_PRIM_FLAGS(tf) & XT_PF_DEGENERATE) || sh.dz == 0.f) return false;
    float ax, ay, az, bx, by, bz, cx, cy, cz;
    shear3(sh, sub(mk(a.x, a.y, a.z), o), ax, ay, az);
    shear3(sh, sub(mk(b.x, b.y, b.z), o), bx, by, bz);
    shear3(sh, sub(mk(c.x, c.y, c.z), o), cx, cy, cz);
    const float U = edge2(cx, by, cy, bx), V = edge2(ax, cy, ay, cx), W = edge2(bx, ay, by, ax);
    const float det = U + V + W;
    if (fabsf(det) * sh.dz * c.w < XT_GRAZE) return true;  // |n.d| small: the reference threshold is 1e-8
    const float inv = 1.f / det;
    const float t = (U * az + V * bz + W * cz) * sh.sz * inv;
    const float m = fminf(fminf(V * inv, W * inv), U * inv);
    if (t < -tsmall || t > best.t * (1.f + XT_TIE)) return false;
    if (m > -XT_BAND && m < XT_BAND) defer(best, t, prim);
    if (m < 0.f) return false;
    if (t < tsmall) {
        defer(best, fmaxf(t, 0.f), prim);
        return false;
    }
    offer(best, t, prim);
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
        defer(best, fmaxf((float)-bb, 0.f), prim);
        if (disc <= 0.0) return;
    } else if (disc <= 0.0) {
        return;
    }
    if (fabs(cc) <= 1e-6 * r2) {  // origin on the surface
        defer(best, 0.f, prim);
        return;
    }
    if (cc < 0.0 || bb >= 0.0) return;  // inside, or sphere behind
    const float t = (float)(cc / (sqrt(disc) - bb));
    if (t > best.t * (1.f + XT_TIE)) return;
    if (t < tsmall) {
        defer(best, t, prim);
        return;
    }
    offer(best, t, prim);
}

__device__ __forceinline__ void plane32(const f3 o, const f3 d, float tsmall, const XtPlane &pl, uint32_t id,
                                        Best &best) {
    const f3 n = mk(pl.normal[0], pl.normal[1], pl.normal[2]);
    const float ndd = dot3(n, d);
    const float nl = sqrtf(dot3(n, n));
    if (fabsf(ndd) < XT_GRAZE * nl) {
        if (nl > 0.f) defer(best, 0.f, id);
        return;
    }
    const float t = dot3(sub(mk(pl.point[0], pl.point[1], pl.point[2]), o), n) / ndd;
    if (t < -tsmall || t > best.t * (1.f + XT_TIE)) return;
    if (t < tsmall) {
        defer(best, fmaxf(t, 0.f), id);
        return;
    }
    offer(best, t, id);
}

// ------------------------------------------------------------ FP64 predicates

struct d3 {
    double x, y, z;
};

// Conservative FP32 screen for the FP64 re-test: true only when the
// watertight FP32 test puts the triangle clearly outside (5% of an edge
// beyond it), clearly behind the origin, or clearly beyond `limit`, with the
// FP32 error (about 4e-7 * distance / (edge * |n.d|) relative) under 1%.
__device__ __forceinline__ bool screen_out32(const f3 o, const Shear &sh, const float4 a, const float4 b,
                                             const float4 c, float limit) {
    if (sh.dz == 0.f) return false;
    float ax, ay, az, bx, by, bz, cx, cy, cz;
    const f3 A = sub(mk(a.x, a.y, a.z), o);
    shear3(sh, A, ax, ay, az);
    shear3(sh, sub(mk(b.x, b.y, b.z), o), bx, by, bz);
    shear3(sh, sub(mk(c.x, c.y, c.z), o), cx, cy, cz);
    const float U = edge2(cx, by, cy, bx), V = edge2(ax, cy, ay, cx), W = edge2(bx, ay, by, ax);
    const float det = U + V + W;
    const float g = fabsf(det) * sh.dz * c.w;  // |n.d|
    const float dist = fmaxf(fmaxf(fabsf(A.x), fabsf(A.y)), fabsf(A.z)) + sqrtf(1.f / c.w);
    if (!(g > 0.f) || 4e-7f * dist * sqrtf(c.w) > 0.01f * g) return false;
    const float inv = 1.f / det;
    const float m = fminf(fminf(U * inv, V * inv), W * inv);
    const float t = (U * az + V * bz + W * cz) * sh.sz * inv;
    return m < -0.05f || t < -0.05f * dist || t > limit * 1.05f;
}

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

// Closest-hit record of a triangle in FP64 from its FP32 vertices: the
// reference's plane-equation t = -n.(o - v0) / n.d (triangle.cc:84-94) and
// the weights (u, v) of v1 and v2 at o + t d. The FP32 test's t loses up to
// about 1e-7 * distance / (edge * |n.d|) relative.
__device__ __forceinline__ void tri_hit64(const f3 o, const f3 d, const float4 a, const float4 b, const float4 c,
                                          float &t, float &u, float &v) {
    const d3 A = {a.x, a.y, a.z};
    const d3 e1 = dsub(d3{b.x, b.y, b.z}, A), e2 = dsub(d3{c.x, c.y, c.z}, A);
    const d3 x = dcross(e1, e2);
    const d3 n = dnorm(x);
    const d3 D = {d.x, d.y, d.z}, O = {o.x, o.y, o.z};
    const double tt = -ddot(n, dsub(O, A)) / ddot(n, D);
    const d3 pa = dsub(dadd(O, dscl(D, tt)), A);
    const double xx = ddot(x, x);
    t = (float)tt;
    u = xx > 0.0 ? (float)(ddot(dcross(pa, e2), x) / xx) : 0.f;
    v = xx > 0.0 ? (float)(ddot(dcross(e1, pa), x) / xx) : 0.f;
}

#pragma clang fp contract(on)
