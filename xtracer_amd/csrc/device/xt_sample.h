// Direction sampling and shading rules of the reference, in FP32.
//   hemisphere(n, dir)  lib/nmath/sample.inl:28-59
//   lobe(n, dir, e)     lib/nmath/sample.inl:61-90
//   diffuse(n)          lib/nmath/sample.inl:92-107 (rejection against the
//                       L4 ball: |p*p|^2 < 1, then normalize(n + p))
// Colour arithmetic clamps every channel at 0 after each operation, like the
// ColorRGBf setters (lib/nimg/color.inl:48-66).
#pragma once
#include "xt_isect.h"
#include "xt_rng.h"

#define XT_PI_F 3.14159265358979323846f
#define XT_DIFFUSE_MAX_TRIES 64

// Sequential draws from one stream.
struct Draws {
    uint32_t stream, i;
    __device__ __forceinline__ float next() { return xt_unit32(xt_draw(stream, i++)); }
};

// Cosine-weighted direction about n (the AO ray). `dir` scales the result by
// max(0, n.dir) before the final normalisation.
__device__ __forceinline__ f3 smp_hemisphere(f3 normal, f3 dir, Draws &r) {
    const float u = r.next();
    const float v = r.next();
    const float phi = acosf(sqrtf(u));
    const float theta = 2.f * XT_PI_F * v;
    const float sp = sinf(phi);
    const f3 dl = mk(cosf(theta) * sp, cosf(phi), sinf(theta) * sp);
    const f3 n = norm3(normal);
    // normalize(cross(n + 1e-8, n)) == normalize(cross((1,1,1), n)) up to 1e-8
    const f3 x = norm3(cross3(mk(1.f, 1.f, 1.f), n));
    const f3 z = norm3(cross3(x, n));
    const f3 w = add(add(scl(x, dl.x), scl(n, dl.y)), scl(z, dl.z));
    const float ndotl = dot3(n, norm3(dir));
    return norm3(scl(w, ndotl > 0.f ? ndotl : 0.f));
}

__device__ __forceinline__ f3 smp_diffuse(f3 normal, Draws &r) {
    f3 p = mk(0.f, 0.f, 0.f);
    for (int k = 0; k < XT_DIFFUSE_MAX_TRIES; ++k) {
        const float a = r.next(), b = r.next(), c = r.next();
        p = mk(2.f * a - 1.f, 2.f * b - 1.f, 2.f * c - 1.f);
        const f3 q = mk(p.x * p.x, p.y * p.y, p.z * p.z);
        float l = sqrtf(dot3(q, q));
        l *= l;
        if (l < 1.f) break;
    }
    return norm3(add(p, normal));
}

// Phong lobe about the mirror direction of `dir` (pointing away from the
// surface) around `normal`.
__device__ __forceinline__ f3 smp_lobe(f3 normal, f3 dir, float e, Draws &r) {
    const f3 i = norm3(dir);
    const f3 n = norm3(normal);
    const f3 refl = sub(scl(n, 2.f * dot3(i, n)), i);
    f3 c0 = mk(1.f, 0.f, 0.f), c1 = mk(0.f, 1.f, 0.f), c2 = mk(0.f, 0.f, 1.f);
    if (1.f - dot3(dir, normal) > 1e-8f) {
        const f3 iv = norm3(cross3(dir, refl));
        const f3 kv = cross3(refl, iv);
        c0 = iv;
        c1 = refl;
        c2 = kv;
    }
    const float u = r.next();
    const float v = r.next();
    const float theta = 2.f * XT_PI_F * u;
    const float phi = acosf(powf(v, 1.f / (e + 1.f)));
    const float sp = sinf(phi);
    const f3 l = mk(cosf(theta) * sp, cosf(phi), sinf(theta) * sp);
    // columns c0, c1, c2: vc = M * l
    const f3 vc = mk(c0.x * l.x + c1.x * l.y + c2.x * l.z, c0.y * l.x + c1.y * l.y + c2.y * l.z,
                     c0.z * l.x + c1.z * l.y + c2.z * l.z);
    float vdotr = dot3(refl, vc);
    if (vdotr < 0.f) vdotr = 0.f;
    return norm3(scl(vc, powf(vdotr, e)));
}

__device__ __forceinline__ f3 cmul(f3 a, f3 b) {
    return mk(fmaxf(a.x * b.x, 0.f), fmaxf(a.y * b.y, 0.f), fmaxf(a.z * b.z, 0.f));
}
__device__ __forceinline__ f3 cscale(f3 a, float s) {
    return mk(fmaxf(a.x * s, 0.f), fmaxf(a.y * s, 0.f), fmaxf(a.z * s, 0.f));
}


