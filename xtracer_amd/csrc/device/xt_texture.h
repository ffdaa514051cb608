// Samplers on the device: textures with nearest / bilinear filtering
// (lib/nimg/sample.cc:8-53 over Pixmap::pixel_ro, pixmap.cc:44-50), the
// gradient (sampler_gradient.cc:11-17), the equirectangular
// (sampler_erp.cc:13-25) and cubemap (sampler_cubemap.cc:14-41)
// environments, each evaluated at a texture coordinate / direction `tc` in
// the reference's float/double mix. The colour a sampler returns is
// ColorRGBf, so a texel's alpha is dropped.
#pragma once
#include <hip/hip_runtime.h>

#include "xt_trace.h"

#pragma clang fp contract(off)

// Pixmap::pixel_ro: coordinates past the edge read the last row / column.
__device__ __forceinline__ f3 texel(const SceneView &S, const XtTexture &t, uint64_t x, uint64_t y) {
    x = x >= t.width ? t.width - 1 : x;
    y = y >= t.height ? t.height - 1 : y;
    const float4 v = S.texels[t.offset + y * t.width + x];
    return mk(v.x, v.y, v.z);
}

// size_t(x) for a non-negative double; NaN and values past 2^64 read as the
// largest size, which pixel_ro clamps to the last texel
__device__ __forceinline__ uint64_t to_size(double x) {
    return x >= 0.0 && x < 18446744073709551616.0 ? (uint64_t)x : ~0ull;
}

__device__ __forceinline__ f3 tex_nearest(const SceneView &S, const XtTexture &t, double u, double v) {
    u = u < 0 ? 0 : (u > 1 ? u - floor(u) : u);
    v = v < 0 ? 0 : (v > 1 ? v - floor(v) : v);
    return texel(S, t, to_size(u * (float)t.width), to_size(v * (float)t.height));
}

// ((int)f) % n as the reference computes it with n a size_t: x86-64's
// float -> int conversion (NaN and values out of range give INT_MIN), the int
// widened to 64 bits, then the unsigned 64-bit remainder. With n below 2^16
// the remainder of 2^64 + i is built from 32-bit remainders.
__device__ __forceinline__ uint32_t wrap_index(float f, uint32_t n) {
    const int i = (f >= -2147483648.f && f < 2147483648.f) ? (int)f : INT_MIN;
    if (i >= 0) return (uint32_t)i % n;
    const uint32_t r32 = (0xffffffffu % n + 1u) % n;  // 2^32 mod n
    const uint32_t r64 = (r32 * r32) % n;              // 2^64 mod n
    const uint32_t neg = (uint32_t)(-(int64_t)i) % n;  // |i| mod n
    return (r64 + n - neg) % n;
}

__device__ __forceinline__ f3 tex_bilinear(const SceneView &S, const XtTexture &t, double u, double v) {
    const uint32_t w = t.width, h = t.height;
    const float fu = (float)(u * (double)w), fv = (float)(v * (double)h);
    const int u1 = (int)wrap_index(fu, w), v1 = (int)wrap_index(fv, h);
    const int u2 = (int)((uint32_t)(u1 + 1) % w), v2 = (int)((uint32_t)(v1 + 1) % h);
    const float fracu = fu - floorf(fu), fracv = fv - floorf(fv);
    const float w1 = (1 - fracu) * (1 - fracv), w2 = fracu * (1 - fracv);
    const float w3 = (1 - fracu) * fracv, w4 = fracu * fracv;
    const f3 c1 = texel(S, t, (uint64_t)(int64_t)u1, (uint64_t)(int64_t)v1);
    const f3 c2 = texel(S, t, (uint64_t)(int64_t)u2, (uint64_t)(int64_t)v1);
    const f3 c3 = texel(S, t, (uint64_t)(int64_t)u1, (uint64_t)(int64_t)v2);
    const f3 c4 = texel(S, t, (uint64_t)(int64_t)u2, (uint64_t)(int64_t)v2);
    // ColorRGBAf * float and + (color.inl), channels kept at or above 0
    auto sc = [](f3 c, float s) { return mk(fmaxf(c.x * s, 0.f), fmaxf(c.y * s, 0.f), fmaxf(c.z * s, 0.f)); };
    auto ad = [](f3 a, f3 b) { return mk(fmaxf(a.x + b.x, 0.f), fmaxf(a.y + b.y, 0.f), fmaxf(a.z + b.z, 0.f)); };
    return ad(ad(ad(sc(c1, w1), sc(c2, w2)), sc(c3, w3)), sc(c4, w4));
}

__device__ __forceinline__ f3 tex_sample(const SceneView &S, int32_t ti, bool bilinear, double u, double v) {
    const XtTexture t = S.textures[ti];
    return bilinear ? tex_bilinear(S, t, u, v) : tex_nearest(S, t, u, v);
}

// Gradient::sample (sampler_gradient.cc:11-17) at the y of the normalized
// direction: (1.0 - t) * a + t * b, t = 0.5f * y + 1.0 rounded to float.
__device__ __forceinline__ f3 gradient_eval(const float *a, const float *b, double dy) {
    const float t = 0.5f * dy + 1.0;
    const float s = (float)(1.0 - t);
    return mk(fmaxf(fmaxf(a[0] * s, 0.f) + fmaxf(b[0] * t, 0.f), 0.f),
              fmaxf(fmaxf(a[1] * s, 0.f) + fmaxf(b[1] * t, 0.f), 0.f),
              fmaxf(fmaxf(a[2] * s, 0.f) + fmaxf(b[2] * t, 0.f), 0.f));
}

// Vector3f::normalized of (x, y, z) in FP64, y component only
__device__ __forceinline__ double unit_y(double x, double y, double z) {
    const double l = sqrt(x * x + y * y + z * z);
    return l != 0.0 ? y / l : y;
}

// ISampler::sample(tc) of a table sampler.
// The record is read field by field from the table (a cubemap's face index
// is a load at a computed address, not an index into a register copy).
__device__ __forceinline__ f3 sampler_eval(const SceneView &S, uint32_t si, double tx, double ty, double tz) {
    const XtSampler &sm = S.samplers[si];
    const uint32_t kind = sm.kind;
    if (kind == XT_SMP_TEXTURE) return tex_sample(S, sm.tex[0], sm.bilinear != 0, tx, ty);
    if (kind == XT_SMP_COLOR) return mk(sm.a[0], sm.a[1], sm.a[2]);
    // the rest read tc as a direction: Vector3f::normalized
    const double l = sqrt(tx * tx + ty * ty + tz * tz);
    const double dx = l != 0.0 ? tx / l : tx, dy = l != 0.0 ? ty / l : ty, dz = l != 0.0 ? tz / l : tz;
    if (kind == XT_SMP_GRADIENT) return gradient_eval(sm.a, sm.b, dy);
    if (kind == XT_SMP_ERP) {
        const float theta = (float)acos(dy);
        const float phi = (float)atan2(dx, -dz);
        const float u = 0.5f + phi / 6.283185307179586232;  // nmath::PI_DOUBLE
        const float v = theta / 3.14159265358979323846;
        return tex_sample(S, sm.tex[0], false, u, v);
    }
    if (kind == XT_SMP_CUBEMAP) {
        // the major axis picks the face; the other two components divided by
        // it give u and v, some of them mirrored (1 - x), per face. One pair
        // of divisions serves all six faces, with the same operands and
        // roundings as the reference's six branches. A zero or NaN direction
        // reads the top face at (0, 0).
        const double ax = fabs(dx), ay = fabs(dy), az = fabs(dz);
        int face = 2;  // top
        double den = 0.0, na = 0.0, nb = 0.0;
        bool flip_u = false, flip_v = false;
        if (ax >= ay && ax >= az) {
            if (dx != 0.0) {
                face = dx > 0.0 ? 1 : 0;
                den = dx, na = dz, nb = dy;
                flip_v = dx > 0.0;
            }
        } else if (ay >= ax && ay >= az) {
            if (dy != 0.0) {
                face = dy > 0.0 ? 2 : 3;
                den = dy, na = dx, nb = dz;
                flip_u = dy < 0.0;
                flip_v = true;
            }
        } else if (az >= ax && az >= ay) {
            if (dz != 0.0) {
                face = dz > 0.0 ? 5 : 4;
                den = dz, na = dx, nb = dy;
                flip_u = true;
                flip_v = dz > 0.0;
            }
        }
        float u = 0.f, v = 0.f;
        if (den != 0.0) {
            const double hu = (na / den + 1.0f) * 0.5f, hv = (nb / den + 1.0f) * 0.5f;
            u = (float)(flip_u ? 1. - hu : hu);
            v = (float)(flip_v ? 1. - hv : hv);
        }
        return tex_sample(S, S.samplers[si].tex[face], false, u, v);
    }
    return mk(0.f, 0.f, 0.f);
}

// Scene::sample_environment (scene.cc:142-145) for the environments that
// read no texture: gradient (sampler_gradient.cc:11-17) or colour
// (sampler_col.cc:23-27); no environment is black. The direction is the
// path's FP32 ray, widened to FP64 for the sampler's double arithmetic.
__device__ __forceinline__ f3 env_constant(const XtEnv &env, f3 dir) {
    if (env.type == XT_ENV_GRADIENT) return gradient_eval(env.a, env.b, unit_y(dir.x, dir.y, dir.z));
    if (env.type == XT_ENV_COLOR) return mk(env.a[0], env.a[1], env.a[2]);
    return mk(0.f, 0.f, 0.f);
}

// Scene::sample_environment with cubemap and ERP environments as well.
__device__ __forceinline__ f3 env_sample(const SceneView &S, f3 dir) {
    const XtEnv &env = S.env;
    if (env.type == XT_ENV_CUBEMAP || env.type == XT_ENV_ERP)
        return sampler_eval(S, env.smp, (double)dir.x, (double)dir.y, (double)dir.z);
    return env_constant(env, dir);
}

#pragma clang fp contract(on)
