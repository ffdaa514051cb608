// Closest-hit traversal of the soup: planes first (linear, like the
// reference's object scan), then the kd-tree (short LDS stack with restart)
// or the octree (LDS stack, children in index order, pruned by the current
// closest t). The leaf test is a template parameter so that the FP32 pass and
// the FP64 re-test share one traversal.
#pragma once
#include "xt_isect.h"

#define XT_KD_STACK 16    // default kd-tree short stack entries per lane (8 B each, LDS)
#define XT_OCT_STACK 80   // octree stack entries per lane (4 B each, LDS); depth <= 10 needs <= 78
#define XT_MAX_STEPS 200000

struct SceneView {
    const float4 *prims;  // 3 float4 per soup primitive (XtPrim)
    const XtPrim64 *prims64;
    const XtPlane *planes;
    const XtKdNode *kd;
    const uint32_t *kd_refs;
    const float4 *kd_leaf;  // 3 float4 per leaf reference: the XtPrim with the soup index as object word
    const XtOctNode *oct;
    const uint32_t *oct_refs;
    const uint32_t *obj_material;
    const XtMaterial *materials;
    const float *vnormals;  // 9 per prim or null
    const double *tc64;     // 6 per prim: texcoords (u, v) of a triangle's corners
    const double *obj_uv;   // 2 per object: u_scale, v_scale
    const float4 *texels;   // all textures' texels
    const XtTexture *textures;
    const XtSampler *samplers;
    const uint32_t *mat_smp;  // 3 per material: sampler of diffuse / specular / emissive, or XT_NONE (colour)
    XtEnv env;
    float kb_lo[3], kb_hi[3];
    uint32_t n_prims, n_planes, n_oct, n_objects;
    uint32_t sampled;  // 1 when a material channel or the environment has a sampler
    uint32_t kd_indexed;  // 1: the FP32 walk reads leaf primitives from `prims` through `kd_refs`
};

struct RayF {
    f3 o, d;
    uint32_t origin;  // primitive the ray leaves from (XT_HIT_NONE: none)
    float self_t;     // > 0: the reference re-hits `origin` at this t
};

#define XT_NONE 0xffffffffu
#define XT_PLANE_BASE 0xffff0000u

__device__ __forceinline__ float safe_inv(float x) {
    return fabsf(x) > 1e-30f ? 1.f / x : copysignf(1e30f, x);
}

// Slab test against an FP32 box; [t0, t1] is narrowed. `grow` widens the
// interval by a relative amount (the boxes are already rounded outward).
__device__ __forceinline__ bool slab(const f3 o, const f3 inv, const float lo[3], const float hi[3], float &t0,
                                     float &t1, float grow) {
    float ax = (lo[0] - o.x) * inv.x, bx = (hi[0] - o.x) * inv.x;
    float ay = (lo[1] - o.y) * inv.y, by = (hi[1] - o.y) * inv.y;
    float az = (lo[2] - o.z) * inv.z, bz = (hi[2] - o.z) * inv.z;
    float tn = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fminf(az, bz));
    float tf = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fmaxf(az, bz));
    tn = tn - fabsf(tn) * grow;
    tf = tf + fabsf(tf) * grow;
    t0 = fmaxf(t0, tn);
    t1 = fminf(t1, tf);
    return t0 <= t1;
}

// Leaf tester interface: test(prim) and limit() (current closest t as float).
struct Tester32 {
    const SceneView &S;
    RayF r;
    float tsmall;
    Best best;
    Shear sh;
    __device__ __forceinline__ float limit() const { return best.t; }
    // one record (a, b, c) of primitive `prim`
    __device__ __forceinline__ void test_rec(uint32_t prim, const float4 a, const float4 b, const float4 c) {
        if (prim == r.origin) return;
        if (XT_PRIM_TYPE(__float_as_uint(a.w)) == XT_PRIM_TRIANGLE) {
            // a grazing triangle is decided by the FP64 re-check
            if (tri32(r.o, sh, tsmall, a, b, c, prim, best)) defer(best, 0.f, prim);
        } else {
            sph32(r.o, r.d, tsmall, a, b, prim, best);
        }
    }
    __device__ __forceinline__ void test(uint32_t prim) {
        const float4 a = S.prims[3 * prim], b = S.prims[3 * prim + 1], c = S.prims[3 * prim + 2];
        test_rec(prim, a, b, c);
    }
    // kd leaf references [k, k + n): each record's three loads go out
    // together, and the next record's while this one is tested (the empty
    // asm keeps the compiler from sinking loads into the branches that use
    // them, which would cost extra memory round trips)
    __device__ __forceinline__ void test_leaf_range(uint32_t k, uint32_t n) {
        if (n == 0) return;
        if (S.kd_indexed) {
            test_leaf_indexed(k, n);
            return;
        }
        float4 a = S.kd_leaf[3 * k], b = S.kd_leaf[3 * k + 1], c = S.kd_leaf[3 * k + 2];
        for (uint32_t j = 1; j <= n; ++j) {
            float4 na = a, nb = b, nc = c;
            if (j < n) {
                const uint32_t q = 3 * (k + j);
                na = S.kd_leaf[q];
                nb = S.kd_leaf[q + 1];
                nc = S.kd_leaf[q + 2];
            }
            asm volatile("" ::"v"(c.x), "v"(c.y), "v"(c.z), "v"(c.w));
            test_rec(__float_as_uint(b.w), a, b, c);
            a = na;
            b = nb;
            c = nc;
        }
    }
    // the same through the reference list: one soup record per reference,
    // the next reference's record loaded while this one is tested
    __device__ __forceinline__ void test_leaf_indexed(uint32_t k, uint32_t n) {
        uint32_t p = S.kd_refs[k];
        float4 a = S.prims[3 * p], b = S.prims[3 * p + 1], c = S.prims[3 * p + 2];
        for (uint32_t j = 1; j <= n; ++j) {
            float4 na = a, nb = b, nc = c;
            uint32_t np = p;
            if (j < n) {
                np = S.kd_refs[k + j];
                na = S.prims[3 * np];
                nb = S.prims[3 * np + 1];
                nc = S.prims[3 * np + 2];
            }
            asm volatile("" ::"v"(c.x), "v"(c.y), "v"(c.z), "v"(c.w));
            test_rec(p, a, b, c);
            a = na;
            b = nb;
            c = nc;
            p = np;
        }
    }
};

struct Tester64 {
    const SceneView &S;
    RayF r;
    d3 o, d;
    double bt;
    uint32_t prim, obj;
    double b0, b1, b2;
    f3 o32;    // the ray rounded to FP32, for the FP32 screen
    Shear sh;
    __device__ __forceinline__ float limit() const {
        return bt == INFINITY ? INFINITY : (float)bt * (1.f + 1e-4f) + 1e-30f;
    }
    __device__ __forceinline__ void offer64(double t, uint32_t p, uint32_t ob, const double *bc) {
        if (t < bt || (t == bt && (ob < obj || (ob == obj && p < prim)))) {
            bt = t;
            prim = p;
            obj = ob;
            b0 = bc ? bc[0] : 0.0;
            b1 = bc ? bc[1] : 0.0;
            b2 = bc ? bc[2] : 0.0;
        }
    }
    __device__ __forceinline__ void test_leaf_range(uint32_t k, uint32_t n) {
        for (uint32_t j = 0; j < n; ++j) test(S.kd_refs[k + j]);
    }
    __device__ __forceinline__ void test(uint32_t p) {
        if (p == r.origin) return;
        const float4 a = S.prims[3 * p], b = S.prims[3 * p + 1];
        const uint32_t tf = __float_as_uint(a.w);
        if (XT_PRIM_TYPE(tf) == XT_PRIM_TRIANGLE &&
            screen_out32(o32, sh, a, b, S.prims[3 * p + 2], limit()))
            return;
        const XtPrim64 q = S.prims64[p];
        if (XT_PRIM_TYPE(tf) == XT_PRIM_TRIANGLE) {
            double bc[3];
            double t = tri64(o, d, q, bc);
            if (t > 0.0) offer64(t, p, __float_as_uint(b.w), bc);
        } else {
            double t = sph64(o, d, q);
            if (t > 0.0) offer64(t, p, __float_as_uint(b.w), nullptr);
        }
    }
};

// kd-tree traversal with a short stack of (node, t_far) entries in LDS and a
// restart from the root when entries were dropped. `stk` points at this
// lane's first entry; entries are `stride` uint2 apart, STACK + 1 of them:
// the ring holds at most STACK entries, so the slot after the head is always
// free. Each descent step writes the far child there and keeps it only when
// both children are visited, which leaves the step without branches.
template <int STACK, class T>
__device__ __forceinline__ bool walk_kd(const SceneView &S, T &tst, const f3 o, const f3 d, uint2 *stk, int stride,
                                        float eps) {
    const f3 inv = mk(safe_inv(d.x), safe_inv(d.y), safe_inv(d.z));
    float t_root0 = 0.f, t_root1 = tst.limit();
    if (!slab(o, inv, S.kb_lo, S.kb_hi, t_root0, t_root1, eps)) return true;
    float tmin = t_root0, tmax = t_root1;
    constexpr int SLOTS = STACK + 1;
    int head = 0, count = 0;  // the ring starts empty at any slot; head stays in [0, SLOTS)
    bool dropped = false;
    const float oa[3] = {o.x, o.y, o.z}, da[3] = {d.x, d.y, d.z}, ia[3] = {inv.x, inv.y, inv.z};
    XtKdNode nd = S.kd[0];
    for (int steps = 0; steps < XT_MAX_STEPS; ++steps) {
        uint32_t axis = nd.word & 3u;
        while (axis != 3u) {
            const uint32_t c = nd.word >> 2;
            // both children in one aligned 16-byte load (pairs start on even
            // nodes), issued as soon as the node is known, so that its
            // latency overlaps the plane test below
            const uint4 pair = *(const uint4 *)(S.kd + c);
            const float split = nd.split;
            const float oo = axis == 0 ? oa[0] : (axis == 1 ? oa[1] : oa[2]);
            const float dd = axis == 0 ? da[0] : (axis == 1 ? da[1] : da[2]);
            const float ii = axis == 0 ? ia[0] : (axis == 1 ? ia[1] : ia[2]);
            const float tp = (split - oo) * ii;
            const bool below = oo < split || (oo == split && dd <= 0.f);
            const uint32_t first = below ? c : c + 1, second = below ? c + 1 : c;
            // tp <= 0: the plane is behind the origin, or the origin lies on
            // it and `below` already picked the side the ray moves into
            const bool only_first = tp > tmax * (1.f + eps) || tp <= 0.f;
            const bool only_second = !only_first && tp < tmin * (1.f - eps);
            const bool both = !only_first && !only_second;
            const int next = head + 1 == SLOTS ? 0 : head + 1;
            stk[next * stride] = make_uint2(second, __float_as_uint(tmax));
            // a push onto a full ring drops its oldest entry
            dropped = dropped || (both && count == STACK);
            count = both && count < STACK ? count + 1 : count;
            head = both ? next : head;
            tmax = both ? fminf(tmax, tp * (1.f + eps)) : tmax;
            const uint32_t node = only_second ? second : first;
            const bool lower = node == c;
            nd.word = lower ? pair.x : pair.z;
            nd.count = lower ? pair.y : pair.w;
            axis = nd.word & 3u;
        }
        const uint32_t first_ref = nd.word >> 2, n = nd.count;
        // the node a pop resumes with (with an empty stack, the root that a
        // restart resumes with), loaded while the leaf is tested
        const XtKdNode pre = S.kd[count > 0 ? stk[head * stride].x : 0u];
        tst.test_leaf_range(first_ref, n);
        const float lim = tst.limit();
        if (lim <= tmax) return true;
        if (count == 0) {
            if (!dropped || tmax >= t_root1) return true;
            // entries were lost: restart from the root beyond the processed range
            tmin = tmax;
            tmax = fminf(t_root1, lim);
            nd = pre;
            dropped = false;
            if (tmin > tmax) return true;
            continue;
        }
        const uint2 e = stk[head * stride];
        head = head == 0 ? SLOTS - 1 : head - 1;
        --count;
        nd = pre;
        tmin = tmax;
        tmax = fminf(__uint_as_float(e.y), lim);
        // the popped range starts beyond the closest hit, and so do all
        // deeper (farther) entries
        if (tmin > tmax * (1.f + eps)) return true;
    }
    return false;
}

// Octree traversal: all children of a visited node are considered in index
// order 7..0 on the stack (so 0 is visited first), pruned by their box entry
// against the current closest t.
template <class T>
__device__ __forceinline__ bool walk_oct(const SceneView &S, T &tst, const f3 o, const f3 d, uint32_t *stk, int stride,
                                         float eps) {
    if (S.n_oct == 0) return true;
    const f3 inv = mk(safe_inv(d.x), safe_inv(d.y), safe_inv(d.z));
    {
        float t0 = 0.f, t1 = tst.limit();
        if (!slab(o, inv, S.oct[0].lo, S.oct[0].hi, t0, t1, eps)) return true;
    }
    int sp = 0;
    stk[0] = 0;
    sp = 1;
    for (int steps = 0; steps < XT_MAX_STEPS; ++steps) {
        if (sp == 0) return true;
        const uint32_t n = stk[(--sp) * stride];
        const XtOctNode nd = S.oct[n];
        for (uint32_t k = 0; k < nd.n_items; ++k) tst.test(S.oct_refs[nd.first_item + k]);
        if (nd.first_child == 0) continue;
        for (int c = 7; c >= 0; --c) {
            const uint32_t ch = nd.first_child + (uint32_t)c;
            const XtOctNode cn = S.oct[ch];
            if (cn.subtree_items == 0) continue;
            float t0 = 0.f, t1 = tst.limit();
            if (!slab(o, inv, cn.lo, cn.hi, t0, t1, eps)) continue;
            if (sp >= XT_OCT_STACK) return false;
            stk[(sp++) * stride] = ch;
        }
    }
    return false;
}
