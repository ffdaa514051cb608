// TEST INFRASTRUCTURE ONLY: the oracle. Nothing in the xtracer_amd product
// links, imports or calls this file; tests, the smoke check and the CPU
// baseline leg of bench.py use it as the checker.
//
// FP64 CPU restatement of the reference's ray-query and integrator path,
// written from the reference sources (cited per function) and compiled with
// -ffp-contract=off so that its roundings are the reference's:
//   primitives          src/xtcore/math/{triangle,sphere,plane,aabb}.cc
//   octree build/walk   src/xtcore/octree.tml:97-249, mesh.cc:30-113
//   Scene::intersection src/xtcore/scene.cc:191-216
//   cameras             src/xtcore/camera/{perspective,erp,ods,cubemap}.cc
//   anti-aliasing       src/xtcore/aa.cc:7-77
//   integrators         src/xtcore/integrator/{stencil,depth,ao,pathtracer,
//                       normal,uv,emission,wireframe}
//   materials/sampling  src/xtcore/material/*.cc, lib/nmath/sample.inl
//   samplers            src/xtcore/sampler_*.cc, lib/nimg/sample.cc:8-50
//   colour clamps       lib/nimg/color.inl, lib/nimg/color.cc
// Every function is pinned against the compiled reference (oracle/_ref) by
// the tests: rays, hits and deterministic images bit for bit, estimators
// within their Monte Carlo bounds.
// Random numbers use the same counter-based streams as the device path
// (xtracer_amd/csrc/device/xt_rng.h restated below), because the reference's
// own glibc rand() stream is not reproducible.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include <omp.h>

namespace orc {

struct V {
    double x = 0, y = 0, z = 0;
};
static inline V mk(double x, double y, double z) { return V{x, y, z}; }
static inline V operator+(V a, V b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V operator-(V a, V b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V operator-(V a) { return mk(-a.x, -a.y, -a.z); }
static inline V operator*(V a, double s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline V operator*(double s, V a) { return mk(a.x * s, a.y * s, a.z * s); }
static inline V operator/(V a, double s) { return mk(a.x / s, a.y / s, a.z / s); }
static inline V operator+(V a, double s) { return mk(a.x + s, a.y + s, a.z + s); }
static inline V operator-(V a, double s) { return mk(a.x - s, a.y - s, a.z - s); }
static inline V mul(V a, V b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline double dot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline V cross(V a, V b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static inline double len(V a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
static inline V nrm(V a) {
    double l = len(a);
    return l != 0 ? mk(a.x / l, a.y / l, a.z / l) : a;
}
static inline double comp(V a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

// ---------------------------------------------------------------- RNG
static inline uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
static inline uint32_t path_key(uint32_t seed, uint32_t job, uint32_t pixel, uint32_t sample) {
    uint32_t h = mix32(seed + 0x9e3779b9u);
    h = mix32(h ^ job);
    h = mix32(h ^ pixel);
    return mix32(h ^ (sample * 0x85ebca6bu + 0x27d4eb2fu));
}
static inline uint32_t stream_of(uint32_t key, uint32_t q) { return mix32(key ^ (q * 0xc2b2ae35u + 0x165667b1u)); }
static inline uint32_t draw(uint32_t st, uint32_t i) { return mix32(st + i * 0x9e3779b9u + 0x7f4a7c15u); }
static inline double unit(uint32_t x) { return (double)x * (1.0 / 4294967295.0); }
struct Draws {
    uint32_t st, i;
    double next() { return unit(draw(st, i++)); }
};

// ---------------------------------------------------------------- scene
struct Tri {
    V v[3];
    V n[3];
    double tc[3][2] = {{0, 0}, {0, 0}, {0, 0}};  // Vector2f tc[3] (triangle.h:29)
};
struct Box {
    V lo, hi;
};
struct ONode {
    Box box;
    std::vector<uint32_t> items;
    int child[8];
};
struct Mesh {
    std::vector<Tri> tris;
    std::vector<Box> boxes;
    std::vector<ONode> nodes;
};
struct Obj {
    uint32_t type;  // 0 plane, 1 sphere, 2 triangle, 3 mesh, other: absent
    uint32_t mat;
    V pn;
    double poff;
    V sc;
    double sr;
    Tri tri;
    int mesh;
    double us = 0, vs = 0;  // uv_scale (surface.h:26)
};
struct Mat {
    uint32_t type;  // 0 lambert, 1 phong, 2 blinn-phong, 3 emissive
    float kd[3], ks[3], ke[3];
    float refl, expo;
    uint32_t smp[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu};  // diffuse, specular, emissive samplers
};
// A texture Pixmap (RGBA float, row 0 at the top) and a sampler
// (sampler_{col,tex,gradient,cubemap,erp}.h).
struct Tex {
    size_t w = 0, h = 0;
    const float *px = nullptr;
};
struct Smp {
    uint32_t kind;  // 1 colour, 2 texture, 3 gradient, 4 cubemap, 5 erp
    uint32_t bilinear;
    int32_t tex[6];  // cubemap faces: left, right, top, bottom, front, back
    float a[3], b[3];
};
struct Scene {
    std::vector<Tex> texs;
    std::vector<Smp> smps;
    uint32_t env_smp = 0;
    std::vector<Obj> objs;
    std::vector<Mesh> meshes;
    std::vector<Mat> mats;
    uint32_t env_type;
    float env_a[3], env_b[3];
};

// AABB3::intersection (aabb.cc:32-71): float slabs on double data, early
// accept when the origin is strictly inside, no entry t.
static bool box_hit(const Box &b, V o, V d) {
    if (o.x > b.lo.x && o.y > b.lo.y && o.z > b.lo.z && o.x < b.hi.x && o.y < b.hi.y && o.z < b.hi.z) return true;
    const V bb[2] = {b.lo, b.hi};
    static const float t0 = 0.0;
    int xs = (int)(d.x < 0.0), ys = (int)(d.y < 0.0);
    float ix = 1.0 / d.x, iy = 1.0 / d.y;
    float tmin = (bb[xs].x - o.x) * ix;
    float tmax = (bb[1 - xs].x - o.x) * ix;
    float tymin = (bb[ys].y - o.y) * iy;
    float tymax = (bb[1 - ys].y - o.y) * iy;
    if ((tmin > tymax) || (tymin > tmax)) return false;
    if (tymin > tmin) tmin = tymin;
    if (tymax < tmax) tmax = tymax;
    int zs = (int)(d.z < 0.0);
    float iz = 1.0 / d.z;
    float tzmin = (bb[zs].z - o.z) * iz;
    float tzmax = (bb[1 - zs].z - o.z) * iz;
    if ((tmin > tzmax) || (tzmin > tmax)) return false;
    if (tzmin > tmin) tmin = tzmin;
    if (tzmax < tmax) tmax = tzmax;
    return tmax > t0;
}

struct Hit {
    double t = INFINITY;
    V point, normal;
    double tu = 0, tv = 0;  // texcoord (hitrecord.h:33; z stays 0)
    int obj = -1;
    long prim = -1;  // triangle index inside its mesh (OBJ face order)
};

// Triangle::intersection / calc_barycentric (triangle.cc:82-184).
static bool tri_hit(const Tri &tr, V o, V d, Hit *h) {
    V normal = nrm(cross(tr.v[1] - tr.v[0], tr.v[2] - tr.v[0]));
    double ndd = dot(normal, d);
    if (std::fabs(ndd) < 1e-8) return false;
    V vo = o - tr.v[0];
    double t = -dot(normal, vo) / ndd;
    if (t < 1e-8) return false;
    V pos = o + d * t;
    V v1 = tr.v[1] - tr.v[0], v2 = tr.v[2] - tr.v[0];
    V x = cross(v1, v2);
    V n2 = nrm(x);
    double area = std::fabs(dot(x, n2)) * 0.5;
    double b0 = 0, b1 = 0, b2 = 0;
    if (!(area < 1e-8)) {
        V p0 = tr.v[0] - pos, p1 = tr.v[1] - pos, p2 = tr.v[2] - pos;
        b0 = std::fabs(dot(cross(p1, p2), n2)) * 0.5 / area;
        b1 = std::fabs(dot(cross(p2, p0), n2)) * 0.5 / area;
        b2 = std::fabs(dot(cross(p0, p1), n2)) * 0.5 / area;
    }
    double s = b0 + b1 + b2;
    if (s < 1.0 - 1e-8 || s > 1.0 + 1e-8) return false;
    if (h) {
        h->t = t;
        h->point = pos;
        // triangle.cc:113: tc[0] * bc.x + tc[1] * bc.y + tc[2] * bc.z
        h->tu = tr.tc[0][0] * b0 + tr.tc[1][0] * b1 + tr.tc[2][0] * b2;
        h->tv = tr.tc[0][1] * b0 + tr.tc[1][1] * b1 + tr.tc[2][1] * b2;
        V pn = tr.n[0] * b0 + tr.n[1] * b1 + tr.n[2] * b2;
        h->normal = len(pn) ? pn : normal;
    }
    return true;
}

// Sphere::intersection (sphere.cc:35-61).
static bool sph_hit(V c, double r, V o, V d, Hit *h, double us = 0, double vs = 0) {
    double b = 2 * dot(o - c, d);
    double cq = dot(c, c) + dot(o, o) + 2 * dot(-c, o) - r * r;
    double disc = (b * b - 4 * cq);
    if (disc > 0.0) {
        double s = std::sqrt(disc);
        double t1 = (-b + s) / 2.0, t2 = (-b - s) / 2.0;
        double t = t1 < t2 ? t1 : t2;
        if (t > 1e-8 && h) {
            h->t = t;
            h->point = o + d * t;
            h->normal = (h->point - c) / r * (t1 * t2 > 0. ? 1. : -1);
            // sphere.cc:53-54
            h->tu = std::asin(h->normal.x / (us != 0.0f ? us : 1.0f)) / 3.14159265358979323846 + 0.5;
            h->tv = std::asin(h->normal.y / (vs != 0.0f ? vs : 1.0f)) / 3.14159265358979323846 + 0.5;
            return true;
        }
    }
    return false;
}

// Plane::intersection (plane.cc:19-61).
static bool plane_hit(V n, double off, V o, V d, Hit *h, double us = 0, double vs = 0) {
    double ndd = dot(n, d);
    if (std::fabs(ndd) < 1e-8) return false;
    V v = mk(std::fabs(n.x), std::fabs(n.y), std::fabs(n.z)) * off;
    V vo = v - o;
    double t = dot(vo, n) / ndd;
    if (t < 1e-8) return false;
    if (h) {
        h->t = t;
        h->point = o + d * t;
        h->normal = n;
        // plane.cc:43-57: the point projected on two tangents, scaled, wrapped
        V nn = nrm(n);
        V uvec = mk(nn.y, nn.z, -nn.x);
        V vvec = cross(uvec, nn);
        double tu = dot(uvec, (v + h->point)) * us;
        double tv = dot(vvec, (v + h->point)) * vs;
        if (tu > 1.f) tu -= (float)(int)tu;
        if (tv > 1.f) tv -= (float)(int)tv;
        if (tu < -1.f) tu -= (float)(int)tu;
        if (tv < -1.f) tv -= (float)(int)tv;
        if (tu < 0.f) tu = 1.f + tu;
        if (tv < 0.f) tv = 1.f + tv;
        h->tu = tu;
        h->tv = tv;
    }
    return true;
}

// Octree<T>::build / subdivide (octree.tml:119-198).
static Box mkbox(V a, V b) {
    Box r;
    r.lo = mk(a.x <= b.x ? a.x : b.x, a.y <= b.y ? a.y : b.y, a.z <= b.z ? a.z : b.z);
    r.hi = mk(a.x >= b.x ? a.x : b.x, a.y >= b.y ? a.y : b.y, a.z >= b.z ? a.z : b.z);
    return r;
}
static bool contains(const Box &b, V p) {
    return (p.x >= b.lo.x) && (p.y >= b.lo.y) && (p.z >= b.lo.z) && (p.x <= b.hi.x) && (p.y <= b.hi.y) &&
           (p.z <= b.hi.z);
}
static void subdivide(Mesh &m, int ni, size_t level, double root_zmax) {
    if (m.nodes[ni].items.size() <= 10 || level >= 10) return;
    Box nb = m.nodes[ni].box;
    V c = (nb.lo + nb.hi) / 2.0;
    Box cb[8] = {mkbox(nb.lo, c),
                 mkbox(mk(c.x, nb.lo.y, nb.lo.z), mk(nb.hi.x, c.y, c.z)),
                 mkbox(mk(c.x, c.y, nb.lo.z), mk(nb.hi.x, nb.hi.y, c.z)),
                 mkbox(mk(nb.lo.x, c.y, nb.lo.z), mk(c.x, nb.hi.y, c.z)),
                 mkbox(mk(nb.lo.x, nb.lo.y, c.z), mk(c.x, c.y, nb.hi.z)),
                 mkbox(mk(c.x, nb.lo.y, c.z), mk(nb.hi.x, c.y, nb.hi.z)),
                 mkbox(c, nb.hi),
                 mkbox(mk(nb.lo.x, c.y, c.z), mk(c.x, nb.hi.y, root_zmax))};
    for (int i = 0; i < 8; ++i) {
        ONode ch;
        ch.box = cb[i];
        for (int k = 0; k < 8; ++k) ch.child[k] = -1;
        std::vector<uint32_t> keep;
        for (uint32_t it : m.nodes[ni].items) {
            const Tri &t = m.tris[it];
            if (contains(ch.box, t.v[0]) && contains(ch.box, t.v[1]) && contains(ch.box, t.v[2])) ch.items.push_back(it);
            else keep.push_back(it);
        }
        m.nodes[ni].items.swap(keep);
        int id = (int)m.nodes.size();
        m.nodes.push_back(ch);
        m.nodes[ni].child[i] = id;
    }
    for (int i = 0; i < 8; ++i) subdivide(m, m.nodes[ni].child[i], level + 1, root_zmax);
}
static void build_mesh(Mesh &m) {
    Box root;  // (0,0,0) box augmented by every triangle box (Octree::clear/add)
    for (const Tri &t : m.tris) {
        Box b;
        b.lo = mk(INFINITY, INFINITY, INFINITY);
        b.hi = mk(-INFINITY, -INFINITY, -INFINITY);
        for (int k = 0; k < 3; ++k) {
            V p = t.v[k];
            if (p.x < b.lo.x) b.lo.x = p.x;
            if (p.y < b.lo.y) b.lo.y = p.y;
            if (p.z < b.lo.z) b.lo.z = p.z;
            if (p.x > b.hi.x) b.hi.x = p.x;
            if (p.y > b.hi.y) b.hi.y = p.y;
            if (p.z > b.hi.z) b.hi.z = p.z;
        }
        if (b.hi.x > root.hi.x) root.hi.x = b.hi.x;
        if (b.lo.x < root.lo.x) root.lo.x = b.lo.x;
        if (b.hi.y > root.hi.y) root.hi.y = b.hi.y;
        if (b.lo.y < root.lo.y) root.lo.y = b.lo.y;
        if (b.hi.z > root.hi.z) root.hi.z = b.hi.z;
        if (b.lo.z < root.lo.z) root.lo.z = b.lo.z;
    }
    m.nodes.clear();
    ONode r;
    r.box = root;
    for (int k = 0; k < 8; ++k) r.child[k] = -1;
    for (size_t i = 0; i < m.tris.size(); ++i) r.items.push_back((uint32_t)i);
    m.nodes.push_back(r);
    if (m.nodes[0].items.size() > 10) subdivide(m, 0, 0, root.hi.z);
}

// Octree::r_intersection (octree.tml:206-249).
static long oct_walk(const Mesh &m, int ni, V o, V d, Hit *out) {
    if (ni < 0 || !box_hit(m.nodes[ni].box, o, d)) return -1;
    Hit best;
    best.t = FLT_MAX;
    long closest = -1;
    for (uint32_t it : m.nodes[ni].items) {
        Hit h;
        if (tri_hit(m.tris[it], o, d, &h) && h.t < best.t) {
            best = h;
            closest = it;
        }
    }
    for (int i = 0; i < 8; ++i) {
        Hit h;
        long c = oct_walk(m, m.nodes[ni].child[i], o, d, &h);
        if (c >= 0 && h.t < best.t) {
            best = h;
            closest = c;
        }
    }
    if (closest >= 0 && out) *out = best;
    return closest;
}

// Scene::intersection (scene.cc:191-216).
static bool scene_hit(const Scene &s, V o, V d, Hit &res) {
    res = Hit();
    bool any = false;
    for (size_t i = 0; i < s.objs.size(); ++i) {
        const Obj &ob = s.objs[i];
        Hit h;
        bool ok = false;
        long prim = -1;
        switch (ob.type) {
            case 0: ok = plane_hit(ob.pn, ob.poff, o, d, &h, ob.us, ob.vs); break;
            case 1: ok = sph_hit(ob.sc, ob.sr, o, d, &h, ob.us, ob.vs); break;
            case 2: ok = tri_hit(ob.tri, o, d, &h); break;
            case 3:
                prim = oct_walk(s.meshes[ob.mesh], 0, o, d, &h);
                ok = prim >= 0;
                break;
            default: break;
        }
        if (ok) {
            any = true;
            if (res.t > h.t) {
                res = h;
                res.obj = (int)i;
                res.prim = ob.type == 3 ? prim : 0;
            }
        }
    }
    return any;
}

// ---------------------------------------------------------------- sampling
// lib/nmath/sample.inl:28-59
static V hemisphere(V normal, V direction, Draws &r) {
    double u = r.next(), v = r.next();
    float phi = std::acos(std::sqrt(u));
    float theta = 2.0 * M_PI * v;
    V dd = mk(std::cos(theta) * std::sin(phi), std::cos(phi), std::sin(theta) * std::sin(phi));
    V n = nrm(normal);
    V xv = nrm(cross(n + 1e-8, n));
    V yv = n;
    V zv = nrm(cross(xv, yv));
    // columns x, y, z; Vector3f::transformed(Matrix3x3f)
    V w = mk(xv.x * dd.x + yv.x * dd.y + zv.x * dd.z, xv.y * dd.x + yv.y * dd.y + zv.y * dd.z,
             xv.z * dd.x + yv.z * dd.y + zv.z * dd.z);
    V dirc = nrm(direction);
    float ndotl = dot(n, dirc);
    return nrm(w * (ndotl > 0 ? ndotl : 0));
}
// lib/nmath/sample.inl:61-90
static V lobe(V normal, V direction, double e, Draws &r) {
    V c0 = mk(1, 0, 0), c1 = mk(0, 1, 0), c2 = mk(0, 0, 1);
    V i = nrm(direction), n = nrm(normal);
    V refl = (2 * dot(i, n) * n) - i;
    if (1.0f - dot(direction, normal) > 1e-8) {
        V iv = nrm(cross(direction, refl));
        V kv = cross(refl, iv);
        c0 = iv;
        c1 = refl;
        c2 = kv;
    }
    double u = r.next(), v = r.next();
    double theta = 2.0 * M_PI * u;
    double phi = std::acos(std::pow(v, 1.0 / (e + 1)));
    V l = mk(std::cos(theta) * std::sin(phi), std::cos(phi), std::sin(theta) * std::sin(phi));
    V vc = mk(c0.x * l.x + c1.x * l.y + c2.x * l.z, c0.y * l.x + c1.y * l.y + c2.y * l.z,
              c0.z * l.x + c1.z * l.y + c2.z * l.z);
    double vdotr = dot(refl, vc);
    if (vdotr < 0.0) vdotr = 0.0;
    return nrm(vc * std::pow(vdotr, e));
}
// lib/nmath/sample.inl:92-107
static V diffuse(V normal, Draws &r) {
    V p;
    for (int k = 0; k < 64; ++k) {
        double a = r.next(), b = r.next(), c = r.next();
        p = 2.0 * mk(a, b, c) - mk(1, 1, 1);
        float l = len(mul(p, p));
        l *= l;
        if (!(l >= 1.0)) break;
    }
    p = p + normal;
    return nrm(p);
}

// ColorRGBf with the setters' clamp at 0 (lib/nimg/color.inl).
struct C3 {
    float r = 0, g = 0, b = 0;
};
static inline float cl(float x) { return x < 0.f ? 0.f : x; }
static inline C3 cmul(C3 a, C3 b) { return C3{cl(a.r * b.r), cl(a.g * b.g), cl(a.b * b.b)}; }
static inline C3 cscale(C3 a, float s) { return C3{cl(a.r * s), cl(a.g * s), cl(a.b * s)}; }
static inline C3 cadd(C3 a, C3 b) { return C3{cl(a.r + b.r), cl(a.g + b.g), cl(a.b + b.b)}; }

// Pixmap::pixel_ro (pixmap.cc:44-50): coordinates past the edge clamp.
static C3 pixel_ro(const Tex &t, size_t x, size_t y) {
    size_t nx = x >= t.w ? t.w - 1 : x;
    size_t ny = y >= t.h ? t.h - 1 : y;
    const float *p = t.px + 4 * (ny * t.w + nx);
    return C3{p[0], p[1], p[2]};
}
// size_t(x) as x86-64 converts a double: NaN and out-of-range give 2^63
static size_t to_size(double x) { return x >= 0.0 && x < 18446744073709551616.0 ? (size_t)x : (size_t)1 << 63; }
// nimg::sample::nearest (sample.cc:8-17)
static C3 nearest(const Tex &t, double u, double v) {
    u = u < 0 ? 0 : (u > 1 ? u - std::floor(u) : u);
    v = v < 0 ? 0 : (v > 1 ? v - std::floor(v) : v);
    size_t x = to_size(u * (float)t.w);
    size_t y = to_size(v * (float)t.h);
    return pixel_ro(t, x, y);
}
// nimg::sample::bilinear (sample.cc:19-50), over ColorRGBAf's * and +
static C3 bilinear(const Tex &t, double u, double v) {
    size_t w = t.w, h = t.h;
    float fu = u * w;
    float fv = v * h;
    int u1 = ((int)fu) % w;
    int v1 = ((int)fv) % h;
    int u2 = (u1 + 1) % w;
    int v2 = (v1 + 1) % h;
    float fracu = fu - floorf(fu);
    float fracv = fv - floorf(fv);
    float w1 = (1 - fracu) * (1 - fracv);
    float w2 = fracu * (1 - fracv);
    float w3 = (1 - fracu) * fracv;
    float w4 = fracu * fracv;
    C3 c1 = pixel_ro(t, u1, v1), c2 = pixel_ro(t, u2, v1), c3 = pixel_ro(t, u1, v2), c4 = pixel_ro(t, u2, v2);
    return cadd(cadd(cadd(cscale(c1, w1), cscale(c2, w2)), cscale(c3, w3)), cscale(c4, w4));
}
// ISampler::sample(tc) (sampler_col.cc:23-27, sampler_tex.cc:40-48,
// sampler_gradient.cc:11-17, sampler_erp.cc:13-25, sampler_cubemap.cc:14-41)
static C3 sample(const Scene &s, const Smp &sm, V tc) {
    if (sm.kind == 1) return C3{sm.a[0], sm.a[1], sm.a[2]};
    if (sm.kind == 2) {
        const Tex &t = s.texs[sm.tex[0]];
        return sm.bilinear ? bilinear(t, tc.x, tc.y) : nearest(t, tc.x, tc.y);
    }
    V dir = nrm(tc);
    if (sm.kind == 3) {
        float t = 0.5f * dir.y + 1.0;
        C3 a{sm.a[0], sm.a[1], sm.a[2]}, b{sm.b[0], sm.b[1], sm.b[2]};
        return cadd(cscale(a, (float)(1.0 - t)), cscale(b, t));
    }
    if (sm.kind == 5) {
        float theta = std::acos(dir.y);
        float phi = std::atan2(dir.x, -dir.z);
        float u = 0.5f + phi / 6.283185307179586232;
        float v = theta / 3.14159265358979323846;
        return nearest(s.texs[sm.tex[0]], u, v);
    }
    if (sm.kind == 4) {
        int face = 2;
        float u = 0.f, v = 0.f;
        if ((std::fabs(dir.x) >= std::fabs(dir.y)) && (std::fabs(dir.x) >= std::fabs(dir.z))) {
            if (dir.x > 0.0f) { face = 1; u = (dir.z / dir.x + 1.0f) * 0.5f; v = 1.f - (dir.y / dir.x + 1.0f) * 0.5f; }
            else if (dir.x < 0.0f) { face = 0; u = (dir.z / dir.x + 1.0f) * 0.5f; v = (dir.y / dir.x + 1.0f) * 0.5f; }
        } else if ((std::fabs(dir.y) >= std::fabs(dir.x)) && (std::fabs(dir.y) >= std::fabs(dir.z))) {
            if (dir.y > 0.0f) { face = 2; u = (dir.x / dir.y + 1.0f) * 0.5f; v = 1.f - (dir.z / dir.y + 1.0f) * 0.5f; }
            else if (dir.y < 0.0f) { face = 3; u = 1. - (dir.x / dir.y + 1.0f) * 0.5f; v = 1.f - (dir.z / dir.y + 1.0f) * 0.5f; }
        } else if ((std::fabs(dir.z) >= std::fabs(dir.x)) && (std::fabs(dir.z) >= std::fabs(dir.y))) {
            if (dir.z > 0.0f) { face = 5; u = 1. - (dir.x / dir.z + 1.0f) * 0.5f; v = 1.f - (dir.y / dir.z + 1.0f) * 0.5f; }
            else if (dir.z < 0.0f) { face = 4; u = 1. - (dir.x / dir.z + 1.0f) * 0.5f; v = (dir.y / dir.z + 1.0f) * 0.5f; }
        }
        return nearest(s.texs[sm.tex[face]], u, v);
    }
    return C3{};
}
// IMaterial::get_sample (material.cc:65-70): the channel's sampler at the
// hit's texcoord, or its constant colour (0 when absent)
static C3 channel(const Scene &s, const Mat &m, int c, const float *konst, double tu, double tv) {
    if (m.smp[c] != 0xffffffffu) return sample(s, s.smps[m.smp[c]], mk(tu, tv, 0.0));
    return C3{konst[0], konst[1], konst[2]};
}

static C3 environment(const Scene &s, V dir) {
    if (s.env_type == 3 || s.env_type == 4) return sample(s, s.smps[s.env_smp], dir);  // cubemap, erp
    if (s.env_type == 1) {  // sampler_gradient.cc:11-17
        V dn = nrm(dir);
        float t = 0.5f * dn.y + 1.0;
        C3 a{s.env_a[0], s.env_a[1], s.env_a[2]}, b{s.env_b[0], s.env_b[1], s.env_b[2]};
        return cadd(cscale(a, (float)(1.0 - t)), cscale(b, t));
    }
    if (s.env_type == 2) return C3{s.env_a[0], s.env_a[1], s.env_a[2]};
    return C3{};
}

// ---------------------------------------------------------------- camera
struct Cam {
    uint32_t kind = 0;  // 0 thin-lens, 1 ods, 2 erp, 3 cubemap (parseutil.cc:262-275)
    V pos, target, up;
    float fov, aperture, flength;
    V rx, ry, rz;
    V orientation;      // ods / erp
    float ipd = 0.f;    // ods
    double rot[3][3];   // Matrix4x4f m; m.rotate(orientation)
};
typedef double M3[3][3];
static void mat3_mul(const M3 &a, const M3 &b, M3 &r) {  // matrix.cc:57-66
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r[i][j] = a[i][0] * b[0][j] + a[i][1] * b[1][j] + a[i][2] * b[2][j];
}
static void cam_setup(Cam &c) {  // perspective.cc:26-36
    c.rz = nrm(c.target - c.pos);
    c.rx = nrm(cross(c.up, c.rz));
    c.ry = nrm(cross(c.rx, c.rz));
    // Matrix4x4f::rotate(euler), matrix.cc:485-492: identity *= Rx*Ry*Rz
    // (the 4x4 product of matrix.cc:414-427, its zero terms included)
    const V &e = c.orientation;
    const M3 xr = {{1, 0, 0}, {0, std::cos(e.x), -std::sin(e.x)}, {0, std::sin(e.x), std::cos(e.x)}};
    const M3 yr = {{std::cos(e.y), 0, std::sin(e.y)}, {0, 1, 0}, {-std::sin(e.y), 0, std::cos(e.y)}};
    const M3 zr = {{std::cos(e.z), -std::sin(e.z), 0}, {std::sin(e.z), std::cos(e.z), 0}, {0, 0, 1}};
    M3 xy, r;
    mat3_mul(xr, yr, xy);
    mat3_mul(xy, zr, r);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            c.rot[i][j] = (i == 0 ? 1.0 : 0.0) * r[0][j] + (i == 1 ? 1.0 : 0.0) * r[1][j] +
                          (i == 2 ? 1.0 : 0.0) * r[2][j] + 0.0 * 0.0;
}
static V rotated(const double m[3][3], V v) {  // Vector3f::transformed, vector.inl:517-523
    return mk(m[0][0] * v.x + m[0][1] * v.y + m[0][2] * v.z + 0.0, m[1][0] * v.x + m[1][1] * v.y + m[1][2] * v.z + 0.0,
              m[2][0] * v.x + m[2][1] * v.y + m[2][2] * v.z + 0.0);
}
static const double PI_D = 3.14159265358979323846;  // nmath::PI, precision.h:101

// erp.cc:11-25
static void primary_erp(const Cam &c, float x, float y, float width, float height, V &o, V &d) {
    float phi = (x / width) * PI_D * 2.f;
    float theta = (y / height) * PI_D;
    o = c.pos;
    d = mk(std::sin((double)theta) * std::cos((double)phi), std::cos((double)theta),
           std::sin((double)theta) * std::sin((double)phi));
    d = rotated(c.rot, d);
}

// ods.cc:15-33
static void primary_ods(const Cam &c, float x, float y, float width, float height, V &o, V &d) {
    float h = height * .5f;
    float scale = c.ipd * .5f * ((y < h) ? -1 : 1);
    if (y > h) y = y - h;
    float theta = (x / width) * 2.f * PI_D;
    float phi = PI_D * .5f - (y / h) * PI_D;
    V u = mk(std::cos((double)theta), 0, std::sin((double)theta));
    o = mk(u.x * (double)scale, u.y * (double)scale, u.z * (double)scale) + c.pos;
    d = mk(std::sin((double)theta) * std::cos((double)phi), std::sin((double)phi),
           -std::cos((double)theta) * std::cos((double)phi));
    d = rotated(c.rot, d);
}

// cubemap.cc:17-60
static void primary_cubemap(const Cam &c, float x, float y, float width, float height, V &o, V &d) {
    size_t face_size = (size_t)(height) / 6;
    size_t face = (size_t)(y) / face_size;
    V up, camdir;
    switch (face) {
        case 0: camdir = mk(1, 0, 0); up = mk(0, 1, 0); break;
        case 1: camdir = mk(-1, 0, 0); up = mk(0, 1, 0); break;
        case 2: camdir = mk(0, 1, 0); up = mk(0, 0, -1); break;
        case 3: camdir = mk(0, -1, 0); up = mk(0, 0, 1); break;
        case 4: camdir = mk(0, 0, 1); up = mk(0, 1, 0); break;
        default: camdir = mk(0, 0, -1); up = mk(0, 1, 0); break;
    }
    y -= (float)(face * face_size);
    o = c.pos;
    V v = mk((2.f * (double)x / (double)width) - 1.f, (2.f * (double)y / (double)face_size) - 1.f, 1.f);
    v = nrm(v);
    V rz = camdir, rx = nrm(cross(up, rz)), ry = nrm(cross(rx, rz));
    const double m[3][3] = {{rx.x, ry.x, rz.x}, {rx.y, ry.y, rz.y}, {rx.z, ry.z, rz.z}};
    d = nrm(rotated(m, v));
}

// The thin lens's two jitters, prng_c(-aperture/2, aperture/2) each
// (perspective.cc:76-78, prng.inl:82: uniform over the closed range), drawn
// from query stream 0 of the sample: the first moves x, the second y.
static void lens_jitter(const Cam &c, uint32_t key, double &jx, double &jy) {
    const double half = c.aperture / 2.f;
    const uint32_t st = stream_of(key, 0);
    jx = unit(draw(st, 0)) * ((half - -half)) + -half;
    jy = unit(draw(st, 1)) * ((half - -half)) + -half;
}

// perspective.cc:38-98 with the lens jitters given (used only when
// flength > 0).
static void primary_persp_at(const Cam &c, float x, float y, float w, float h, double jx, double jy, V &o, V &d) {
    double aspect = (double)w / (double)h;
    o = c.pos;
    V dir = mk((2.0 * (double)x / (double)w) - 1.0, ((2.0 * (double)y / (double)h) - 1.0) / aspect,
               1.0 / std::tan(c.fov * 0.017453292519943 / 2.0));
    V org = o;
    if (c.flength > 0) {
        // the lens sample lies on the image plane (z = dir.z), the ray aims at
        // the focal point flength further along the unjittered direction
        V fo = dir;
        fo.x += jx;
        fo.y += jy;
        V fp = dir + (double)c.flength * nrm(dir);
        dir = fp - fo;
        org = fo;
    }
    auto xf = [&](V v) {
        return mk(c.rx.x * v.x + c.ry.x * v.y + c.rz.x * v.z + 0.0, c.rx.y * v.x + c.ry.y * v.y + c.rz.y * v.z + 0.0,
                  c.rx.z * v.x + c.ry.z * v.y + c.rz.z * v.z + 0.0);
    };
    d = nrm(xf(dir));
    if (c.flength > 0) o = xf(org) + c.pos;
}

static void primary_persp(const Cam &c, float x, float y, float w, float h, uint32_t key, V &o, V &d) {
    double jx = 0.0, jy = 0.0;
    if (c.flength > 0) lens_jitter(c, key, jx, jy);
    primary_persp_at(c, x, y, w, h, jx, jy, o, d);
}

// antialiasing::gen_samples_random (aa.cc:34-48): sub-sample `a` of pixel
// (x, y) sits at pixel + (float(prng_c(0, 1)), float(prng_c(0, 1))), the
// Vector2f sum in double, handed to get_primary_ray as floats. The two draws
// come from stream 0 of a key of their own per (pixel, a).
static void aa_random_draws(uint32_t seed, uint32_t job_key, uint32_t pix, uint32_t a, double &ux, double &uy) {
    const uint32_t k2 = path_key(seed, job_key, pix, 0xfff00000u + a);
    const uint32_t st = stream_of(k2, 0);
    ux = unit(draw(st, 0));
    uy = unit(draw(st, 1));
}

static void aa_random_at(uint32_t x, uint32_t y, double ux, double uy, float &cx, float &cy) {
    const float rx = (float)ux, ry = (float)uy;
    cx = (float)((double)x + (double)rx);
    cy = (float)((double)y + (double)ry);
}

// ICamera::get_primary_ray of the camera's kind.
static void primary(const Cam &c, float x, float y, float w, float h, uint32_t key, V &o, V &d) {
    switch (c.kind) {
        case 1: primary_ods(c, x, y, w, h, o, d); break;
        case 2: primary_erp(c, x, y, w, h, o, d); break;
        case 3: primary_cubemap(c, x, y, w, h, o, d); break;
        default: primary_persp(c, x, y, w, h, key, o, d); break;
    }
}

struct Job {
    uint32_t integrator, width, height, samples, aa, aa_mode, rdepth, seed, job_key;
};

struct Ctx {
    const Scene *s;
    const Cam *cam;
    Job job;
    uint64_t *queries;
};

// pathtracer::Integrator::eval (pathtracer/integrator.cc:21-47); q is the
// 1-based index of the query this call makes.
static C3 eval(const Ctx &cx, size_t depth, C3 in_i, V o, V d, uint32_t key, uint32_t q, uint64_t &nq) {
    if (depth == 0) return C3{};
    Hit h;
    ++nq;
    bool hit = scene_hit(*cx.s, o, d, h);
    if (!hit) return environment(*cx.s, d);
    const Obj &ob = cx.s->objs[h.obj];
    const Mat &m = cx.s->mats[ob.mat];
    Draws r{stream_of(key, q), 0};
    C3 f;
    V no, nd;
    bool cont = true;
    if (m.type == 3) {
        f = channel(*cx.s, m, 2, m.ke, h.tu, h.tv);
        cont = false;
    } else if (m.type == 4) {
        // dielectric.cc:24-32 sets only the ray: the intensity keeps
        // ColorRGBf's default (0, 0, 0), so value = in.I * 0 and the product
        // with the refracted ray's (finite) value stays 0; the recursion's
        // queries change nothing and are not made here.
        f = C3{0.f, 0.f, 0.f};
        cont = false;
    } else if (m.type == 0) {  // lambert.cc:30-40
        no = h.point + h.normal * 1e-8;
        nd = nrm(diffuse(h.normal, r));
        f = cscale(channel(*cx.s, m, 0, m.kd, h.tu, h.tv), (float)dot(h.normal, nd));
    } else {  // phong.cc:35-55, blinnphong.cc:34-61
        no = h.point + h.normal * 1e-8;
        double s = m.refl;
        double k = r.next();
        if (k > s) {
            f = channel(*cx.s, m, 0, m.kd, h.tu, h.tv);
            nd = diffuse(h.normal, r);
        } else {
            f = channel(*cx.s, m, 1, m.ks, h.tu, h.tv);
            nd = lobe(h.normal, -d, m.expo, r);
            if (m.type == 2) nd = nrm(nd);
        }
    }
    C3 value = cmul(in_i, f);
    if (cont) value = cmul(value, eval(cx, depth - 1, f, no, nd, key, q + 1, nq));
    return value;
}

}  // namespace orc

using namespace orc;

extern "C" {

// Scene description handed in by the tests (objects in scan order).
struct or_scene_desc {
    uint32_t n_objects;
    const uint32_t *obj_type;
    const uint32_t *obj_material;
    const double *obj_param;     // 9 doubles per object: plane n, off | sphere c, r | triangle v0..v2
    const uint32_t *mesh_first;  // per object: first triangle in mesh arrays (meshes only)
    const uint32_t *mesh_count;
    const double *mesh_v;        // 9 doubles per triangle, OBJ face order
    const float *mesh_n;         // 9 floats per triangle or NULL
    const double *obj_tc;        // 6 doubles per object (triangles: u, v per corner) or NULL
    const double *mesh_tc;       // 6 doubles per mesh triangle or NULL
    const double *obj_uv_scale;  // 2 doubles per object or NULL
    uint32_t n_materials;
    const uint32_t *mat_type;
    const float *mat_param;      // 11 floats: kd[3], ks[3], ke[3], reflectance, exponent
    uint32_t env_type;
    const float *env;            // a[3], b[3]
    uint32_t n_textures;
    const uint32_t *tex_info;    // 4 per texture: offset (texels), width, height, 0
    const float *texels;         // 4 floats per texel
    uint32_t n_samplers;
    const uint32_t *samplers;    // 14 words each (see Smp)
    const uint32_t *mat_smp;     // 3 per material or NULL
    uint32_t env_smp;
};

// Sampler hooks for distribution tests: kind 0 hemisphere(n, n), 1 diffuse(n),
// 2 lobe(n, dir, e). out: 3 doubles per sample; draws of sample i come from
// stream (seed, i). Returns the total number of draws consumed.
uint64_t or_sample(int kind, const double *n, const double *dir, double e, uint32_t seed, uint64_t count,
                   double *out) {
    uint64_t used = 0;
    for (uint64_t i = 0; i < count; ++i) {
        Draws r{stream_of(path_key(seed, 0, (uint32_t)i, 0), 1), 0};
        V nn = mk(n[0], n[1], n[2]);
        V v = kind == 0 ? hemisphere(nn, nn, r) : kind == 1 ? diffuse(nn, r) : lobe(nn, mk(dir[0], dir[1], dir[2]), e, r);
        out[3 * i] = v.x;
        out[3 * i + 1] = v.y;
        out[3 * i + 2] = v.z;
        used += r.i;
    }
    return used;
}

void *or_build(const or_scene_desc *in) {
    Scene *s = new Scene;
    for (uint32_t i = 0; i < in->n_objects; ++i) {
        Obj o{};
        o.type = in->obj_type[i];
        o.mat = in->obj_material[i];
        const double *p = in->obj_param + 9 * (size_t)i;
        o.mesh = -1;
        if (o.type == 0) {
            o.pn = mk(p[0], p[1], p[2]);
            o.poff = p[3];
        } else if (o.type == 1) {
            o.sc = mk(p[0], p[1], p[2]);
            o.sr = p[3];
        } else if (o.type == 2) {
            for (int k = 0; k < 3; ++k) o.tri.v[k] = mk(p[3 * k], p[3 * k + 1], p[3 * k + 2]);
            if (in->obj_tc)
                for (int k = 0; k < 6; ++k) o.tri.tc[k / 2][k % 2] = in->obj_tc[6 * (size_t)i + k];
        } else if (o.type == 3) {
            Mesh m;
            for (uint32_t t = 0; t < in->mesh_count[i]; ++t) {
                size_t g = (size_t)in->mesh_first[i] + t;
                Tri tr;
                for (int k = 0; k < 3; ++k) {
                    tr.v[k] = mk(in->mesh_v[9 * g + 3 * k], in->mesh_v[9 * g + 3 * k + 1], in->mesh_v[9 * g + 3 * k + 2]);
                    if (in->mesh_n)
                        tr.n[k] = mk(in->mesh_n[9 * g + 3 * k], in->mesh_n[9 * g + 3 * k + 1], in->mesh_n[9 * g + 3 * k + 2]);
                    if (in->mesh_tc) {
                        tr.tc[k][0] = in->mesh_tc[6 * g + 2 * k];
                        tr.tc[k][1] = in->mesh_tc[6 * g + 2 * k + 1];
                    }
                }
                m.tris.push_back(tr);
            }
            build_mesh(m);
            o.mesh = (int)s->meshes.size();
            s->meshes.push_back(std::move(m));
        }
        if (in->obj_uv_scale) {
            o.us = in->obj_uv_scale[2 * (size_t)i];
            o.vs = in->obj_uv_scale[2 * (size_t)i + 1];
        }
        s->objs.push_back(o);
    }
    for (uint32_t i = 0; i < in->n_materials; ++i) {
        Mat m{};
        m.type = in->mat_type[i];
        const float *p = in->mat_param + 11 * (size_t)i;
        for (int k = 0; k < 3; ++k) {
            m.kd[k] = p[k];
            m.ks[k] = p[3 + k];
            m.ke[k] = p[6 + k];
        }
        m.refl = p[9];
        m.expo = p[10];
        s->mats.push_back(m);
    }
    s->env_type = in->env_type;
    for (int k = 0; k < 3; ++k) {
        s->env_a[k] = in->env ? in->env[k] : 0.f;
        s->env_b[k] = in->env ? in->env[3 + k] : 0.f;
    }
    for (uint32_t t = 0; t < in->n_textures; ++t) {
        Tex x;
        x.w = in->tex_info[4 * t + 1];
        x.h = in->tex_info[4 * t + 2];
        x.px = in->texels + 4 * (size_t)in->tex_info[4 * t];
        s->texs.push_back(x);
    }
    for (uint32_t k = 0; k < in->n_samplers; ++k) {
        Smp sm;
        std::memcpy(&sm, in->samplers + 14 * (size_t)k, sizeof sm);
        s->smps.push_back(sm);
    }
    if (in->mat_smp)
        for (size_t i = 0; i < s->mats.size(); ++i)
            for (int c = 0; c < 3; ++c) s->mats[i].smp[c] = in->mat_smp[3 * i + c];
    s->env_smp = in->env_smp;
    return s;
}

void or_free(void *h) { delete (Scene *)h; }

// Mesh octree statistics for object `obj` (must be a mesh).
int or_mesh_octree(void *h, uint32_t obj, uint64_t *nodes, uint64_t *preorder) {
    Scene *s = (Scene *)h;
    if (obj >= s->objs.size() || s->objs[obj].type != 3) return 1;
    const Mesh &m = s->meshes[s->objs[obj].mesh];
    if (nodes) *nodes = m.nodes.size();
    if (preorder) {
        size_t k = 0;
        std::vector<int> st{0};
        while (!st.empty()) {
            int n = st.back();
            st.pop_back();
            for (uint32_t it : m.nodes[n].items) preorder[k++] = it;
            for (int c = 7; c >= 0; --c)
                if (m.nodes[n].child[c] >= 0) st.push_back(m.nodes[n].child[c]);
        }
    }
    return 0;
}

// Closest hit per ray (double SoA in: o xyz, d xyz). Outputs t (INF on a
// miss), object index (-1), mesh-local triangle (-1 unless a mesh), point and
// normal (3 doubles each, may be NULL).
void or_trace(void *h, uint64_t n, const double *rays, double *t, int32_t *obj, int64_t *prim, double *point,
              double *normal, int threads) {
    Scene *s = (Scene *)h;
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 256)
    for (long long i = 0; i < (long long)n; ++i) {
        const double *r = rays + 6 * (size_t)i;
        Hit hh;
        bool ok = scene_hit(*s, mk(r[0], r[1], r[2]), mk(r[3], r[4], r[5]), hh);
        t[i] = ok ? hh.t : INFINITY;
        obj[i] = ok ? hh.obj : -1;
        prim[i] = ok ? hh.prim : -1;
        if (point) {
            point[3 * i] = hh.point.x;
            point[3 * i + 1] = hh.point.y;
            point[3 * i + 2] = hh.point.z;
        }
        if (normal) {
            normal[3 * i] = hh.normal.x;
            normal[3 * i + 1] = hh.normal.y;
            normal[3 * i + 2] = hh.normal.z;
        }
    }
}

struct or_job {
    uint32_t integrator;  // 0 stencil, 1 depth, 2 ao, 3 pathtracer, 4 normal, 6 emission, 7 wireframe
    uint32_t width, height, samples, aa, aa_mode, rdepth, seed, job_key;
    uint32_t tile_size;   // 32 (context.cc defaults)
    uint32_t tile_stride; // render every k-th tile in row-major order (1: all)
    int threads;
    double cam[9];        // position, target, up
    float fov, aperture, flength;
    uint32_t cam_kind;    // 0 thin-lens, 1 ods, 2 erp, 3 cubemap
    double orientation[3];
    float ipd;
};

// Renders with the reference's per-sample float read-modify-write (tile.cc,
// color.cc clamps) into rgba (W*H*4, zero-initialised by the caller). When
// mean/m2 are given (W*H*3 doubles each) the per-pixel sample mean and sum of
// squared deviations (Welford, FP64) of the per-sample values are returned
// too. Returns the number of executed scene queries.
static Cam cam_of(const or_job *jb) {
    Cam cam;
    cam.pos = mk(jb->cam[0], jb->cam[1], jb->cam[2]);
    cam.target = mk(jb->cam[3], jb->cam[4], jb->cam[5]);
    cam.up = mk(jb->cam[6], jb->cam[7], jb->cam[8]);
    cam.fov = jb->fov;
    cam.aperture = jb->aperture;
    cam.flength = jb->flength;
    cam.kind = jb->cam_kind;
    cam.orientation = mk(jb->orientation[0], jb->orientation[1], jb->orientation[2]);
    cam.ipd = jb->ipd;
    cam_setup(cam);
    return cam;
}

// Primary rays of the job's camera at n sample coordinates xy (float pairs);
// a thin lens draws from the stream of key 0.
void or_camera_rays(const or_job *jb, uint64_t n, const float *xy, double *o, double *d) {
    const Cam cam = cam_of(jb);
    for (uint64_t i = 0; i < n; ++i) {
        V a, b;
        primary(cam, xy[2 * i], xy[2 * i + 1], (float)jb->width, (float)jb->height, 0u, a, b);
        o[3 * i] = a.x;
        o[3 * i + 1] = a.y;
        o[3 * i + 2] = a.z;
        d[3 * i] = b.x;
        d[3 * i + 1] = b.y;
        d[3 * i + 2] = b.z;
    }
}

// Thin-lens primary rays with the lens jitters given (2 doubles per ray, the
// values prng_c returned): the camera arithmetic alone, for comparison with
// the reference's rays under recorded draws.
void or_thin_lens_rays(const or_job *jb, uint64_t n, const float *xy, const double *jitter, double *o, double *d) {
    const Cam cam = cam_of(jb);
    for (uint64_t i = 0; i < n; ++i) {
        V a, b;
        primary_persp_at(cam, xy[2 * i], xy[2 * i + 1], (float)jb->width, (float)jb->height, jitter[2 * i],
                         jitter[2 * i + 1], a, b);
        o[3 * i] = a.x;
        o[3 * i + 1] = a.y;
        o[3 * i + 2] = a.z;
        d[3 * i] = b.x;
        d[3 * i + 1] = b.y;
        d[3 * i + 2] = b.z;
    }
}

// The lens jitters the oracle draws for sample keys `keys` (2 per key).
void or_lens_jitter(const or_job *jb, uint64_t n, const uint32_t *keys, double *jitter) {
    const Cam cam = cam_of(jb);
    for (uint64_t i = 0; i < n; ++i) lens_jitter(cam, keys[i], jitter[2 * i], jitter[2 * i + 1]);
}

// RANDOM anti-aliasing: with draws NULL, the oracle's own draws for the
// `level` sub-samples of each pixel (pixel index y * width + x) are written
// to draws_out (2 per sub-sample, may be NULL); with draws given, those are
// used. coords receives the float sample coordinates (2 per sub-sample).
void or_aa_random(const or_job *jb, uint64_t n, const uint32_t *pixels, const double *draws, double *draws_out,
                  float *coords) {
    const uint32_t L = jb->aa;
    for (uint64_t i = 0; i < n; ++i) {
        const uint32_t x = pixels[2 * i], y = pixels[2 * i + 1];
        for (uint32_t a = 0; a < L; ++a) {
            const uint64_t k = i * L + a;
            double ux, uy;
            if (draws) {
                ux = draws[2 * k];
                uy = draws[2 * k + 1];
            } else {
                aa_random_draws(jb->seed, jb->job_key, y * jb->width + x, a, ux, uy);
            }
            if (draws_out) {
                draws_out[2 * k] = ux;
                draws_out[2 * k + 1] = uy;
            }
            aa_random_at(x, y, ux, uy, coords[2 * k], coords[2 * k + 1]);
        }
    }
}

// The sample key of (pixel, sample) the device and the oracle derive their
// random streams from.
uint32_t or_path_key(uint32_t seed, uint32_t job_key, uint32_t pixel, uint32_t sample) {
    return path_key(seed, job_key, pixel, sample);
}

// IMaterial::shade (material/lambert.cc:9-28, phong.cc:11-33,
// blinnphong.cc:9-32, emissive.cc:8-18, dielectric.cc:9-21): the light of one
// point emitter through object oi's material. q: 16 doubles per query (hit
// point, normal, texcoord, camera position, emitter position, object index);
// ei: emitter intensity, io: the intensity argument, in and out (3 floats
// each, ColorRGBf: clamped at 0); ret: shade()'s result, or -1 for an object
// without a material.
int or_shade(void *h, uint64_t n, const double *q, const float *ei, float *io, int32_t *ret) {
    const Scene &s = *(const Scene *)h;
    for (uint64_t k = 0; k < n; ++k) {
        const double *a = q + 16 * k;
        const uint32_t oi = (uint32_t)a[15];
        ret[k] = -1;
        if (oi >= s.objs.size() || s.objs[oi].mat >= s.mats.size()) continue;
        const Mat &m = s.mats[s.objs[oi].mat];
        const V P = mk(a[0], a[1], a[2]), N = mk(a[3], a[4], a[5]), TC = mk(a[6], a[7], a[8]);
        const V cam = mk(a[9], a[10], a[11]), L = mk(a[12], a[13], a[14]);
        auto chan = [&](int c, const float *konst) {
            return m.smp[c] != 0xffffffffu ? sample(s, s.smps[m.smp[c]], TC) : C3{konst[0], konst[1], konst[2]};
        };
        const C3 E{cl(ei[3 * k]), cl(ei[3 * k + 1]), cl(ei[3 * k + 2])};
        C3 I{cl(io[3 * k]), cl(io[3 * k + 1]), cl(io[3 * k + 2])};
        int r = 1;
        auto mx = [](double x, double y) { return x > y ? x : y; };  // nmath::max (mutil.inl:41)
        if (m.type == 0) {  // Lambert: I += E * d * kd when d > 0
            const V ld = nrm(L - P);
            const double d = dot(ld, N);
            if (d > 0) I = cadd(I, cmul(cscale(E, (float)d), chan(0, m.kd)));
        } else if (m.type == 1) {  // Phong: I += E * (d kd + ks max(0, r.v)^e)
            const V ld = nrm(L - P);
            const double d = mx(0, dot(ld, N));
            const V v = nrm(cam - P);
            const V li = nrm(ld), nn = nrm(N);
            const V rf = nrm((2 * dot(li, nn)) * nn - li);  // Vector3f::reflected (vector.inl:455-460)
            const double rmv = mx(0, dot(rf, v));
            const float pw = (float)powl((long double)rmv, (long double)m.expo);
            I = cadd(I, cmul(E, cadd(cscale(chan(0, m.kd), (float)d), cscale(chan(1, m.ks), pw))));
        } else if (m.type == 2) {  // Blinn-Phong: I = E * (d kd + ks max(h.n, 0)^e)
            const V ld = nrm(L - P);
            const double d = mx(dot(ld, N), 0);
            const V hv = nrm(ld + nrm(cam - P));
            const double rmv = mx(dot(hv, N), 0);
            const float pw = (float)powl((long double)rmv, (long double)m.expo);
            I = cmul(E, cadd(cscale(chan(0, m.kd), (float)d), cscale(chan(1, m.ks), pw)));
        } else if (m.type == 3) {  // emissive: I = ke, returns false
            I = chan(2, m.ke);
            r = 0;
        } else {  // dielectric: I = 0
            I = C3{};
        }
        io[3 * k] = I.r;
        io[3 * k + 1] = I.g;
        io[3 * k + 2] = I.b;
        ret[k] = r;
    }
    return 0;
}

uint64_t or_render(void *h, const or_job *jb, float *rgba, double *mean, double *m2) {
    Scene *s = (Scene *)h;
    const Cam cam = cam_of(jb);
    const uint32_t W = jb->width, H = jb->height, T = jb->tile_size ? jb->tile_size : 32;
    const uint32_t tx = W / T + (W % T ? 1 : 0), ty = H / T + (H % T ? 1 : 0);
    const uint32_t ntiles = tx * ty, stride = jb->tile_stride ? jb->tile_stride : 1;
    const uint32_t L = jb->aa;
    uint32_t n_aa;
    float aw;
    if (jb->aa_mode == 0) {
        n_aa = L == 0 ? 0 : L * L;
        aw = L <= 1 ? 1.f : 1.f / (L * L);
    } else {
        n_aa = L;
        aw = L ? 1.f / L : 1.f;
    }
    const double weight = (double)aw * (jb->samples ? 1. / jb->samples : 1.);
    Ctx cx{s, &cam, {jb->integrator, W, H, jb->samples, L, jb->aa_mode, jb->rdepth, jb->seed, jb->job_key}, nullptr};
    uint64_t total_q = 0;
    if (jb->threads > 0) omp_set_num_threads(jb->threads);
#pragma omp parallel for schedule(dynamic) reduction(+ : total_q)
    for (long long ti = 0; ti < (long long)ntiles; ti += stride) {
        const uint32_t x0 = (uint32_t)(ti % tx) * T, y0 = (uint32_t)(ti / tx) * T;
        const uint32_t x1 = std::min(x0 + T, W), y1 = std::min(y0 + T, H);
        uint64_t nq = 0;
        for (uint32_t x = x0; x < x1; ++x) {  // produce(): x outer, y inner (aa.cc:55-56)
            for (uint32_t y = y0; y < y1; ++y) {
                const uint32_t pix = y * W + x;
                float *px = rgba + 4 * (size_t)pix;
                double mu[3] = {0, 0, 0}, q2[3] = {0, 0, 0};
                uint32_t cnt = 0;
                for (uint32_t a = 0; a < n_aa; ++a) {
                    float cxp, cyp;
                    if (jb->aa_mode == 1) {
                        double ux, uy;
                        aa_random_draws(jb->seed, jb->job_key, pix, a, ux, uy);
                        aa_random_at(x, y, ux, uy, cxp, cyp);
                    } else if (L <= 1) {
                        cxp = (float)((double)x + 0.5);
                        cyp = (float)((double)y + 0.5);
                    } else {
                        float sub = 1.f / L, off = sub * 0.5f;
                        cxp = (float)((double)x + (double)(a % L) * (double)sub + (double)off);
                        cyp = (float)((double)y + (double)(a / L) * (double)sub + (double)off);
                    }
                    for (uint32_t k = 0; k < jb->samples; ++k) {
                        const uint32_t sidx = a * jb->samples + k;
                        const uint32_t key = path_key(jb->seed, jb->job_key, pix, sidx);
                        V o, d;
                        primary(cam, cxp, cyp, (float)W, (float)H, key, o, d);
                        double val[3] = {0, 0, 0};
                        if (jb->integrator == 0) {  // stencil/integrator.cc:9-34
                            Hit hh;
                            ++nq;
                            if (scene_hit(*s, o, d, hh)) {
                                float na = (float)((double)px[3] + weight);
                                px[0] = px[1] = px[2] = 1.f;
                                px[3] = na < 0.f ? 0.f : (na > 1.f ? 1.f : na);
                                val[0] = val[1] = val[2] = 1;
                            }
                        } else if (jb->integrator == 1) {  // depth/integrator.cc:11-55
                            Hit hh;
                            ++nq;
                            if (scene_hit(*s, o, d, hh)) {
                                float dep = 1. / std::log(len(o - hh.point));
                                float r = (float)((double)px[0] + (double)dep * weight);
                                r = r < 0.f ? 0.f : r;
                                px[0] = px[1] = px[2] = r;
                                val[0] = val[1] = val[2] = dep;
                            }
                            px[3] = 1.f;
                        } else if (jb->integrator == 2) {  // ao/integrator.cc:12-53
                            Hit hh;
                            float w = 1.0f;
                            ++nq;
                            if (scene_hit(*s, o, d, hh)) {
                                Draws r{stream_of(key, 1), 0};
                                V ad = hemisphere(hh.normal, hh.normal, r);
                                V ao = hh.point + hh.normal * 1e-8;
                                Hit h2;
                                ++nq;
                                if (scene_hit(*s, ao, ad, h2)) {
                                    float dist = len(h2.point - ao);
                                    w = std::min(dist, 100.0f) / 100.0f;
                                }
                            }
                            for (int c = 0; c < 3; ++c) {
                                float v = (float)((double)px[c] + (double)w * weight);
                                px[c] = v < 0.f ? 0.f : v;
                                val[c] = w;
                            }
                            px[3] = 1.f;
                        } else if (jb->integrator == 4) {  // normal/integrator.cc:11-41
                            Hit hh;
                            ++nq;
                            V acc = mk((double)px[0], (double)px[1], (double)px[2]) * 2.0 - 1.0;
                            if (scene_hit(*s, o, d, hh)) {
                                V nn = (hh.normal + acc) * 0.5;
                                acc = (nrm(nn) + 1.0) * 0.5;
                                val[0] = hh.normal.x;
                                val[1] = hh.normal.y;
                                val[2] = hh.normal.z;
                            }
                            // ColorRGBAf(acc, 1): negative channels clamp to 0
                            const float fr = (float)acc.x, fg = (float)acc.y, fb = (float)acc.z;
                            px[0] = fr < 0.f ? 0.f : fr;
                            px[1] = fg < 0.f ? 0.f : fg;
                            px[2] = fb < 0.f ? 0.f : fb;
                            px[3] = 1.f;
                        } else if (jb->integrator == 5) {  // uv/integrator.cc:11-44
                            Hit hh;
                            ++nq;
                            V acc = mk((double)px[0], (double)px[1], (double)px[2]);
                            const double alpha = (double)px[3];
                            double alpha_sample = 0.f;
                            if (scene_hit(*s, o, d, hh)) {
                                // (texcoord * 0.5f + 0.5f) * weight, texcoord.z = 0
                                acc = acc + mk(hh.tu * 0.5f + 0.5f, hh.tv * 0.5f + 0.5f, 0.0 * 0.5f + 0.5f) * weight;
                                alpha_sample += weight;
                                val[0] = hh.tu;
                                val[1] = hh.tv;
                            }
                            // ColorRGBAf(acc, 0.) then a(alpha + weight * alpha_sample)
                            const float fr = (float)acc.x, fg = (float)acc.y, fb = (float)acc.z;
                            px[0] = fr < 0.f ? 0.f : fr;
                            px[1] = fg < 0.f ? 0.f : fg;
                            px[2] = fb < 0.f ? 0.f : fb;
                            const float na = (float)(alpha + weight * alpha_sample);
                            px[3] = na < 0.f ? 0.f : (na > 1.f ? 1.f : na);
                        } else if (jb->integrator == 6) {  // emission/integrator.cc:11-40
                            Hit hh;
                            ++nq;
                            if (scene_hit(*s, o, d, hh)) {
                                const Obj &ob = s->objs[hh.obj];
                                if (ob.mat != 0xffffffffu) {
                                    const Mat &m = s->mats[ob.mat];
                                    const float wf = (float)weight;  // operator*(ColorRGBf, float)
                                    const C3 ke = channel(*s, m, 2, m.ke, hh.tu, hh.tv);
                                    const float kv[3] = {ke.r, ke.g, ke.b};
                                    for (int c = 0; c < 3; ++c) {
                                        const float e = cl(kv[c] * wf);
                                        px[c] = cl(px[c] + e);
                                        val[c] = kv[c];
                                    }
                                }
                            }
                            px[3] = 1.f;  // ColorRGBAf = ColorRGBf sets alpha 1
                        } else if (jb->integrator == 7) {  // wireframe/integrator.cc:11-36
                            Hit hh;
                            ++nq;
                            if (scene_hit(*s, o, d, hh)) {
                                float na = (float)((double)px[3] + weight);
                                px[0] = px[1] = px[2] = 1.f;
                                px[3] = na < 0.f ? 0.f : (na > 1.f ? 1.f : na);
                                val[0] = val[1] = val[2] = 1;
                            }
                        } else {  // pathtracer/integrator.cc:51-74
                            C3 L3 = eval(cx, jb->rdepth, C3{1, 1, 1}, o, d, key, 1, nq);
                            C3 add = cscale(L3, (float)weight);
                            px[0] = cl(px[0] + add.r);
                            px[1] = cl(px[1] + add.g);
                            px[2] = cl(px[2] + add.b);
                            px[3] = 1.f;
                            val[0] = L3.r;
                            val[1] = L3.g;
                            val[2] = L3.b;
                        }
                        ++cnt;
                        for (int c = 0; c < 3; ++c) {
                            double dlt = val[c] - mu[c];
                            mu[c] += dlt / cnt;
                            q2[c] += dlt * (val[c] - mu[c]);
                        }
                    }
                }
                if (mean)
                    for (int c = 0; c < 3; ++c) mean[3 * (size_t)pix + c] = mu[c];
                if (m2)
                    for (int c = 0; c < 3; ++c) m2[3 * (size_t)pix + c] = q2[c];
            }
        }
        total_q += nq;
    }
    return total_q;
}

}  // extern "C"
