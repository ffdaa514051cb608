#include "flat.h"

#include <chrono>
#include <cmath>
#include <cstring>
#include <unordered_map>

namespace xt {

namespace {

// Triangle area exactly as Triangle::calc_barycentric computes it
// (src/xtcore/math/triangle.cc:149-166); below 1e-8 the triangle is never hit.
double ref_area(const dvec3 &a, const dvec3 &b, const dvec3 &c) {
    dvec3 v1 = b - a, v2 = c - a;
    dvec3 x = cross(v1, v2);
    dvec3 n = normalized(x);
    return std::fabs(dot(x, n)) * 0.5;
}

float round_down(double v) {
    float f = (float)v;
    if ((double)f > v) f = std::nextafter(f, -INFINITY);
    return f;
}
float round_up(double v) {
    float f = (float)v;
    if ((double)f < v) f = std::nextafter(f, INFINITY);
    return f;
}

void set3(float *d, const dvec3 &v) {
    d[0] = (float)v.x;
    d[1] = (float)v.y;
    d[2] = (float)v.z;
}
void set3d(double *d, const dvec3 &v) {
    d[0] = v.x;
    d[1] = v.y;
    d[2] = v.z;
}

void tri_finish(XtPrim &p, const dvec3 v[3], bool has_vn) {
    uint32_t flags = has_vn ? XT_PF_VNORMALS : 0u;
    if (ref_area(v[0], v[1], v[2]) < 1e-8) flags |= XT_PF_DEGENERATE;
    dvec3 e1(p.p1[0] - (double)p.p0[0], p.p1[1] - (double)p.p0[1], p.p1[2] - (double)p.p0[2]);
    dvec3 e2(p.p2[0] - (double)p.p0[0], p.p2[1] - (double)p.p0[1], p.p2[2] - (double)p.p0[2]);
    double nl = length(cross(e1, e2));
    p.aux = (flags & XT_PF_DEGENERATE) || nl == 0.0 ? 0.f : (float)(1.0 / nl);
    p.tf = XT_PRIM_TRIANGLE | (flags << 4);
}

struct MeshTri {
    dvec3 v[3];
    float n[9];
    double tc[6];  // u, v per corner
    bool has_n;
};

void mesh_tris(const MeshData &m, std::vector<MeshTri> &out) {
    const size_t nt = m.tri_count();
    out.resize(nt);
    for (size_t i = 0; i < nt; ++i) {
        MeshTri &t = out[i];
        t.has_n = false;
        // a corner without a texture index keeps Triangle()'s texcoord
        // (triangle.cc:12-17: (0, 0), (1, 0), (0, 0)); mesh.cc:56-58
        static const double tc_default[6] = {0, 0, 1, 0, 0, 0};
        for (int k = 0; k < 3; ++k) {
            int vi = m.iv[3 * i + k], ni = m.in[3 * i + k], ti = m.it[3 * i + k];
            t.v[k] = vi >= 0 ? dvec3(m.v[3 * vi], m.v[3 * vi + 1], m.v[3 * vi + 2]) : dvec3(0, 0, 0);
            t.tc[2 * k] = ti >= 0 ? (double)m.uv[2 * ti] : tc_default[2 * k];
            t.tc[2 * k + 1] = ti >= 0 ? (double)m.uv[2 * ti + 1] : tc_default[2 * k + 1];
            if (ni >= 0) {
                t.n[3 * k] = m.n[3 * ni];
                t.n[3 * k + 1] = m.n[3 * ni + 1];
                t.n[3 * k + 2] = m.n[3 * ni + 2];
                t.has_n = true;
            } else {
                t.n[3 * k] = t.n[3 * k + 1] = t.n[3 * k + 2] = 0.f;
            }
        }
    }
}

Box64 tri_box(const dvec3 v[3]) {
    Box64 b;
    b.lo = dvec3(INFINITY, INFINITY, INFINITY);
    b.hi = dvec3(-INFINITY, -INFINITY, -INFINITY);
    for (int k = 0; k < 3; ++k)
        for (int a = 0; a < 3; ++a) {
            if (v[k][a] < b.lo[a]) b.lo[a] = v[k][a];
            if (v[k][a] > b.hi[a]) b.hi[a] = v[k][a];
        }
    return b;
}

}  // namespace

int flatten_scene(const Scene &sc, FlatScene &out, const KdParams &kdp, std::string *err) {
    out = FlatScene();
    for (const Image &im : sc.textures) {
        out.textures.push_back(XtTexture{(uint32_t)(out.texels.size() / 4), im.width, im.height, 0u});
        out.texels.insert(out.texels.end(), im.rgba.begin(), im.rgba.end());
    }
    auto add_sampler = [&](const SamplerDesc &sd) {
        XtSampler x{};
        x.kind = sd.kind;
        x.bilinear = sd.bilinear;
        for (int k = 0; k < 6; ++k) x.tex[k] = sd.tex[k];
        std::memcpy(x.a, sd.a, sizeof x.a);
        std::memcpy(x.b, sd.b, sizeof x.b);
        out.samplers.push_back(x);
        return (uint32_t)out.samplers.size() - 1;
    };
    // materials, in handle order
    std::unordered_map<uint64_t, uint32_t> mat_index;
    for (auto &kv : sc.materials) {
        const Material &m = kv.second;
        XtMaterial x{};
        x.type = m.type == MAT_LAMBERT ? XT_MAT_LAMBERT
                 : m.type == MAT_PHONG ? XT_MAT_PHONG
                 : m.type == MAT_BLINNPHONG ? XT_MAT_BLINNPHONG
                 : m.type == MAT_DIELECTRIC ? XT_MAT_DIELECTRIC
                 : XT_MAT_EMISSIVE;
        std::memcpy(x.kd, m.diffuse, sizeof x.kd);
        std::memcpy(x.ks, m.specular, sizeof x.ks);
        std::memcpy(x.ke, m.emissive, sizeof x.ke);
        x.reflectance = m.reflectance;
        x.exponent = m.exponent;
        mat_index[kv.first] = (uint32_t)out.materials.size();
        out.materials.push_back(x);
        // a colour channel stays in kd / ks / ke; other samplers go to the table
        for (int c = 0; c < 3; ++c)
            out.mat_smp.push_back(m.smp[c].kind == SMP_NONE || m.smp[c].kind == SMP_COLOR ? XT_NO_MATERIAL
                                                                                       : add_sampler(m.smp[c]));
        out.material_names.push_back(m.name);
        out.material_supported.push_back(m.unsupported.empty() ? 1 : 0);
    }
    out.env.type = sc.env.type == ENV_GRADIENT ? XT_ENV_GRADIENT
                   : sc.env.type == ENV_COLOR  ? XT_ENV_COLOR
                   : sc.env.type == ENV_CUBEMAP ? XT_ENV_CUBEMAP
                   : sc.env.type == ENV_ERP     ? XT_ENV_ERP
                                                : XT_ENV_NONE;
    std::memcpy(out.env.a, sc.env.a, sizeof out.env.a);
    std::memcpy(out.env.b, sc.env.b, sizeof out.env.b);
    out.env.smp = (sc.env.type == ENV_CUBEMAP || sc.env.type == ENV_ERP) ? add_sampler(sc.env.smp) : 0u;
    out.env_supported = sc.env.type != ENV_UNSUPPORTED;

    // per-geometry pre-ordered triangle lists, built once per mesh geometry
    auto t0 = std::chrono::steady_clock::now();
    std::unordered_map<uint64_t, std::vector<uint32_t>> mesh_order;
    std::unordered_map<uint64_t, std::vector<MeshTri>> mesh_data;
    for (auto &kv : sc.objects) {
        auto g = sc.geometries.find(kv.second.geometry);
        if (g == sc.geometries.end() || g->second.type != GEOM_MESH) continue;
        if (mesh_order.count(g->first)) continue;
        std::vector<MeshTri> &tris = mesh_data[g->first];
        mesh_tris(g->second.mesh, tris);
        std::vector<Box64> boxes(tris.size());
        for (size_t i = 0; i < tris.size(); ++i) boxes[i] = tri_box(tris[i].v);
        OctreeBuild ob;
        octree_build(boxes, 10, 10, ob);
        std::vector<uint32_t> &order = mesh_order[g->first];
        octree_preorder(ob, order);
        size_t maxl = 0;
        for (auto &n : ob.nodes) maxl = std::max<size_t>(maxl, n.level);
        out.mesh_octrees.push_back({g->first, ob.nodes.size(), maxl, tris.size()});
    }

    // objects in scan order
    out.obj_uv_scale.assign(2 * sc.objects.size(), 0.0);
    for (auto &kv : sc.objects) {
        const Object &o = kv.second;
        FlatObject fo;
        fo.handle = o.handle;
        fo.name = o.name;
        auto mi = mat_index.find(o.material);
        fo.material = mi == mat_index.end() ? XT_NO_MATERIAL : mi->second;
        const uint32_t oi = (uint32_t)out.objects.size();
        auto gi = sc.geometries.find(o.geometry);
        if (gi == sc.geometries.end()) {
            // Scene::intersection skips objects whose geometry is missing
            fo.geom_type = 0xffffffffu;
            out.objects.push_back(fo);
            continue;
        }
        const Geometry &g = gi->second;
        fo.geom_type = g.type;
        fo.first_prim = (uint32_t)out.prims.size();
        out.obj_uv_scale[2 * oi] = g.u_scale;
        out.obj_uv_scale[2 * oi + 1] = g.v_scale;
        if (g.type == GEOM_PLANE) {
            XtPlane p{};
            dvec3 pt = dvec3(std::fabs(g.normal.x), std::fabs(g.normal.y), std::fabs(g.normal.z)) * g.offset;
            set3(p.normal, g.normal);
            set3(p.point, pt);
            set3d(p.normal64, g.normal);
            set3d(p.point64, pt);
            p.object = oi;
            fo.plane = (int32_t)out.planes.size();
            out.planes.push_back(p);
        } else if (g.type == GEOM_SPHERE) {
            XtPrim p{};
            set3(p.p0, g.center);
            p.p1[0] = (float)g.radius;
            p.p1[1] = (float)(g.radius * g.radius);
            p.p1[2] = g.radius != 0.0 ? (float)(1.0 / g.radius) : 0.f;
            p.object = oi;
            p.tf = XT_PRIM_SPHERE | ((g.radius == 0.0 ? XT_PF_TINY_SPHERE : 0u) << 4);
            XtPrim64 q{};
            set3d(q.a, g.center);
            q.b[0] = g.radius;
            out.prims.push_back(p);
            out.prims64.push_back(q);
            out.tc64.insert(out.tc64.end(), 6, 0.0);
            out.prim_local.push_back(0);
            fo.prim_count = 1;
        } else if (g.type == GEOM_TRIANGLE) {
            XtPrim p{};
            set3(p.p0, g.v[0]);
            set3(p.p1, g.v[1]);
            set3(p.p2, g.v[2]);
            p.object = oi;
            tri_finish(p, g.v, false);
            XtPrim64 q{};
            set3d(q.a, g.v[0]);
            set3d(q.b, g.v[1]);
            set3d(q.c, g.v[2]);
            out.prims.push_back(p);
            out.prims64.push_back(q);
            for (int k = 0; k < 3; ++k) out.tc64.insert(out.tc64.end(), {g.tc[k][0], g.tc[k][1]});
            out.prim_local.push_back(0);
            fo.prim_count = 1;
        } else {
            const std::vector<uint32_t> &order = mesh_order[g.handle];
            const std::vector<MeshTri> &tris = mesh_data[g.handle];
            for (uint32_t li : order) {
                const MeshTri &t = tris[li];
                XtPrim p{};
                set3(p.p0, t.v[0]);
                set3(p.p1, t.v[1]);
                set3(p.p2, t.v[2]);
                p.object = oi;
                tri_finish(p, t.v, t.has_n);
                XtPrim64 q{};
                set3d(q.a, t.v[0]);
                set3d(q.b, t.v[1]);
                set3d(q.c, t.v[2]);
                out.prims.push_back(p);
                out.prims64.push_back(q);
                out.tc64.insert(out.tc64.end(), t.tc, t.tc + 6);
                out.prim_local.push_back(li);
            }
            fo.prim_count = (uint32_t)order.size();
        }
        out.objects.push_back(fo);
    }
    // vertex normals, only materialised when some triangle has them
    bool any_vn = false;
    for (auto &p : out.prims) any_vn |= (XT_PRIM_FLAGS(p.tf) & XT_PF_VNORMALS) != 0;
    if (any_vn) {
        out.vnormals.assign(out.prims.size() * 9, 0.f);
        for (const FlatObject &fo : out.objects) {
            if (fo.geom_type != GEOM_MESH) continue;
            auto gi = sc.geometries.find(sc.objects.at(fo.handle).geometry);
            const std::vector<MeshTri> &tris = mesh_data[gi->first];
            for (uint32_t k = 0; k < fo.prim_count; ++k) {
                uint32_t s = fo.first_prim + k;
                std::memcpy(&out.vnormals[9 * (size_t)s], tris[out.prim_local[s]].n, 9 * sizeof(float));
            }
        }
    }
    auto t1 = std::chrono::steady_clock::now();

    // soup boxes
    const size_t np = out.prims.size();
    std::vector<Box64> boxes(np);
    std::vector<Box32> boxes32(np);
    for (size_t i = 0; i < np; ++i) {
        const XtPrim64 &q = out.prims64[i];
        Box64 b;
        if (XT_PRIM_TYPE(out.prims[i].tf) == XT_PRIM_SPHERE) {
            double r = std::fabs(q.b[0]);
            b.lo = dvec3(q.a[0] - r, q.a[1] - r, q.a[2] - r);
            b.hi = dvec3(q.a[0] + r, q.a[1] + r, q.a[2] + r);
        } else {
            dvec3 v[3] = {dvec3(q.a[0], q.a[1], q.a[2]), dvec3(q.b[0], q.b[1], q.b[2]), dvec3(q.c[0], q.c[1], q.c[2])};
            b = tri_box(v);
        }
        boxes[i] = b;
        for (int a = 0; a < 3; ++a) {
            boxes32[i].lo[a] = round_down(b.lo[a]);
            boxes32[i].hi[a] = round_up(b.hi[a]);
        }
    }
    // octree over the soup (reference limits and child layout)
    {
        OctreeBuild ob;
        octree_build(boxes, 10, 10, ob);
        std::vector<uint32_t> ident(np);
        for (size_t i = 0; i < np; ++i) ident[i] = (uint32_t)i;
        octree_flatten(ob, ident, out.oct_nodes, out.oct_refs);
    }
    auto t2 = std::chrono::steady_clock::now();
    std::vector<uint8_t> is_tri(np);
    for (size_t i = 0; i < np; ++i) is_tri[i] = XT_PRIM_TYPE(out.prims[i].tf) == XT_PRIM_TRIANGLE;
    static_assert(sizeof(XtPrim64) == 9 * sizeof(double), "XtPrim64 is v0 v1 v2 as 9 doubles");
    out.kd_deferred = kdp.on_device;
    out.kd_params = kdp;
    if (!kdp.on_device)
        kdtree_build(boxes32, np ? (const double *)out.prims64.data() : nullptr, is_tri.data(), kdp, out.kd);
    auto t3 = std::chrono::steady_clock::now();
    out.build_seconds_octree = std::chrono::duration<double>(t1 - t0).count() + std::chrono::duration<double>(t2 - t1).count();
    out.build_seconds_kd = std::chrono::duration<double>(t3 - t2).count();
    (void)err;
    return 0;
}

}  // namespace xt
