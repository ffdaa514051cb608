#include "scene.h"

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <unordered_map>
#include <algorithm>
#include <fstream>
#include <sstream>

#include "ncf.h"

namespace xt {

uint64_t Scene::intern(const std::string &name) {
    if (name.empty()) return 0;  // an empty string pools to handle 0 (strpool_inject)
    auto it = names.find(name);
    if (it != names.end()) return it->second;
    uint64_t h = names.size() + 1;
    names.emplace(name, h);
    return h;
}

const std::string *Scene::name_of(uint64_t handle) const {
    for (auto &kv : names)
        if (kv.second == handle) return &kv.first;
    return nullptr;
}

namespace {

double numf(const char *v, double def = 0.0) { return v ? std::strtod(v, nullptr) : def; }

// vec3(...) literals go through sscanf("%f") into float, groups through strtod
// (parseutil.cc:117-145).
dvec3 vec3_of(const ncf::Node *node, const char *name, dvec3 def = dvec3(0, 0, 0)) {
    if (!node) return def;
    if (const ncf::Node *g = node->find_group(name)) {
        const char *x = g->property("x"), *y = g->property("y"), *z = g->property("z");
        return dvec3(x ? std::strtod(x, nullptr) : def.x, y ? std::strtod(y, nullptr) : def.y,
                     z ? std::strtod(z, nullptr) : def.z);
    }
    if (const char *val = node->property(name)) {
        float a, b, c;
        if (std::sscanf(val, "vec3(%f,%f,%f)", &a, &b, &c) == 3) return dvec3(a, b, c);
    }
    return def;
}

void col3_of(const ncf::Node *node, const char *name, float out[3]) {
    float def[3] = {out[0], out[1], out[2]};
    if (!node) return;
    if (const ncf::Node *g = node->find_group(name)) {
        const char *r = g->property("r"), *gg = g->property("g"), *b = g->property("b");
        out[0] = r ? (float)std::strtod(r, nullptr) : def[0];
        out[1] = gg ? (float)std::strtod(gg, nullptr) : def[1];
        out[2] = b ? (float)std::strtod(b, nullptr) : def[2];
        return;
    }
    if (const char *val = node->property(name)) {
        float a, b, c;
        if (std::sscanf(val, "col3(%f,%f,%f)", &a, &b, &c) == 3) {
            out[0] = a;
            out[1] = b;
            out[2] = c;
        }
    }
    // the ColorRGBf constructor keeps channels at or above 0 (color.cc:6-10)
    for (int k = 0; k < 3; ++k) out[k] = out[k] < 0.f ? 0.f : out[k];
}

std::string dir_of(const std::string &path) {
    size_t p = path.find_last_of('/');
    return p == std::string::npos ? std::string() : path.substr(0, p + 1);
}

bool parse_bool(const char *v) {
    if (!v) return false;
    std::string s = v;
    for (auto &c : s) c = (char)std::tolower((unsigned char)c);
    return s == "1" || s == "yes" || s == "true";
}

// deserialize_camera and its four kinds (parseutil.cc:205-275): every
// property read with its default (vec3 0, number 0) when absent.
int load_camera(Scene &sc, ncf::Node *p) {
    const char *type = p->property("type");
    const std::string t = type ? type : "";
    Camera cam;
    if (t == "thin-lens") cam.kind = 0;
    else if (t == "ods") cam.kind = 1;
    else if (t == "erp") cam.kind = 2;
    else if (t == "cubemap") cam.kind = 3;
    else {
        sc.error = "unsupported camera type for " + p->name();
        return 1;
    }
    cam.name = p->name();
    cam.handle = sc.intern(p->name());
    cam.position = vec3_of(p, "position");
    if (cam.kind == 0) {
        cam.target = vec3_of(p, "target");
        cam.up = vec3_of(p, "up");
        cam.fov = (float)numf(p->property("fov"));
        cam.flength = (float)numf(p->property("flength"));
        cam.aperture = (float)numf(p->property("aperture"));
    } else if (cam.kind == 1 || cam.kind == 2) {
        cam.orientation = vec3_of(p, "orientation");
        if (cam.kind == 1) cam.ipd = (float)numf(p->property("ipd"));
    }
    sc.cameras[cam.handle] = cam;
    return 0;
}

// Euler rotation matrix Rx*Ry*Rz (lib/nmath/matrix.cc:517-524).
dmat3 euler(const dvec3 &e) {
    dmat3 rx, ry, rz;
    rx.m[1][1] = std::cos(e.x); rx.m[1][2] = -std::sin(e.x);
    rx.m[2][1] = std::sin(e.x); rx.m[2][2] = std::cos(e.x);
    ry.m[0][0] = std::cos(e.y); ry.m[0][2] = std::sin(e.y);
    ry.m[2][0] = -std::sin(e.y); ry.m[2][2] = std::cos(e.y);
    rz.m[0][0] = std::cos(e.z); rz.m[0][1] = -std::sin(e.z);
    rz.m[1][0] = std::sin(e.z); rz.m[1][1] = std::cos(e.z);
    return mul(mul(rx, ry), rz);
}

// Mesh modifiers in the order parseutil.cc:332-353 applies them
// (lib/nmesh/transform.cc:8-54, lib/nmesh/invnormals.cc).
int parse_sampler(Scene &sc, const std::string &type, ncf::Node *cfg, SamplerDesc &out);

// Red channel of a cubemap sampler at direction (x, y, z): Cubemap::sample
// (sampler_cubemap.cc:14-41) over nimg::sample::nearest (sample.cc:8-17) and
// Pixmap::pixel_ro's edge clamp, in the reference's float / double mix.
float cubemap_red(const Scene &sc, const SamplerDesc &cm, double x, double y, double z) {
    const double l = std::sqrt(x * x + y * y + z * z);
    if (l != 0) {
        x /= l;
        y /= l;
        z /= l;
    }
    int face = 2;
    float u = 0.f, v = 0.f;
    if (std::fabs(x) >= std::fabs(y) && std::fabs(x) >= std::fabs(z)) {
        if (x > 0.0f) {
            face = 1;
            u = (z / x + 1.0f) * 0.5f;
            v = 1.f - (y / x + 1.0f) * 0.5f;
        } else if (x < 0.0f) {
            face = 0;
            u = (z / x + 1.0f) * 0.5f;
            v = (y / x + 1.0f) * 0.5f;
        }
    } else if (std::fabs(y) >= std::fabs(x) && std::fabs(y) >= std::fabs(z)) {
        if (y > 0.0f) {
            face = 2;
            u = (x / y + 1.0f) * 0.5f;
            v = 1.f - (z / y + 1.0f) * 0.5f;
        } else if (y < 0.0f) {
            face = 3;
            u = 1. - (x / y + 1.0f) * 0.5f;
            v = 1.f - (z / y + 1.0f) * 0.5f;
        }
    } else if (std::fabs(z) >= std::fabs(x) && std::fabs(z) >= std::fabs(y)) {
        if (z > 0.0f) {
            face = 5;
            u = 1. - (x / z + 1.0f) * 0.5f;
            v = 1.f - (y / z + 1.0f) * 0.5f;
        } else if (z < 0.0f) {
            face = 4;
            u = 1. - (x / z + 1.0f) * 0.5f;
            v = (y / z + 1.0f) * 0.5f;
        }
    }
    const Image &im = sc.textures[cm.tex[face]];
    double du = u, dv = v;
    du = du < 0 ? 0 : (du > 1 ? du - std::floor(du) : du);
    dv = dv < 0 ? 0 : (dv > 1 ? dv - std::floor(dv) : dv);
    const double fx = du * (float)im.width, fy = dv * (float)im.height;
    size_t px = fx >= 0.0 && fx < 18446744073709551616.0 ? (size_t)fx : (size_t)1 << 63;
    size_t py = fy >= 0.0 && fy < 18446744073709551616.0 ? (size_t)fy : (size_t)1 << 63;
    px = px >= im.width ? im.width - 1 : px;
    py = py >= im.height ? im.height - 1 : py;
    return im.rgba[4 * (py * im.width + px)];
}

int apply_modifiers(Scene &sc, MeshData &m, ncf::Node *mods) {
    dvec3 rot = vec3_of(mods, "rotation", dvec3(0, 0, 0));
    dvec3 scl = vec3_of(mods, "scale", dvec3(1, 1, 1));
    dvec3 tsl = vec3_of(mods, "translation", dvec3(0, 0, 0));
    dmat3 r = euler(rot);
    auto rotate = [&](std::vector<float> &a) {
        for (size_t i = 0; i + 2 < a.size(); i += 3) {
            dvec3 v(a[i], a[i + 1], a[i + 2]);
            dvec3 o = apply(r, v);
            a[i] = (float)(o.x + 0.0);
            a[i + 1] = (float)(o.y + 0.0);
            a[i + 2] = (float)(o.z + 0.0);
        }
    };
    rotate(m.v);
    rotate(m.n);
    for (size_t i = 0; i + 2 < m.v.size(); i += 3) {
        m.v[i] = (float)(m.v[i] * scl.x);
        m.v[i + 1] = (float)(m.v[i + 1] * scl.y);
        m.v[i + 2] = (float)(m.v[i + 2] * scl.z);
    }
    for (size_t i = 0; i + 2 < m.v.size(); i += 3) {
        m.v[i] = (float)(m.v[i] + tsl.x);
        m.v[i + 1] = (float)(m.v[i + 1] + tsl.y);
        m.v[i + 2] = (float)(m.v[i + 2] + tsl.z);
    }
    if (mods->has_property("flip_normals") && parse_bool(mods->property("flip_normals")))
        for (auto &x : m.n) x = -x;
    // extrude (extrude.cc:9-27): each vertex moves to 10 * its direction from
    // the origin * the red channel of the cubemap in that direction; the
    // cubemap's faces are read for this alone
    if (mods->has_group("extrude")) {
        const size_t keep = sc.textures.size();
        SamplerDesc cm;
        if (parse_sampler(sc, "cubemap", mods->group("extrude"), cm)) return 1;
        for (size_t i = 0; i + 2 < m.v.size(); i += 3) {
            const double x = m.v[i], y = m.v[i + 1], z = m.v[i + 2];
            const double l = std::sqrt(x * x + y * y + z * z);
            const double dx = l != 0 ? x / l : x, dy = l != 0 ? y / l : y, dz = l != 0 ? z / l : z;
            const double c = cubemap_red(sc, cm, x, y, z);
            m.v[i] = (float)(10. * dx * c);
            m.v[i + 1] = (float)(10. * dy * c);
            m.v[i + 2] = (float)(10. * dz * c);
        }
        sc.textures.resize(keep);
    }
    return 0;
}

// gen(icosahedron) (lib/nmesh/icosahedron.cc:10-70): the twelve vertices
// (+-1, +-k, 0) and their cyclic permutations, k the golden ratio rounded to
// float, normalised in FP64 and stored as float; the twenty faces of the
// icosahedron in the generator's order. The generator leaves the faces'
// normal and texcoord indices unset; here they are absent.
void gen_icosahedron(MeshData &m) {
    const double k = (float)((1.f + std::sqrt(5.0)) / 2.f);
    const double c[12][3] = {{-1, k, 0}, {1, k, 0}, {-1, -k, 0}, {1, -k, 0}, {0, -1, k}, {0, 1, k},
                             {0, -1, -k}, {0, 1, -k}, {k, 0, -1}, {k, 0, 1}, {-k, 0, -1}, {-k, 0, 1}};
    for (const auto &v : c) {
        const double len = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        for (int a = 0; a < 3; ++a) m.v.push_back((float)(v[a] / len));
    }
    static const int32_t faces[60] = {0, 11, 5, 0, 5, 1,  0, 1, 7,   0, 7, 10, 0, 10, 11, 1, 5, 9,  5, 11, 4,
                                      11, 10, 2, 10, 7, 6, 7, 1, 8,  3, 9, 4,  3, 4, 2,   3, 2, 6,  3, 6, 8,
                                      3, 8, 9,  4, 9, 5,  2, 4, 11,  6, 2, 10, 8, 6, 7,   9, 8, 1};
    m.iv.assign(faces, faces + 60);
    m.in.assign(60, -1);
    m.it.assign(60, -1);
}

// gen(plane) (lib/nmesh/plane.cc:11-60): a unit square in y = 0 centred on
// the origin, (res + 1)^2 vertices stepped by 1/res in FP64, normal +y,
// texcoords x + 0.5, z + 0.5; two triangles per cell (v0, v1, v2), (v0, v2,
// v3) with v1 the next row.
void gen_plane(MeshData &m, size_t res) {
    double px = -0.5, pz = -0.5;
    for (size_t i = 0; i < res + 1; ++i) {
        for (size_t j = 0; j < res + 1; ++j) {
            m.v.insert(m.v.end(), {(float)px, 0.f, (float)pz});
            m.n.insert(m.n.end(), {0.f, 1.f, 0.f});
            m.uv.insert(m.uv.end(), {(float)(px + 0.5), (float)(pz + 0.5)});
            px += 1.0 / res;
        }
        px = -0.5;
        pz += 1.0 / res;
    }
    for (size_t i = 0; i < res; ++i) {
        for (size_t j = 0; j < res; ++j) {
            const int32_t v0 = (int32_t)(j + i * (res + 1)), v1 = v0 + (int32_t)res + 1, v2 = v0 + (int32_t)res + 2,
                          v3 = v0 + 1;
            for (int32_t v : {v0, v1, v2, v0, v2, v3}) {
                m.iv.push_back(v);
                m.in.push_back(v);
                m.it.push_back(v);
            }
        }
    }
}

int load_geometry(Scene &sc, ncf::Node *p) {
    const char *type_c = p->property("type");
    std::string type = type_c ? type_c : "";
    Geometry g;
    g.name = p->name();
    g.handle = sc.intern(p->name());
    if (type == "plane") {
        g.type = GEOM_PLANE;
        g.normal = vec3_of(p, "normal");
        g.offset = numf(p->property("distance"));
    } else if (type == "sphere" || type == "point") {
        g.type = GEOM_SPHERE;
        g.center = vec3_of(p, "position");
        g.radius = type == "point" ? 0.0 : numf(p->property("radius"));
    } else if (type == "triangle") {
        g.type = GEOM_TRIANGLE;
        ncf::Node *vd = p->group("vecdata");
        const char *names[3] = {"v0", "v1", "v2"};
        for (int k = 0; k < 3; ++k) {
            g.v[k] = vec3_of(vd, names[k]);
            double u = 0, v = 0;
            if (const ncf::Node *tg = vd->find_group(names[k])) {
                const char *uu = tg->property("u"), *vv = tg->property("v");
                u = uu ? std::strtod(uu, nullptr) : 0.0;
                v = vv ? std::strtod(vv, nullptr) : 0.0;
            } else if (const char *val = vd->property(names[k])) {
                float a, b;
                if (std::sscanf(val, "tex2(%f,%f)", &a, &b) == 2) {
                    u = a;
                    v = b;
                }
            }
            g.tc[k][0] = u;
            g.tc[k][1] = v;
        }
    } else if (type == "mesh") {
        g.type = GEOM_MESH;
        const char *src = p->property("source");
        std::string f = src ? src : "";
        char gen[200] = {0};
        if (std::sscanf(f.c_str(), "gen(%199[a-zA-Z])", gen) > 0) {
            // procedural meshes (parseutil.cc:295-314); an unknown generator
            // leaves the mesh empty
            if (!std::strcmp(gen, "icosahedron")) {
                gen_icosahedron(g.mesh);
            } else if (!std::strcmp(gen, "plane")) {
                const char *r = p->property("resolution");
                long res = r ? std::strtol(r, nullptr, 0) : 0;
                gen_plane(g.mesh, res < 1 ? 1 : (size_t)res);
            }
        } else {
            std::string err;
            if (read_obj(dir_of(sc.source) + f, g.mesh, &err)) {
                sc.error = "failed to load mesh " + f + ": " + err;
                return 1;
            }
        }
        if (p->has_group("modifiers"))
            if (apply_modifiers(sc, g.mesh, p->group("modifiers"))) return 1;
    } else {
        sc.error = "unsupported geometry type " + type + " for " + p->name();
        return 1;
    }
    g.u_scale = numf(p->property("u_scale"));
    g.v_scale = numf(p->property("v_scale"));
    sc.geometries[g.handle] = std::move(g);
    return 0;
}

// Loads texture `file` (relative to the scene's directory) into sc.textures,
// flipped as the texture sampler's flip_x / flip_y ask (parseutil.cc:472-500,
// nimg/transform.cc); returns its index or -1 with sc.error set.
int32_t load_texture_at(Scene &sc, const std::string &path, bool flip_x, bool flip_y) {
    Image im;
    std::string err;
    if (load_image(path, im, &err)) {
        sc.error = "texture " + path + ": " + err;
        return -1;
    }
    const uint32_t w = im.width, h = im.height;
    if (flip_x)
        for (uint32_t y = 0; y < h; ++y)
            for (uint32_t x = 0; x < w / 2; ++x)
                for (int c = 0; c < 4; ++c)
                    std::swap(im.rgba[4 * ((size_t)y * w + x) + c], im.rgba[4 * ((size_t)y * w + w - 1 - x) + c]);
    if (flip_y)
        for (uint32_t y = 0; y < h / 2; ++y)
            for (uint32_t x = 0; x < 4 * w; ++x)
                std::swap(im.rgba[(size_t)y * 4 * w + x], im.rgba[(size_t)(h - 1 - y) * 4 * w + x]);
    sc.textures.push_back(std::move(im));
    return (int32_t)sc.textures.size() - 1;
}

int32_t load_texture(Scene &sc, const std::string &file, bool flip_x, bool flip_y) {
    return load_texture_at(sc, dir_of(sc.source) + file, flip_x, flip_y);
}

// deserialize_{rgba,gradient,texture,cubemap,erp} (parseutil.cc:147-203,
// 472-500) into a SamplerDesc; `cfg` holds the sampler's properties.
int parse_sampler(Scene &sc, const std::string &type, ncf::Node *cfg, SamplerDesc &out) {
    out = SamplerDesc{};
    if (type == "color") {
        out.kind = SMP_COLOR;
        col3_of(cfg, "value", out.a);
    } else if (type == "gradient") {
        out.kind = SMP_GRADIENT;
        col3_of(cfg, "a", out.a);
        col3_of(cfg, "b", out.b);
    } else if (type == "texture") {
        out.kind = SMP_TEXTURE;
        const char *src = cfg->property("source"), *flt = cfg->property("filtering");
        out.bilinear = flt && std::string(flt) == "bilinear" ? 1u : 0u;
        out.tex[0] = load_texture(sc, src ? src : "", parse_bool(cfg->property("flip_x")),
                                  parse_bool(cfg->property("flip_y")));
        if (out.tex[0] < 0) return 1;
    } else if (type == "erp") {
        out.kind = SMP_ERP;
        const char *src = cfg->property("source");
        out.tex[0] = load_texture(sc, src ? src : "", false, false);
        if (out.tex[0] < 0) return 1;
    } else if (type == "cubemap") {
        out.kind = SMP_CUBEMAP;
        // posx right, posy top, posz front, negx left, negy bottom, negz back
        static const char *names[6] = {"negx", "posx", "posy", "negy", "posz", "negz"};
        for (int f = 0; f < 6; ++f) {
            const char *src = cfg->property(names[f]);
            out.tex[f] = load_texture(sc, src ? src : "", false, false);
            if (out.tex[f] < 0) return 1;
        }
    }
    return 0;
}

int load_material(Scene &sc, ncf::Node *p) {
    const char *type_c = p->property("type");
    std::string type = type_c ? type_c : "";
    Material m;
    m.name = p->name();
    m.handle = sc.intern(p->name());
    if (type == "lambert") m.type = MAT_LAMBERT;
    else if (type == "phong") m.type = MAT_PHONG;
    else if (type == "blinn_phong") m.type = MAT_BLINNPHONG;
    else if (type == "emissive") m.type = MAT_EMISSIVE;
    else if (type == "dielectric") m.type = MAT_DIELECTRIC;
    else {
        sc.error = "unsupported material type " + type + " for " + p->name();
        return 1;
    }
    ncf::Node *props = p->group("properties");
    ncf::Node *samplers = props->group("samplers");
    ncf::Node *scalars = props->group("scalars");
    for (size_t i = 0; i < samplers->count_groups(); ++i) {
        ncf::Node *e = samplers->group_at(i);
        const std::string &nm = samplers->group_name_at(i);
        const char *st = e->property("type");
        std::string stype = st ? st : "";
        float *dst = nm == "diffuse" ? m.diffuse : nm == "specular" ? m.specular : nm == "emissive" ? m.emissive : nullptr;
        const int slot = nm == "diffuse" ? 0 : nm == "specular" ? 1 : nm == "emissive" ? 2 : -1;
        if (nm == "emissive") m.is_emissive = true;
        if (stype == "color") {
            float c[3] = {0, 0, 0};
            col3_of(e, "value", c);
            if (dst) std::memcpy(dst, c, sizeof c);
        }
        if (slot >= 0 && parse_sampler(sc, stype, e, m.smp[slot])) return 1;
    }
    for (size_t i = 0; i < scalars->count_properties(); ++i) {
        const std::string &nm = scalars->property_name_at(i);
        float v = (float)std::strtod(scalars->property_at(i).c_str(), nullptr);
        if (nm == "reflectance") m.reflectance = v;
        else if (nm == "exponent") m.exponent = v;
        else if (nm == "ior") m.ior = v;
        else if (nm == "emissive") m.is_emissive = true;
    }
    sc.materials[m.handle] = m;
    return 0;
}

// A material channel of an OBJ's MTL material (create_sampler,
// parseutil.cc:566-586): its texture, relative to the OBJ's directory with
// '\\' read as '/', mirrored left to right; or its colour.
int mtl_channel(Scene &sc, const std::string &dir, const std::string &tex, const float col[3], float dst[3],
                SamplerDesc &smp) {
    smp = SamplerDesc{};
    if (!tex.empty()) {
        std::string path = dir + tex;
        std::replace(path.begin(), path.end(), '\\', '/');
        smp.kind = SMP_TEXTURE;
        smp.tex[0] = load_texture_at(sc, path, true, false);
        for (int k = 0; k < 3; ++k) dst[k] = 0.f;
        return smp.tex[0] < 0 ? 1 : 0;
    }
    smp.kind = SMP_COLOR;
    for (int k = 0; k < 3; ++k) smp.a[k] = dst[k] = col[k] < 0.f ? 0.f : col[k];
    return 0;
}

// One shape of an OBJ file as a mesh: only the attributes its triangles use,
// in first-use order, with the indices remapped (the values are unchanged).
void shape_mesh(const ObjFile &o, const ObjShape &sh, MeshData &m) {
    auto remap = [](const std::vector<int32_t> &idx, const std::vector<float> &src, int width,
                    std::vector<float> &dst_vals, std::vector<int32_t> &dst_idx) {
        std::unordered_map<int32_t, int32_t> seen;
        const int32_t count = (int32_t)(src.size() / width);
        dst_idx.reserve(idx.size());
        for (int32_t i : idx) {
            if (i < 0 || i >= count) {
                dst_idx.push_back(-1);
                continue;
            }
            auto it = seen.find(i);
            if (it == seen.end()) {
                it = seen.emplace(i, (int32_t)(dst_vals.size() / width)).first;
                dst_vals.insert(dst_vals.end(), src.begin() + (size_t)i * width, src.begin() + (size_t)(i + 1) * width);
            }
            dst_idx.push_back(it->second);
        }
    };
    remap(sh.iv, o.v, 3, m.v, m.iv);
    remap(sh.in, o.n, 3, m.n, m.in);
    remap(sh.it, o.uv, 2, m.uv, m.it);
}

// An object imported from an OBJ file with its MTL materials
// (create_object(Scene*, filepath, prefix), parseutil.cc:588-682). Each MTL
// material becomes `prefix + name`: emissive when it has an emission colour
// above 0 or an emission texture, Blinn-Phong otherwise (diffuse, specular,
// ior = Ni, exponent = Ns); a repeated name keeps its first material. Each
// shape k (from 1) becomes geometry and object `k + prefix + shape name`
// with the material of its first face, taken, as the reference does, by
// the face's material index among the distinct names. A file that cannot be
// read adds nothing (the reference logs a warning and goes on).
int load_external_object(Scene &sc, ncf::Node *p) {
    const char *src = p->property("source"), *pre = p->property("prefix");
    const std::string path = dir_of(sc.source) + (src ? src : "");
    const std::string prefix = pre ? pre : "";
    const std::string dir = dir_of(path);
    ObjFile o;
    std::string err;
    if (read_obj_file(path, dir, o, &err)) return 0;
    std::vector<uint64_t> matids;
    for (const ObjMaterial &om : o.materials) {
        const std::string name = prefix + om.name;
        const uint64_t id = sc.intern(name);
        if (std::find(matids.begin(), matids.end(), id) != matids.end()) continue;
        matids.push_back(id);
        Material m;
        m.name = name;
        m.handle = id;
        const bool emissive = om.ke[0] > 0.f || om.ke[1] > 0.f || om.ke[2] > 0.f || !om.map_ke.empty();
        if (emissive) {
            m.type = MAT_EMISSIVE;
            m.is_emissive = true;
            if (mtl_channel(sc, dir, om.map_ke, om.ke, m.emissive, m.smp[2])) return 1;
        } else {
            m.type = MAT_BLINNPHONG;
            if (mtl_channel(sc, dir, om.map_kd, om.kd, m.diffuse, m.smp[0])) return 1;
            if (mtl_channel(sc, dir, om.map_ks, om.ks, m.specular, m.smp[1])) return 1;
            m.ior = om.ior;
            m.exponent = om.shininess;
        }
        sc.materials[id] = m;
    }
    int count = 0;
    for (const ObjShape &sh : o.shapes) {
        ++count;
        const std::string name = std::to_string(count) + prefix + sh.name;
        const uint64_t id = sc.intern(name);
        const int mi = sh.mat.empty() ? -1 : sh.mat[0];
        if (mi < 0 || (size_t)mi >= matids.size()) {
            sc.error = "shape " + name + " of " + path + " has no material";
            return 1;
        }
        Geometry g;
        g.name = name;
        g.handle = id;
        g.type = GEOM_MESH;
        shape_mesh(o, sh, g.mesh);
        sc.geometries[id] = std::move(g);
        Object ob;
        ob.name = name;
        ob.handle = id;
        ob.geometry = id;
        ob.material = matids[mi];
        sc.objects[id] = ob;
    }
    return 0;
}

int load_object(Scene &sc, ncf::Node *p) {
    if (p->has_property("source")) return load_external_object(sc, p);
    Object o;
    o.name = p->name();
    o.handle = sc.intern(p->name());
    const char *g = p->property("geometry");
    const char *m = p->property("material");
    o.geometry = sc.intern(g ? g : "");
    o.material = sc.intern(m ? m : "");
    sc.objects[o.handle] = o;
    return 0;
}

}  // namespace

int load_scn(Scene &sc, const std::string &path, const std::vector<std::string> &mods) {
    ncf::Node root;
    size_t line = 0;
    int rc = root.parse(path, &line);
    if (rc != ncf::OK) {
        sc.error = "failed to parse the scene, line " + std::to_string(line) + ": " + ncf::error_message(rc);
        return LOAD_PARSE;
    }
    for (std::string mod : mods) {
        if (mod.find_last_of(':') == std::string::npos) continue;
        ncf::Node *node = &root;
        while (mod.find('.') != std::string::npos && mod.find(':') > mod.find('.')) {
            size_t d = mod.find('.');
            node = node->group(mod.substr(0, d));
            mod = mod.substr(d + 1);
        }
        size_t c = mod.find(':');
        node->set_property(mod.substr(0, c), mod.substr(c + 1));
    }
    sc.source = path;
    const char *t = root.property("title");
    const char *d = root.property("description");
    const char *v = root.property("version");
    sc.title = t ? t : "";
    sc.description = d ? d : "";
    sc.version = v ? v : "";

    ncf::Node *env = root.group("environment");
    ncf::Node *cfg = env->group("config");
    const char *et = env->property("type");
    std::string etype = et ? et : "";
    if (etype == "gradient") {
        sc.env.type = ENV_GRADIENT;
        col3_of(cfg, "a", sc.env.a);
        col3_of(cfg, "b", sc.env.b);
    } else if (etype == "color") {
        sc.env.type = ENV_COLOR;
        col3_of(cfg, "value", sc.env.a);
    } else if (etype == "cubemap" || etype == "erp") {
        sc.env.type = etype == "cubemap" ? ENV_CUBEMAP : ENV_ERP;
        if (parse_sampler(sc, etype, cfg, sc.env.smp)) return 1;
    }

    const char *sections[] = {"camera", "geometry", "material", "texture", "object"};
    for (const char *sec : sections) {
        ncf::Node *s = root.group(sec);
        for (size_t i = 0; i < s->count_groups(); ++i) {
            ncf::Node *n = s->group_at(i);
            int r = 0;
            if (!std::strcmp(sec, "camera")) r = load_camera(sc, n);
            else if (!std::strcmp(sec, "geometry")) r = load_geometry(sc, n);
            else if (!std::strcmp(sec, "material")) r = load_material(sc, n);
            else if (!std::strcmp(sec, "object")) r = load_object(sc, n);
            if (r) {
                if (sc.error.empty()) sc.error = std::string("failed to load ") + n->name();
                return LOAD_FAIL;
            }
        }
    }
    return LOAD_OK;
}

namespace {
// Parses "v", "v/vt", "v//vn" or "v/vt/vn" into 0-based indices (tinyobj's
// parseTriple / fixIndex: negative indices count from the end, 0 reads 0).
bool parse_corner(const char *s, int nv, int nt, int nn, int out[3]) {
    int vals[3] = {0, 0, 0};
    bool has[3] = {false, false, false};
    int k = 0;
    const char *p = s;
    while (*p && k < 3) {
        if (*p == '/') {
            ++k;
            ++p;
            continue;
        }
        char *end;
        long x = std::strtol(p, &end, 10);
        if (end == p) return false;
        vals[k] = (int)x;
        has[k] = true;
        p = end;
    }
    const int counts[3] = {nv, nt, nn};
    for (int i = 0; i < 3; ++i) {
        if (!has[i]) {
            out[i] = -1;
            continue;
        }
        const int x = vals[i];
        out[i] = x > 0 ? x - 1 : (x == 0 ? 0 : counts[i] + x);
    }
    return has[0];
}

bool is_space(char c) { return c == ' ' || c == '\t'; }

// The line without its trailing newline, and the offset of its first token.
const char *line_token(std::string &line) {
    while (!line.empty() && (line.back() == '\n' || line.back() == '\r')) line.pop_back();
    const char *t = line.c_str();
    t += std::strspn(t, " \t");
    return t;
}

// First whitespace-delimited word after `t` (sscanf "%s").
std::string first_word(const char *t) {
    t += std::strspn(t, " \t\r\n");
    return std::string(t, std::strcspn(t, " \t\r\n"));
}

void read3(const char *t, float out[3]) {
    out[0] = out[1] = out[2] = 0.f;
    std::sscanf(t, "%f %f %f", &out[0], &out[1], &out[2]);
}

// LoadMtl (tinyobjloader 1.0): `newmtl` starts a material and stores the
// previous one when it has a name; the last one is stored at the end in any
// case, so an unreadable file still yields one unnamed default material.
// The first material of a name is the one `usemtl` finds.
void read_mtl(const std::string &path, std::vector<ObjMaterial> &mats, std::map<std::string, int> &index) {
    std::ifstream in(path);
    ObjMaterial m;
    std::string line;
    auto flush = [&]() {
        index.insert({m.name, (int)mats.size()});
        mats.push_back(m);
    };
    while (in.good() && std::getline(in, line)) {
        while (!line.empty() && (line.back() == ' ' || line.back() == '\t')) line.pop_back();
        const char *t = line_token(line);
        if (!*t || *t == '#') continue;
        if (!std::strncmp(t, "newmtl", 6) && is_space(t[6])) {
            if (!m.name.empty()) flush();
            m = ObjMaterial();
            m.name = first_word(t + 7);
        } else if (t[0] == 'K' && t[1] == 'd' && is_space(t[2])) {
            read3(t + 2, m.kd);
        } else if (t[0] == 'K' && t[1] == 's' && is_space(t[2])) {
            read3(t + 2, m.ks);
        } else if (t[0] == 'K' && t[1] == 'e' && is_space(t[2])) {
            read3(t + 2, m.ke);
        } else if (t[0] == 'N' && t[1] == 'i' && is_space(t[2])) {
            m.ior = (float)std::strtod(t + 2, nullptr);
        } else if (t[0] == 'N' && t[1] == 's' && is_space(t[2])) {
            m.shininess = (float)std::strtod(t + 2, nullptr);
        } else if (!std::strncmp(t, "map_Kd", 6) && is_space(t[6])) {
            m.map_kd = t + 7;
        } else if (!std::strncmp(t, "map_Ks", 6) && is_space(t[6])) {
            m.map_ks = t + 7;
        } else if (!std::strncmp(t, "map_Ke", 6) && is_space(t[6])) {
            m.map_ke = t + 7;
        }
    }
    flush();
}
}  // namespace

int read_obj_file(const std::string &path, const std::string &mtl_dir, ObjFile &o, std::string *err) {
    std::ifstream in(path);
    if (!in.good()) {
        if (err) *err = "cannot open " + path;
        return 1;
    }
    std::map<std::string, int> mat_index;
    std::vector<std::vector<int>> group;  // faces not yet in a shape: 3 ints per corner
    std::string name;
    int material = -1;
    ObjShape shape;
    // exportFaceGroupToShape: the pending faces, fan-triangulated, join the
    // shape with the current material; false when there were none
    auto flush_faces = [&]() -> bool {
        if (group.empty()) return false;
        for (const std::vector<int> &f : group) {
            const size_t corners = f.size() / 3;
            for (size_t k = 2; k < corners; ++k) {
                const size_t idx[3] = {0, k - 1, k};
                for (size_t c : idx) {
                    shape.iv.push_back(f[3 * c]);
                    shape.it.push_back(f[3 * c + 1]);
                    shape.in.push_back(f[3 * c + 2]);
                }
                shape.mat.push_back(material);
            }
        }
        shape.name = name;
        return true;
    };
    // a `g` / `o` line or the end: the shape is kept only when faces were
    // pending (faces an earlier `usemtl` moved into it are dropped otherwise)
    auto end_shape = [&]() {
        if (flush_faces()) o.shapes.push_back(shape);
        shape = ObjShape();
        group.clear();
    };
    std::string line;
    std::vector<int> face;
    while (std::getline(in, line)) {
        const char *s = line_token(line);
        if (!*s || *s == '#') continue;
        if (s[0] == 'v' && is_space(s[1])) {
            float x = 0, y = 0, z = 0;
            std::sscanf(s + 2, "%f %f %f", &x, &y, &z);
            o.v.insert(o.v.end(), {x, y, z});
        } else if (s[0] == 'v' && s[1] == 'n' && is_space(s[2])) {
            float x = 0, y = 0, z = 0;
            std::sscanf(s + 3, "%f %f %f", &x, &y, &z);
            o.n.insert(o.n.end(), {x, y, z});
        } else if (s[0] == 'v' && s[1] == 't' && is_space(s[2])) {
            float u = 0, v = 0;
            std::sscanf(s + 3, "%f %f", &u, &v);
            o.uv.insert(o.uv.end(), {u, v});
        } else if (s[0] == 'f' && is_space(s[1])) {
            std::istringstream ss(s + 2);
            std::string tok;
            face.clear();
            const int nv = (int)o.v.size() / 3, nt = (int)o.uv.size() / 2, nn = (int)o.n.size() / 3;
            while (ss >> tok) {
                int c[3];
                if (!parse_corner(tok.c_str(), nv, nt, nn, c)) {
                    if (err) *err = "bad face: " + line;
                    return 1;
                }
                face.insert(face.end(), {c[0], c[1], c[2]});
            }
            // two corners make no triangle (but count as a pending face);
            // fewer have no meaning
            if (face.size() < 6) {
                if (err) *err = "face with fewer than two corners: " + line;
                return 1;
            }
            group.push_back(face);
        } else if (!std::strncmp(s, "usemtl", 6) && is_space(s[6])) {
            auto it = mat_index.find(first_word(s + 7));
            const int id = it == mat_index.end() ? -1 : it->second;
            if (id != material) {
                flush_faces();
                group.clear();
                material = id;
            }
        } else if (!std::strncmp(s, "mtllib", 6) && is_space(s[6])) {
            read_mtl(mtl_dir + first_word(s + 7), o.materials, mat_index);
        } else if (s[0] == 'g' && is_space(s[1])) {
            end_shape();
            std::istringstream ss(s + 2);
            std::string w;
            name = (ss >> w) ? w : "";
        } else if (s[0] == 'o' && is_space(s[1])) {
            end_shape();
            name = first_word(s + 2);
        }
    }
    end_shape();
    return 0;
}

int read_obj(const std::string &path, MeshData &m, std::string *err) {
    ObjFile o;
    if (read_obj_file(path, "", o, err)) return 1;
    m.v = std::move(o.v);
    m.n = std::move(o.n);
    m.uv = std::move(o.uv);
    for (const ObjShape &sh : o.shapes) {
        m.iv.insert(m.iv.end(), sh.iv.begin(), sh.iv.end());
        m.it.insert(m.it.end(), sh.it.begin(), sh.it.end());
        m.in.insert(m.in.end(), sh.in.begin(), sh.in.end());
    }
    // an index past its array (or before it, from a negative index) reads
    // as absent, as the imported-object path reads it: vertex (0, 0, 0),
    // Triangle()'s default texcoord, no vertex normal
    auto bound = [](std::vector<int> &idx, size_t count) {
        for (int &i : idx)
            if (i < 0 || (size_t)i >= count) i = -1;
    };
    bound(m.iv, m.v.size() / 3);
    bound(m.it, m.uv.size() / 2);
    bound(m.in, m.n.size() / 3);
    return 0;
}

}  // namespace xt
