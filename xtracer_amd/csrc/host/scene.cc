#include "scene.h"

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
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
    if (mods->has_group("extrude")) {
        sc.error = "the extrude mesh modifier is not supported";
        return 1;
    }
    return 0;
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
            sc.error = "procedural mesh generators are not supported: " + f;
            return 1;
        }
        std::string err;
        if (read_obj(dir_of(sc.source) + f, g.mesh, &err)) {
            sc.error = "failed to load mesh " + f + ": " + err;
            return 1;
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
int32_t load_texture(Scene &sc, const std::string &file, bool flip_x, bool flip_y) {
    Image im;
    std::string err;
    if (load_image(dir_of(sc.source) + file, im, &err)) {
        sc.error = "texture " + file + ": " + err;
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

int load_object(Scene &sc, ncf::Node *p) {
    if (p->has_property("source")) {
        sc.error = "external object sources are not supported: " + p->name();
        return 1;
    }
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
// Parses "v", "v/vt", "v//vn" or "v/vt/vn" into 0-based indices.
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
        int x = vals[i];
        out[i] = x > 0 ? x - 1 : counts[i] + x;
    }
    return has[0];
}
}  // namespace

int read_obj(const std::string &path, MeshData &m, std::string *err) {
    std::ifstream in(path);
    if (!in.good()) {
        if (err) *err = "cannot open " + path;
        return 1;
    }
    std::string line;
    std::vector<int> face;
    while (std::getline(in, line)) {
        const char *s = line.c_str();
        while (*s == ' ' || *s == '\t') ++s;
        if (s[0] == 'v' && (s[1] == ' ' || s[1] == '\t')) {
            float x = 0, y = 0, z = 0;
            std::sscanf(s + 2, "%f %f %f", &x, &y, &z);
            m.v.insert(m.v.end(), {x, y, z});
        } else if (s[0] == 'v' && s[1] == 'n') {
            float x = 0, y = 0, z = 0;
            std::sscanf(s + 3, "%f %f %f", &x, &y, &z);
            m.n.insert(m.n.end(), {x, y, z});
        } else if (s[0] == 'v' && s[1] == 't') {
            float u = 0, v = 0;
            std::sscanf(s + 3, "%f %f", &u, &v);
            m.uv.insert(m.uv.end(), {u, v});
        } else if (s[0] == 'f' && (s[1] == ' ' || s[1] == '\t')) {
            std::istringstream ss(s + 2);
            std::string tok;
            face.clear();
            const int nv = (int)m.v.size() / 3, nt = (int)m.uv.size() / 2, nn = (int)m.n.size() / 3;
            while (ss >> tok) {
                int c[3];
                if (!parse_corner(tok.c_str(), nv, nt, nn, c)) {
                    if (err) *err = "bad face: " + line;
                    return 1;
                }
                face.insert(face.end(), {c[0], c[1], c[2]});
            }
            const size_t corners = face.size() / 3;
            for (size_t k = 2; k < corners; ++k) {  // fan triangulation, as tinyobj does
                const size_t idx[3] = {0, k - 1, k};
                for (size_t j : idx) {
                    m.iv.push_back(face[3 * j]);
                    m.it.push_back(face[3 * j + 1]);
                    m.in.push_back(face[3 * j + 2]);
                }
            }
        }
    }
    return 0;
}

}  // namespace xt
