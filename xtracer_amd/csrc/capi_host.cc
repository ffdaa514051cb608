// Host half of the C ABI: scene loading, flattening and tree builds, and the
// FP64 camera set-up. Needs no GPU.
#include <cmath>
#include <cstring>
#include <exception>
#include <new>

#include "../../include/xtracer_amd.h"
#include "capi_internal.h"
#include "host/image_load.h"
#include "host/image_io.h"

using namespace xt;

extern "C" {

int xt_scene_load(const char *scn_path, const char *const *mods, int n_mods, xt_scene **out) {
    if (!scn_path || !out || n_mods < 0) return XT_E_INVALID;
    *out = nullptr;
    xt_scene *s = new (std::nothrow) xt_scene;
    if (!s) return XT_E_INVALID;
    std::vector<std::string> m;
    for (int i = 0; i < n_mods; ++i)
        if (mods && mods[i]) m.push_back(mods[i]);
    int rc;
    try {
        rc = load_scn(s->scene, scn_path, m);
    } catch (const std::exception &e) {
        s->scene.error = e.what();
        rc = LOAD_FAIL;
    }
    s->error = s->scene.error;
    *out = s;  // returned even on failure so that xt_scene_error can be read
    if (rc == LOAD_PARSE) return XT_E_IO;
    if (rc != LOAD_OK) return s->error.find("failed to load mesh") != std::string::npos ? XT_E_IO : XT_E_SCENE;
    return XT_OK;
}

int xt_scene_build(xt_scene *s, const xt_build_params *p) {
    if (!s) return XT_E_INVALID;
    KdParams kp;
    if (p) {
        if (p->kd_max_depth > 0) kp.max_depth = p->kd_max_depth;
        if (p->kd_leaf_prims > 0) kp.leaf_prims = p->kd_leaf_prims;
        if (p->kd_cost_intersect > 0.f) kp.cost_intersect = p->kd_cost_intersect;
        if (p->threads > 0) kp.threads = p->threads;
        kp.on_device = p->kd_on_device == 1;
    }
    try {
        std::string err;
        if (flatten_scene(s->scene, s->flat, kp, &err)) {
            s->error = err;
            return XT_E_SCENE;
        }
    } catch (const std::exception &e) {
        s->error = e.what();
        return XT_E_SCENE;
    }
    s->built = true;
    return XT_OK;
}

void xt_scene_free(xt_scene *s) { delete s; }

const char *xt_scene_error(const xt_scene *s) { return s ? s->error.c_str() : "null scene"; }

int xt_scene_get_info(const xt_scene *s, xt_scene_info *info) {
    if (!s || !info) return XT_E_INVALID;
    std::memset(info, 0, sizeof *info);
    const FlatScene &f = s->flat;
    info->n_objects = (uint32_t)(s->built ? f.objects.size() : s->scene.objects.size());
    info->n_prims = (uint32_t)f.prims.size();
    info->n_planes = (uint32_t)f.planes.size();
    info->n_materials = (uint32_t)(s->built ? f.materials.size() : s->scene.materials.size());
    info->n_cameras = (uint32_t)s->scene.cameras.size();
    info->n_kd_nodes = (uint32_t)f.kd.nodes.size();
    info->n_kd_refs = (uint32_t)f.kd.refs.size();
    info->kd_depth = (uint32_t)f.kd.depth;
    info->n_oct_nodes = (uint32_t)f.oct_nodes.size();
    info->n_oct_refs = (uint32_t)f.oct_refs.size();
    info->env_type = s->scene.env.type;
    info->any_vertex_normals = f.vnormals.empty() ? 0u : 1u;
    info->build_seconds_octree = f.build_seconds_octree;
    info->build_seconds_kd = f.build_seconds_kd;
    return XT_OK;
}

int xt_scene_get_object(const xt_scene *s, uint32_t index, xt_object_info *info) {
    if (!s || !info || !s->built) return XT_E_INVALID;
    if (index >= s->flat.objects.size()) return XT_E_NOTFOUND;
    const FlatObject &o = s->flat.objects[index];
    std::memset(info, 0, sizeof *info);
    std::strncpy(info->name, o.name.c_str(), sizeof(info->name) - 1);
    info->handle = o.handle;
    info->geom_type = o.geom_type;
    info->material = o.material;
    info->first_prim = o.first_prim;
    info->prim_count = o.prim_count;
    info->plane = o.plane;
    return XT_OK;
}

int xt_scene_get_arrays(const xt_scene *s, void *prims, double *prims64, uint32_t *prim_local, void *planes,
                        void *materials, float env[7]) {
    if (!s || !s->built) return XT_E_INVALID;
    const FlatScene &f = s->flat;
    if (prims && !f.prims.empty()) std::memcpy(prims, f.prims.data(), f.prims.size() * sizeof(XtPrim));
    if (prims64 && !f.prims64.empty()) std::memcpy(prims64, f.prims64.data(), f.prims64.size() * sizeof(XtPrim64));
    if (prim_local && !f.prim_local.empty())
        std::memcpy(prim_local, f.prim_local.data(), f.prim_local.size() * sizeof(uint32_t));
    if (planes && !f.planes.empty()) std::memcpy(planes, f.planes.data(), f.planes.size() * sizeof(XtPlane));
    if (materials && !f.materials.empty())
        std::memcpy(materials, f.materials.data(), f.materials.size() * sizeof(XtMaterial));
    if (env) {
        env[0] = (float)f.env.type;
        for (int i = 0; i < 3; ++i) {
            env[1 + i] = f.env.a[i];
            env[4 + i] = f.env.b[i];
        }
    }
    return XT_OK;
}

int xt_scene_get_object_params(const xt_scene *s, uint32_t index, double out[9]) {
    if (!s || !s->built || !out) return XT_E_INVALID;
    if (index >= s->flat.objects.size()) return XT_E_NOTFOUND;
    for (int k = 0; k < 9; ++k) out[k] = 0.0;
    const auto &obj = s->scene.objects.at(s->flat.objects[index].handle);
    auto g = s->scene.geometries.find(obj.geometry);
    if (g == s->scene.geometries.end()) return XT_OK;
    const Geometry &geo = g->second;
    if (geo.type == GEOM_PLANE) {
        out[0] = geo.normal.x;
        out[1] = geo.normal.y;
        out[2] = geo.normal.z;
        out[3] = geo.offset;
    } else if (geo.type == GEOM_SPHERE) {
        out[0] = geo.center.x;
        out[1] = geo.center.y;
        out[2] = geo.center.z;
        out[3] = geo.radius;
    } else if (geo.type == GEOM_TRIANGLE) {
        for (int k = 0; k < 3; ++k)
            for (int a = 0; a < 3; ++a) out[3 * k + a] = geo.v[k][a];
    }
    return XT_OK;
}

int xt_scene_get_vnormals(const xt_scene *s, float *out) {
    if (!s || !s->built || !out) return XT_E_INVALID;
    if (s->flat.vnormals.empty()) return XT_E_NOTFOUND;
    std::memcpy(out, s->flat.vnormals.data(), s->flat.vnormals.size() * sizeof(float));
    return XT_OK;
}

int xt_scene_get_texcoords(const xt_scene *s, double *prim_tc, double *uv_scale) {
    if (!s || !s->built) return XT_E_INVALID;
    const FlatScene &f = s->flat;
    if (prim_tc && !f.tc64.empty()) std::memcpy(prim_tc, f.tc64.data(), f.tc64.size() * sizeof(double));
    if (uv_scale && !f.obj_uv_scale.empty())
        std::memcpy(uv_scale, f.obj_uv_scale.data(), f.obj_uv_scale.size() * sizeof(double));
    return XT_OK;
}

int xt_scene_get_samplers(const xt_scene *s, uint32_t counts[3], uint32_t *tex_info, float *texels, uint32_t *samplers,
                          uint32_t *mat_smp, uint32_t *env_smp) {
    if (!s || !s->built || !counts) return XT_E_INVALID;
    static_assert(sizeof(XtSampler) == 14 * sizeof(uint32_t), "XtSampler is 14 words");
    const FlatScene &f = s->flat;
    counts[0] = (uint32_t)f.textures.size();
    counts[1] = (uint32_t)f.texels.size();
    counts[2] = (uint32_t)f.samplers.size();
    if (tex_info && !f.textures.empty()) std::memcpy(tex_info, f.textures.data(), f.textures.size() * sizeof(XtTexture));
    if (texels && !f.texels.empty()) std::memcpy(texels, f.texels.data(), f.texels.size() * sizeof(float));
    if (samplers && !f.samplers.empty()) std::memcpy(samplers, f.samplers.data(), f.samplers.size() * sizeof(XtSampler));
    if (mat_smp && !f.mat_smp.empty()) std::memcpy(mat_smp, f.mat_smp.data(), f.mat_smp.size() * sizeof(uint32_t));
    if (env_smp) *env_smp = f.env.smp;
    return XT_OK;
}

int xt_scene_get_kdtree(const xt_scene *s, uint32_t *nodes, uint32_t *refs, float bounds[6]) {
    if (!s || !s->built) return XT_E_INVALID;
    const KdTree &k = s->flat.kd;
    if (nodes && !k.nodes.empty()) std::memcpy(nodes, k.nodes.data(), k.nodes.size() * sizeof(XtKdNode));
    if (refs && !k.refs.empty()) std::memcpy(refs, k.refs.data(), k.refs.size() * sizeof(uint32_t));
    if (bounds)
        for (int a = 0; a < 3; ++a) {
            bounds[a] = k.bounds.lo[a];
            bounds[3 + a] = k.bounds.hi[a];
        }
    return XT_OK;
}

int xt_scene_mesh_octree(const xt_scene *s, uint32_t k, uint64_t *geometry, uint64_t *nodes, uint64_t *max_level,
                         uint64_t *triangles) {
    if (!s || !s->built) return XT_E_INVALID;
    if (k >= s->flat.mesh_octrees.size()) return XT_E_NOTFOUND;
    const auto &m = s->flat.mesh_octrees[k];
    if (geometry) *geometry = m.geometry;
    if (nodes) *nodes = m.nodes;
    if (max_level) *max_level = m.max_level;
    if (triangles) *triangles = m.triangles;
    return XT_OK;
}

// Perspective::calculate_transform (perspective.cc:26-36) and the projection
// constant 1 / tan(fov * RADIAN / 2) (perspective.cc:53), in FP64.
}  // extern "C"

namespace {
// Matrix4x4f m; m.rotate(euler) as ODS / ERP build it (matrix.cc:485-492,
// 414-427): Rx*Ry*Rz from double cos/sin, the 3x3 products summed in the
// reference's order, then identity *= R, whose sums turn -0 entries into +0.
void euler_rotation(const dvec3 &e, double out[9]) {
    typedef double M3[3][3];
    const M3 xr = {{1, 0, 0}, {0, std::cos(e.x), -std::sin(e.x)}, {0, std::sin(e.x), std::cos(e.x)}};
    const M3 yr = {{std::cos(e.y), 0, std::sin(e.y)}, {0, 1, 0}, {-std::sin(e.y), 0, std::cos(e.y)}};
    const M3 zr = {{std::cos(e.z), -std::sin(e.z), 0}, {std::sin(e.z), std::cos(e.z), 0}, {0, 0, 1}};
    auto mul = [](const M3 &a, const M3 &b, M3 &r) {
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) r[i][j] = a[i][0] * b[0][j] + a[i][1] * b[1][j] + a[i][2] * b[2][j];
    };
    M3 xy, r;
    mul(xr, yr, xy);
    mul(xy, zr, r);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {  // row i of the identity times column j of R as a 4x4
            double v = (i == 0 ? 1.0 : 0.0) * r[0][j];
            v = v + (i == 1 ? 1.0 : 0.0) * r[1][j];
            v = v + (i == 2 ? 1.0 : 0.0) * r[2][j];
            out[3 * i + j] = v + 0.0;  // the fourth term, 0 * 0
        }
}
}  // namespace

extern "C" {

int xt_camera_make(const double position[3], const double target[3], const double up[3], float fov_deg,
                   float aperture, float flength, xt_camera *out) {
    if (!position || !target || !up || !out) return XT_E_INVALID;
    std::memset(out, 0, sizeof *out);
    const dvec3 pos(position[0], position[1], position[2]);
    const dvec3 tg(target[0], target[1], target[2]);
    const dvec3 u(up[0], up[1], up[2]);
    const dvec3 rz = normalized(tg - pos);
    const dvec3 rx = normalized(cross(u, rz));
    const dvec3 ry = normalized(cross(rx, rz));
    const double m[9] = {rx.x, ry.x, rz.x, rx.y, ry.y, rz.y, rx.z, ry.z, rz.z};
    std::memcpy(out->basis, m, sizeof m);
    for (int i = 0; i < 3; ++i) out->position[i] = position[i];
    const double RADIAN = 0.017453292519943;  // lib/nmath/precision.h
    out->dir_z = 1.0 / std::tan((double)fov_deg * RADIAN / 2.0);
    out->aperture = aperture;
    out->flength = flength;
    out->thin_lens = flength > 0.f ? 1u : 0u;
    return XT_OK;
}

int xt_scene_camera(const xt_scene *s, const char *camera_name, xt_camera *out) {
    if (!s || !out) return XT_E_INVALID;
    const Camera *cam = nullptr;
    if (!camera_name || !*camera_name) {
        if (!s->scene.cameras.empty()) cam = &s->scene.cameras.begin()->second;
    } else {
        for (auto &kv : s->scene.cameras)
            if (kv.second.name == camera_name) cam = &kv.second;
    }
    if (!cam) return XT_E_NOTFOUND;
    const double p[3] = {cam->position.x, cam->position.y, cam->position.z};
    if (cam->kind == XT_CAMERA_THIN_LENS) {
        const double t[3] = {cam->target.x, cam->target.y, cam->target.z};
        const double u[3] = {cam->up.x, cam->up.y, cam->up.z};
        return xt_camera_make(p, t, u, cam->fov, cam->aperture, cam->flength, out);
    }
    std::memset(out, 0, sizeof *out);
    out->kind = cam->kind;
    for (int i = 0; i < 3; ++i) out->position[i] = p[i];
    out->ipd = cam->ipd;
    if (cam->kind == XT_CAMERA_ODS || cam->kind == XT_CAMERA_ERP) euler_rotation(cam->orientation, out->basis);
    return XT_OK;
}


int xt_image_write(const char *path, const float *rgba, uint32_t width, uint32_t height) {
    if (!path || !rgba || !width || !height) return XT_E_INVALID;
    const int rc = xt::write_image(path, rgba, width, height);
    return rc == 0 ? XT_OK : rc == 2 ? XT_E_INVALID : XT_E_IO;
}

int xt_image_read(const char *path, float *rgba, uint32_t *width, uint32_t *height) {
    if (!path || !width || !height) return XT_E_INVALID;
    xt::Image img;
    int rc;
    try {
        rc = xt::load_image(path, img);
    } catch (const std::exception &) {  // e.g. a header declaring more pixels than memory holds
        return XT_E_INVALID;
    }
    if (rc) return rc == 1 ? XT_E_IO : XT_E_INVALID;
    *width = img.width;
    *height = img.height;
    if (rgba) std::memcpy(rgba, img.rgba.data(), img.rgba.size() * sizeof(float));
    return XT_OK;
}

}  // extern "C"
