// Host-side scene model and .scn loader.
//
// The loader mirrors xtcore::io::scn::load (src/xtcore/parseutil.cc:697-822):
// sections are visited camera -> geometry -> material -> texture -> object,
// each in alphabetical group order, and every asset name is registered in a
// first-seen order that stands in for the string-pool handles the reference
// keys its asset maps with (src/xtcore/strpool.cc:27-34). Objects are kept
// sorted by that order, which is the order Scene::intersection scans them in
// (src/xtcore/scene.cc:191-216) and therefore the tie-break order of hits.
#pragma once
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "image_load.h"

#include "xt_vec.h"

namespace xt {

enum LoadStatus { LOAD_OK = 0, LOAD_FAIL = 1, LOAD_PARSE = 2 };

struct Camera {
    std::string name;
    uint64_t handle = 0;
    uint32_t kind = 0;  // XT_CAMERA_*: thin-lens, ods, erp, cubemap
    dvec3 position, target{0, 0, 1}, up{0, 1, 0};
    float fov = 0.f, aperture = 0.f, flength = 0.f;  // float members, perspective.h:22-26
    dvec3 orientation;                               // ods / erp euler angles (ods.h:18-19, erp.h:17)
    float ipd = 0.f;                                 // ods
};

enum GeomType : uint32_t { GEOM_PLANE = 0, GEOM_SPHERE = 1, GEOM_TRIANGLE = 2, GEOM_MESH = 3 };

// Triangle soup of a mesh after modifiers; coordinates are float, as in
// nmesh::attrib_t (lib/nmesh/structs.h:56-60).
struct MeshData {
    std::vector<float> v, n, uv;      // flat xyz / xyz / uv arrays
    std::vector<int32_t> iv, in, it;  // 3 per triangle, -1 when absent
    size_t tri_count() const { return iv.size() / 3; }
};

struct Geometry {
    std::string name;
    uint64_t handle = 0;
    GeomType type = GEOM_PLANE;
    dvec3 normal;          // plane
    double offset = 0;     // plane "distance"
    dvec3 center;          // sphere / point
    double radius = 0;     // sphere (0 for a point)
    dvec3 v[3];            // triangle
    double tc[3][2] = {{0, 0}, {1, 0}, {0, 0}};  // triangle texcoords (triangle.cc:12-17 defaults)
    MeshData mesh;         // mesh
    double u_scale = 0, v_scale = 0;
};

enum MatType : uint32_t { MAT_LAMBERT = 0, MAT_PHONG = 1, MAT_BLINNPHONG = 2, MAT_EMISSIVE = 3, MAT_DIELECTRIC = 4 };

// A sampler of a material channel or of the environment
// (src/xtcore/sampler_{col,tex,gradient,cubemap,erp}.cc). Textures index
// Scene::textures; a cubemap's faces are in CUBEMAP_FACE order: left, right,
// top, bottom, front, back (sampler_cubemap.h:11-19).
enum SamplerKind : uint32_t { SMP_NONE = 0, SMP_COLOR = 1, SMP_TEXTURE = 2, SMP_GRADIENT = 3, SMP_CUBEMAP = 4, SMP_ERP = 5 };
struct SamplerDesc {
    SamplerKind kind = SMP_NONE;
    float a[3] = {0, 0, 0}, b[3] = {0, 0, 0};  // colour value, or gradient ends
    int32_t tex[6] = {-1, -1, -1, -1, -1, -1};
    uint32_t bilinear = 0;                      // texture filtering (parseutil.cc:489-495)
};

// Constant-colour material parameters. A sampler that is absent reads as
// (0,0,0) and a scalar as 0 (src/xtcore/material.cc:58-70).
struct Material {
    std::string name;
    uint64_t handle = 0;
    MatType type = MAT_LAMBERT;
    float diffuse[3] = {0, 0, 0}, specular[3] = {0, 0, 0}, emissive[3] = {0, 0, 0};
    float reflectance = 0, exponent = 0, ior = 0;
    bool is_emissive = false;           // an "emissive" sampler or scalar exists
    SamplerDesc smp[3];                 // diffuse, specular, emissive
    std::vector<std::string> unsupported;  // samplers whose type the GPU path does not evaluate
};

struct Object {
    std::string name;
    uint64_t handle = 0;
    uint64_t geometry = 0, material = 0;  // name handles
};

enum EnvType : uint32_t { ENV_NONE = 0, ENV_GRADIENT = 1, ENV_COLOR = 2, ENV_UNSUPPORTED = 3, ENV_CUBEMAP = 4, ENV_ERP = 5 };
struct Environment {
    EnvType type = ENV_NONE;
    float a[3] = {0, 0, 0}, b[3] = {0, 0, 0};
    SamplerDesc smp;  // cubemap / erp
};

struct Scene {
    std::string source, title, description, version;
    Environment env;
    std::map<uint64_t, Camera> cameras;
    std::map<uint64_t, Geometry> geometries;
    std::map<uint64_t, Material> materials;
    std::map<uint64_t, Object> objects;  // iteration order == scan order
    std::map<std::string, uint64_t> names;
    std::vector<Image> textures;  // texture images the samplers read
    std::string error;

    uint64_t intern(const std::string &name);  // first-seen order handle
    const std::string *name_of(uint64_t handle) const;
};

// Loads a .scn file. `mods` are "group.group.prop:value" overrides applied
// after parsing (parseutil.cc:715-738). Returns a LoadStatus.
int load_scn(Scene &scene, const std::string &path, const std::vector<std::string> &mods);

// Wavefront OBJ as nmesh::io::import::obj reads it through tinyobjloader
// 1.0 (lib/nmesh/obj.cc:14-131): attributes, shapes split at `g` / `o`,
// fan-triangulated faces with the `usemtl` material of each, and the MTL
// materials of `mtllib` files.
struct ObjShape {
    std::string name;
    std::vector<int32_t> iv, in, it;  // 3 per triangle, -1 when absent
    std::vector<int32_t> mat;         // per triangle, -1: none
};
struct ObjMaterial {
    std::string name;
    float kd[3] = {0, 0, 0}, ks[3] = {0, 0, 0}, ke[3] = {0, 0, 0};
    float shininess = 1.f, ior = 1.f;
    std::string map_kd, map_ks, map_ke;
};
struct ObjFile {
    std::vector<float> v, n, uv;
    std::vector<ObjShape> shapes;
    std::vector<ObjMaterial> materials;
};
// Reads `path`; `mtl_dir` prefixes mtllib names. Returns 0 on success.
int read_obj_file(const std::string &path, const std::string &mtl_dir, ObjFile &out, std::string *err);

// All shapes of an OBJ file as one triangle soup (a `mesh` geometry,
// Mesh::build_octree(object_t&), mesh.cc:75-113). Returns 0 on success.
int read_obj(const std::string &path, MeshData &out, std::string *err);

}  // namespace xt
