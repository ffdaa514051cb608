// Flattened, device-ready form of a loaded Scene.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../xt_layout.h"
#include "kdtree.h"
#include "octree.h"
#include "scene.h"

namespace xt {

#define XT_NO_MATERIAL 0xffffffffu

struct FlatObject {
    uint64_t handle = 0;
    std::string name;
    uint32_t geom_type = 0;     // GeomType
    uint32_t material = XT_NO_MATERIAL;
    uint32_t first_prim = 0, prim_count = 0;  // soup range (bounded geometry)
    int32_t plane = -1;         // index into planes (planes only)
};

struct FlatScene {
    std::vector<FlatObject> objects;  // scan order
    std::vector<XtPrim> prims;        // soup in tie order
    std::vector<XtPrim64> prims64;
    std::vector<float> vnormals;      // 9 floats per prim when any triangle has vertex normals
    std::vector<double> tc64;         // 6 per prim: a triangle's texcoords (u, v per corner), 0 for spheres
    std::vector<double> obj_uv_scale; // 2 per object: the geometry's u_scale, v_scale
    std::vector<uint32_t> prim_local; // index of the primitive inside its geometry (OBJ face order)
    std::vector<XtPlane> planes;
    std::vector<XtMaterial> materials;
    std::vector<std::string> material_names;
    std::vector<uint8_t> material_supported;  // 0 when a sampler cannot be evaluated on the GPU
    XtEnv env{};
    bool env_supported = true;
    std::vector<float> texels;          // 4 floats per texel, all textures
    std::vector<XtTexture> textures;
    std::vector<XtSampler> samplers;    // non-constant material channels and the environment
    std::vector<uint32_t> mat_smp;      // 3 per material (diffuse, specular, emissive): sampler or XT_NO_MATERIAL

    // acceleration structures over the soup
    KdTree kd;
    std::vector<XtOctNode> oct_nodes;
    std::vector<uint32_t> oct_refs;

    // Per-mesh reference octrees (octree.tml rules over OBJ face order),
    // keyed by geometry handle: node count and deepest level, for inspection.
    struct MeshOctreeInfo {
        uint64_t geometry;
        size_t nodes, max_level, triangles;
    };
    std::vector<MeshOctreeInfo> mesh_octrees;

    double build_seconds_octree = 0, build_seconds_kd = 0;
    bool kd_deferred = false;  // the kd-tree is built on the device when the scene is uploaded
    KdParams kd_params;        // its parameters
};

// Builds the flat form. `kd` parameters control the kd-tree; the octree uses
// the reference limits (10 items per node, depth 10).
int flatten_scene(const Scene &sc, FlatScene &out, const KdParams &kd, std::string *err);

}  // namespace xt
