// Plain-old-data layouts shared by the host builders and the HIP kernels.
// Everything a kernel reads from HBM is described here; sizes are asserted so
// that host and device agree byte for byte.
#pragma once
#include <stdint.h>

#define XT_PRIM_TRIANGLE 0u
#define XT_PRIM_SPHERE 1u

// Primitive flag bits (XtPrim::flags).
#define XT_PF_DEGENERATE 1u   // triangle area below 1e-8: never hit (triangle.cc:161-166)
#define XT_PF_VNORMALS 2u     // triangle carries per-vertex normals
#define XT_PF_TINY_SPHERE 4u  // radius <= 0: never hit (disc <= 0 by Cauchy-Schwarz)

// One bounded primitive of the scene soup, 48 bytes, FP32. Soup index order
// equals the reference's closest-hit tie order: objects in scan order and,
// inside a mesh, triangles in octree pre-order.
//   triangle: p0 = v0, p1 = v1, p2 = v2 (exact: vertex data is float),
//             aux = 1 / |(v1 - v0) x (v2 - v0)| (0 when degenerate)
//   sphere:   p0 = centre, p1 = (radius, radius^2, 1 / radius)
// tf = type | (flags << 4).
struct XtPrim {
    float p0[3];
    uint32_t tf;
    float p1[3];
    uint32_t object;  // index into the object table (scan order)
    float p2[3];
    float aux;
};
#define XT_PRIM_TYPE(tf) ((tf) & 15u)
#define XT_PRIM_FLAGS(tf) ((tf) >> 4)

// FP64 copy used only by the near-edge / grazing re-test: the values the
// reference computes with (triangle vertices, or sphere centre + radius).
struct XtPrim64 {
    double a[3], b[3], c[3];
};

// Infinite plane (plane.cc:19-61). `point` is abs(normal) * distance, the
// point the reference intersects against.
struct XtPlane {
    float normal[3];
    uint32_t object;
    float point[3];
    uint32_t pad;
    double normal64[3];
    double point64[3];
};

#define XT_MAT_LAMBERT 0u
#define XT_MAT_PHONG 1u
#define XT_MAT_BLINNPHONG 2u
#define XT_MAT_EMISSIVE 3u
#define XT_MAT_DIELECTRIC 4u

struct XtMaterial {
    uint32_t type;
    float kd[3];
    float ks[3];
    float ke[3];
    float reflectance;
    float exponent;
};

#define XT_ENV_CUBEMAP 3u
#define XT_ENV_ERP 4u
#define XT_ENV_NONE 0u
#define XT_ENV_GRADIENT 1u
#define XT_ENV_COLOR 2u

struct XtEnv {
    uint32_t type;
    float a[3];
    float b[3];
    uint32_t smp;  // cubemap / erp: index into the sampler table
};

// Samplers of material channels and the environment
// (src/xtcore/sampler_{col,tex,gradient,cubemap,erp}.cc).
#define XT_SMP_COLOR 1u
#define XT_SMP_TEXTURE 2u
#define XT_SMP_GRADIENT 3u
#define XT_SMP_CUBEMAP 4u
#define XT_SMP_ERP 5u
struct XtSampler {
    uint32_t kind;
    uint32_t bilinear;  // texture filtering
    int32_t tex[6];     // texture; erp; cubemap faces left, right, top, bottom, front, back
    float a[3], b[3];   // colour, gradient ends
};
// A texture image: width * height float4 texels (byte / 255, row 0 at the
// top) starting at `offset` in the texel array.
struct XtTexture {
    uint32_t offset, width, height, pad;
};

// kd-tree node, 8 bytes.
//   inner: word = (first_child << 2) | axis, split = plane position;
//          children are stored as a pair at first_child, first_child + 1,
//          first_child even (the pair is one aligned 16-byte record).
//   leaf:  word = (first_ref << 2) | 3, count = number of prim refs.
struct XtKdNode {
    uint32_t word;
    union {
        float split;
        uint32_t count;
    };
};

// Octree node in breadth-first order (children of a node are 8 consecutive
// nodes). Box rounded outward to FP32. Items are a contiguous range of the
// soup because the soup is laid out in the octree's pre-order.
struct XtOctNode {
    float lo[3];
    uint32_t first_item;
    float hi[3];
    uint32_t n_items;
    uint32_t first_child;  // 0 when the node has no children
    uint32_t subtree_items;
    uint32_t pad[2];
};

#ifdef __cplusplus
static_assert(sizeof(XtPrim) == 48, "XtPrim layout");
static_assert(sizeof(XtPrim64) == 72, "XtPrim64 layout");
static_assert(sizeof(XtPlane) == 80, "XtPlane layout");
static_assert(sizeof(XtMaterial) == 48, "XtMaterial layout");
static_assert(sizeof(XtKdNode) == 8, "XtKdNode layout");
static_assert(sizeof(XtOctNode) == 48, "XtOctNode layout");
#endif
