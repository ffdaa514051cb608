/*
 * xtracer_amd: C ABI of the MI355X ray-traversal and integrator path.
 *
 * It replaces, below the reference's IIntegrator::render()
 * (src/xtcore/integrator.cc:29-53), the OpenMP tile loop, the per-sample
 * Integrator::render_tile() of the ao / depth / stencil / pathtracer
 * integrators and every Scene::intersection() query they make
 * (src/xtcore/scene.cc:191-216). Scene loading (io::scn::load,
 * src/xtcore/parseutil.cc:697-822) and the octree build (src/xtcore/octree.tml,
 * src/xtcore/mesh.cc:30-113) stay host-side and are exposed here too.
 *
 * Conventions: every call returns XT_OK (0) or an XT_E_* code, like the
 * reference loader's int returns; no C++ exception crosses this boundary.
 * Buffers are caller-owned; device memory behind an xt_dev is owned by it.
 * One xt_dev is driven by one host thread at a time.
 */
#ifndef XTRACER_AMD_H
#define XTRACER_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define XT_OK 0
#define XT_E_INVALID 1   /* bad argument */
#define XT_E_IO 2        /* scene file missing or unparsable (loader codes 1 / 2) */
#define XT_E_SCENE 3     /* scene content the GPU path does not evaluate */
#define XT_E_DEVICE 4    /* HIP runtime error */
#define XT_E_BUDGET 5    /* job not admitted: HBM budget */
#define XT_E_NOTFOUND 6  /* unknown scene id, ticket or camera */
#define XT_E_STATE 7     /* call not valid in the current state */

#define XT_ACCEL_KDTREE 0
#define XT_ACCEL_OCTREE 1

/* integrators, src/xtcore/integrator/{stencil,depth,ao,pathtracer,normal,uv,
 * emission,wireframe}; normal and uv always accumulate as the reference does
 * (XT_ACCUM_REFERENCE): their per-sample updates are not sums */
#define XT_INTEGRATOR_STENCIL 0
#define XT_INTEGRATOR_DEPTH 1
#define XT_INTEGRATOR_AO 2
#define XT_INTEGRATOR_PATHTRACER 3
#define XT_INTEGRATOR_NORMAL 4
#define XT_INTEGRATOR_UV 5
#define XT_INTEGRATOR_EMISSION 6
#define XT_INTEGRATOR_WIREFRAME 7

/* antialiasing::SAMPLE_DISTRIBUTION (src/xtcore/aa.h) */
#define XT_AA_GRID 0
#define XT_AA_RANDOM 1

/* per-pixel accumulation */
#define XT_ACCUM_DEFAULT 0   /* FP64 for ao/pathtracer, REFERENCE for the others (normal: always) */
#define XT_ACCUM_REFERENCE 1 /* the reference's float read-modify-write per sample */
#define XT_ACCUM_FP64 2      /* compensated (Neumaier) FP64 sum, rounded once */

/* hit primitive ids returned by xt_trace */
#define XT_HIT_NONE 0xffffffffu
#define XT_HIT_PLANE_BASE 0xffff0000u /* XT_HIT_PLANE_BASE + k: plane k (k < 65533) */

typedef struct xt_scene xt_scene; /* host scene: loaded, flattened, trees built */
typedef struct xt_dev xt_dev;     /* one HIP device with resident scenes */

/* ---------------------------------------------------------------- host side */

/* io::scn::load (parseutil.cc:697). mods: "group.group.prop:value" strings.
 * Returns XT_OK, XT_E_IO (parse error / missing file) or XT_E_SCENE. */
int xt_scene_load(const char *scn_path, const char *const *mods, int n_mods, xt_scene **out);

typedef struct xt_build_params {
    int kd_max_depth;     /* 0: 8 + 1.3 log2(N) */
    int kd_leaf_prims;    /* 0: 4 */
    float kd_cost_intersect; /* 0: 12 */
    int threads;          /* 0: all cores */
    int kd_on_device;     /* 1: xt_scene_build leaves the kd-tree out and xt_scene_upload builds it on the
                             device (xt_scene_build_kd_device with these parameters) */
} xt_build_params;

/* Flattens the scene (objects in scan order, mesh triangles in octree
 * pre-order) and builds the octree (reference rules) and the SAH kd-tree
 * (src/xtcore/bvh.cc:13-110 is the reference's empty stub of a tree over
 * the scene; its README.md:47 lists a kd-tree). */
int xt_scene_build(xt_scene *s, const xt_build_params *p);
void xt_scene_free(xt_scene *s);
const char *xt_scene_error(const xt_scene *s);

typedef struct xt_scene_info {
    uint32_t n_objects, n_prims, n_planes, n_materials, n_cameras;
    uint32_t n_kd_nodes, n_kd_refs, kd_depth, n_oct_nodes, n_oct_refs;
    uint32_t env_type;
    uint32_t any_vertex_normals;
    double build_seconds_octree, build_seconds_kd;
} xt_scene_info;
int xt_scene_get_info(const xt_scene *s, xt_scene_info *info);

typedef struct xt_object_info {
    char name[64];
    uint64_t handle;      /* first-seen name order (string-pool order) */
    uint32_t geom_type;   /* 0 plane, 1 sphere, 2 triangle, 3 mesh, ~0 missing */
    uint32_t material;    /* index, ~0 if missing */
    uint32_t first_prim, prim_count;
    int32_t plane;
} xt_object_info;
int xt_scene_get_object(const xt_scene *s, uint32_t index, xt_object_info *info);

/* Copies of the flattened arrays (any pointer may be NULL):
 * prims: n_prims * 12 floats (XtPrim), prims64: n_prims * 9 doubles,
 * prim_local: index of each prim inside its geometry (OBJ face order),
 * planes: n_planes * 20 words (XtPlane), materials: n_materials * 12 words. */
int xt_scene_get_arrays(const xt_scene *s, void *prims, double *prims64, uint32_t *prim_local, void *planes,
                        void *materials, float env[7]);

/* Geometry parameters of object `index` as loaded (FP64): plane normal +
 * distance, sphere centre + radius, triangle v0 v1 v2 (9 doubles). */
int xt_scene_get_object_params(const xt_scene *s, uint32_t index, double out[9]);
/* Per-vertex normals of the soup, 9 floats per prim (zeros where absent);
 * XT_E_NOTFOUND when no triangle carries vertex normals. */
int xt_scene_get_vnormals(const xt_scene *s, float *out);
/* Texture coordinates, for the uv integrator: prim_tc (6 doubles per soup
 * primitive: u, v of each triangle corner, zeros for spheres) as the
 * reference's Triangle::tc holds them (explicit triangles: parseutil.cc:395-397;
 * mesh corners: mesh.cc:56-58, Triangle() defaults for corners without one,
 * triangle.cc:12-17), and uv_scale (2 doubles per object: the geometry's
 * u_scale, v_scale, parseutil.cc:410-413). Either pointer may be NULL. */
int xt_scene_get_texcoords(const xt_scene *s, double *prim_tc, double *uv_scale);
/* Samplers and the textures they read, as io::scn::load builds them
 * (parseutil.cc:147-203, 428-500; img.cc:17-66 for the texels): counts[0..2]
 * receive the numbers of textures, texel floats and table samplers; then,
 * where not NULL: tex_info (offset in texels / 4, width, height, 0 per
 * texture), texels (4 floats per texel, row 0 at the top), samplers (14
 * words each: kind (1 colour, 2 texture, 3 gradient, 4 cubemap, 5 erp),
 * bilinear, 6 texture indices (cubemap faces left, right, top, bottom,
 * front, back), then a[3], b[3] as floats), mat_smp (3 per material, the
 * table sampler of diffuse / specular / emissive or 0xffffffff for a colour
 * or absent one) and env_smp (the environment's table sampler). */
int xt_scene_get_samplers(const xt_scene *s, uint32_t counts[3], uint32_t *tex_info, float *texels, uint32_t *samplers,
                          uint32_t *mat_smp, uint32_t *env_smp);

/* The kd-tree as uploaded: n_kd_nodes * 2 words (XtKdNode), n_kd_refs
 * primitive references and the root box (lo xyz, hi xyz). */
int xt_scene_get_kdtree(const xt_scene *s, uint32_t *nodes, uint32_t *refs, float bounds[6]);

/* Device kd-tree build of a resident scene (no counterpart in the
 * reference, which has no kd-tree: bvh.cc:13-110 is a stub). The host build's
 * binned SAH with triangle clipping, run level by level on the device: the
 * tree is the one xt_scene_build makes on one thread (params->threads = 1),
 * node for node and in the same depth-first order. It replaces the scene's
 * kd-tree (node order as the depth-first build, whatever xt_tuning.kd_layout
 * says). XT_E_STATE while served jobs use the scene. params may be NULL
 * (defaults); threads and kd_on_device are not read. */
typedef struct xt_kd_build_stats {
    double ms;                /* device time of the build (HIP events), 0 for a host-built tree */
    uint32_t n_nodes, n_refs; /* of the tree the scene holds */
    uint32_t depth, leaves, levels;
    uint32_t pad;
} xt_kd_build_stats;
int xt_scene_build_kd_device(xt_dev *d, uint32_t scene_id, const xt_build_params *params, xt_kd_build_stats *stats);
/* The resident kd-tree: n_nodes * 2 words, n_refs references, the root box
 * (lo xyz, hi xyz), and the stats (sizes first: NULL arrays copy nothing). */
int xt_scene_get_kd_device(xt_dev *d, uint32_t scene_id, uint32_t *nodes, uint32_t *refs, float bounds[6],
                           xt_kd_build_stats *stats);

/* Per-mesh reference octree statistics (node count, deepest level). */
int xt_scene_mesh_octree(const xt_scene *s, uint32_t k, uint64_t *geometry, uint64_t *nodes, uint64_t *max_level,
                         uint64_t *triangles);

/* Camera kinds, as the .scn "type" names them (parseutil.cc:262-275). */
#define XT_CAMERA_THIN_LENS 0u  /* Perspective, src/xtcore/camera/perspective.cc:26-98 */
#define XT_CAMERA_ODS 1u        /* omni-directional stereo, camera/ods.cc:15-33 */
#define XT_CAMERA_ERP 2u        /* equirectangular, camera/erp.cc:11-25 */
#define XT_CAMERA_CUBEMAP 3u    /* vertical strip +x -x +y -y +z -z, camera/cubemap.cc:17-60;
                                   the image height must be a multiple of 6 */

/* A camera, set up on the host in FP64 the way the reference sets it up;
 * the device evaluates its primary rays in FP64 with the reference's
 * float/double mix. */
typedef struct xt_camera {
    double position[3];
    double basis[9];      /* thin-lens: row-major m_transform 3x3; ODS / ERP: the
                             orientation's rotation Rx*Ry*Rz (matrix.cc:485-492),
                             row-major; cubemap: unused */
    double dir_z;         /* thin-lens: 1 / tan(fov * RADIAN / 2) */
    double aperture, flength;
    uint32_t thin_lens;   /* thin-lens with flength > 0 */
    uint32_t kind;        /* XT_CAMERA_* */
    double ipd;           /* ODS: the float inter-pupillary distance */
} xt_camera;
/* camera_name NULL or "": the first camera in name order (workspace.cc:33-38) */
int xt_scene_camera(const xt_scene *s, const char *camera_name, xt_camera *out);
/* A thin-lens camera from an explicit pose (serving requests with their own
 * views). */
int xt_camera_make(const double position[3], const double target[3], const double up[3], float fov_deg,
                   float aperture, float flength, xt_camera *out);

/* Writes an assembled RGBA float framebuffer (width*height*4, row-major) as
 * the file type its extension names: .png .bmp .tga (8-bit, the reference's
 * conversion: v*255, set to 255 above 255, truncated; BMP composites alpha
 * over magenta like the reference's writer), .hdr (Radiance RGBE), .exr
 * (OpenEXR, FLOAT B G R, uncompressed), .pfm, .ppm. Replaces
 * nimg::io::save::{png,bmp,tga,hdr,exr} (lib/nimg/img.cc:72-318) as the GUI
 * (src/frontend/gui/action.cc:95-110) and CLI (src/frontend/cli/xtracer.cc:
 * 140-150) call them. XT_E_INVALID for an unknown extension, XT_E_IO when the
 * file cannot be written. */
int xt_image_write(const char *path, const float *rgba, uint32_t width, uint32_t height);

/* Reads a texture image as the scene loader does for `texture`, `cubemap`
 * and `erp` samplers: RGBA floats (byte / 255, alpha 1 for RGB files), row 0
 * at the top. Replaces nimg::io::load::image (lib/nimg/img.cc:17-66) for the
 * files it reads as 8-bit RGB / RGBA: PNG (RGB, RGBA, palette; any depth;
 * Adam7; tRNS), TGA (raw or
 * run-length, 24 / 32 bit), BMP (uncompressed 24 / 32 bit), PPM (P6, maxval
 * 255), Radiance HDR (converted to 8 bits as stb_image does), baseline and
 * progressive JPEG (stb_image's integer IDCT, chroma upsampling and YCbCr
 * conversion). With rgba NULL only *width / *height are set. XT_E_IO when the file
 * cannot be read, XT_E_INVALID for a format or variant not decoded. */
int xt_image_read(const char *path, float *rgba, uint32_t *width, uint32_t *height);

/* ----------------------------------------------------------------- device */

/* hbm_budget_bytes caps the device memory the handle allocates (scenes,
 * wavefront and job buffers); 0 means 90% of the device's memory. */
int xt_dev_create(xt_dev **out, int hip_device, uint64_t hbm_budget_bytes);
void xt_dev_destroy(xt_dev *d);
const char *xt_last_error(const xt_dev *d);

/* Launch-shape knobs (the autotuner's search space). In xt_dev_set_tuning
 * a 0 field keeps the current value; xt_dev_get_tuning reports the values in
 * force (trace_blocks_per_cu 0: as many blocks as a CU keeps resident). */
#define XT_TUNING_AUTO 0xffffffffu /* trace_blocks_per_cu: back to resident fill */
typedef struct xt_tuning {
    uint32_t kd_stack;        /* kd short-stack entries per lane: 8, 12, 16 or 24 (12) */
    uint32_t trace_block;     /* lanes per trace block: 128 or 256 (256) */
    uint32_t wave_paths;      /* paths in flight per wavefront pass, 4096..2^28 (64M: a 1080p 16 spp frame in one pass) */
    uint32_t trace_blocks_per_cu; /* persistent trace grid: blocks per CU, <= 64 (resident fill) */
    uint32_t trace_segments;  /* 1: one work cursor over a trace queue; 8: the queue cut into 8 contiguous
                                 segments, one per XCD, each drained by its own XCD's waves first (1) */
    uint32_t ray_order;       /* camera rays of a pass: 1 sample by sample over the image; 2 all samples of
                                 an 8x8 tile together; 3 all samples of a pixel together (1) */
    uint32_t kd_layout;       /* kd-tree node order of scenes uploaded afterwards: 1 depth-first; 2 child
                                 pairs grouped into 128-byte lines three levels deep (1) */
    uint32_t kd_leaf;         /* kd leaf primitives in the FP32 walk: 1 a copy of each referenced record in
                                 leaf order; 2 read from the soup through the reference list (1) */
    uint32_t frame_lanes;     /* one-job renders: 1 a pass's paths in one group on the device stream; 2..4
                                 its pixels cut into that many groups whose rounds run on streams of their
                                 own, so that one group's shading and launch tails overlap another's trace;
                                 images are the same bits either way (2) */
} xt_tuning;
int xt_dev_set_tuning(xt_dev *d, const xt_tuning *t);
int xt_dev_get_tuning(const xt_dev *d, xt_tuning *t);

/* Makes a built scene resident in HBM. The scene may be freed afterwards. */
int xt_scene_upload(xt_dev *d, const xt_scene *s, uint32_t *scene_id);
int xt_scene_release(xt_dev *d, uint32_t scene_id);

/* Closest-hit queries: the surface of Scene::intersection. Rays are FP32 SoA
 * (origin, unit direction); hits receive t, the primitive id (soup index,
 * XT_HIT_PLANE_BASE + plane, or XT_HIT_NONE), the barycentrics (u, v) of a
 * triangle hit, and flags (bit 0: re-tested in FP64). Host pointers. */
typedef struct xt_rays {
    const float *ox, *oy, *oz, *dx, *dy, *dz;
} xt_rays;
typedef struct xt_hits {
    float *t;
    uint32_t *prim;
    float *u, *v;
    uint32_t *flags;
} xt_hits;
int xt_trace(xt_dev *d, uint32_t scene_id, int accel, const xt_rays *rays, xt_hits *hits, uint64_t n);

/* Same on device-resident buffers: ray0[i] = (ox, oy, oz, origin) and
 * ray1[i] = (dx, dy, dz, 0) as float4, where `origin` holds the bits of the
 * primitive id the ray leaves from (XT_HIT_NONE for a free ray); hit[i] =
 * (t, prim bits, u, v) as float4.
 * Runs on the device's stream; *kernel_ms (optional) gets the trace kernel
 * time measured with HIP events on that stream. */
int xt_trace_device(xt_dev *d, uint32_t scene_id, int accel, const void *ray0, const void *ray1, void *hit,
                    uint32_t *flags, uint64_t n, float *kernel_ms);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). xt_trace_device with its scratch (the FP64 re-check and
 * re-test lists and the counters, 20 B per ray) in a fixed device array of
 * XT_TRACE_STATIC_MAX_RAYS rays that each GPU holds until the process ends;
 * it calls no allocator. n above that limit: XT_E_INVALID. The scene must be
 * uploaded. Every call on one GPU shares that GPU's array, so calls that
 * overlap on one GPU (other handles, other threads) race. Synchronous on the
 * device's stream. */
#define XT_TRACE_STATIC_MAX_RAYS (1u << 20)
int xt_trace_device_static(xt_dev *d, uint32_t scene_id, int accel, const void *ray0, const void *ray1, void *hit,
                           uint32_t *flags, uint64_t n);

typedef struct xt_job {
    uint32_t scene_id;
    uint32_t integrator;   /* XT_INTEGRATOR_* */
    uint32_t width, height;
    uint32_t samples;      /* params_t::samples */
    uint32_t aa;           /* params_t::aa (level) */
    uint32_t aa_mode;      /* XT_AA_GRID / XT_AA_RANDOM */
    uint32_t rdepth;       /* params_t::rdepth */
    uint32_t accel;        /* XT_ACCEL_* */
    uint32_t accum;        /* XT_ACCUM_* */
    uint32_t seed;
    uint32_t job_key;      /* distinguishes jobs in the random streams */
    uint32_t samples_per_pass; /* 0: auto (wavefront size) */
    uint32_t pad;
    xt_camera camera;
} xt_job;

typedef struct xt_job_stats {
    uint64_t queries;      /* executed closest-hit queries (Scene::intersection calls) */
    uint64_t retests;      /* rays re-tested in FP64 */
    uint64_t samples;      /* camera samples */
    double ms_total;       /* device time of the render, HIP events */
    double ms_trace;       /* sum of trace kernel times */
    uint32_t trace_launches;
    uint32_t passes;
    double latency_ms;     /* serving: completion - submission */
    double queue_ms;       /* serving: admission - submission */
    uint64_t trace_fail;   /* FP32 traversals that ran out of steps or stack (re-tested in FP64) */
    uint64_t retest_fail;  /* FP64 re-test traversals that ran out of steps (closest hit so far kept) */
    double ms_trace_busy;  /* time in which at least one trace kernel ran: the union of the timed launches'
                              intervals (equal to ms_trace when the launches do not overlap) */
} xt_job_stats;

/* IMaterial::shade (src/xtcore/material/lambert.cc:9-28, phong.cc:11-33,
 * blinnphong.cc:9-32, emissive.cc:8-18, dielectric.cc:9-21): the light one
 * point emitter sends from a hit through the hit object's material, as the
 * reference's materials define it (none of its integrators calls it).
 * intensity: 3 floats per query, read and written as shade() updates its
 * `intensity` argument (Lambert and Phong add to it; Blinn-Phong, emissive and
 * dielectric assign it); result (optional): shade()'s return value, 1 or 0,
 * or -1 for an object index out of range or an object without a material
 * (intensity left as it was). Host buffers. */
typedef struct xt_shade_query {
    double point[3];     /* hit_record.point */
    double normal[3];    /* hit_record.normal, as the hit record holds it */
    double texcoord[3];  /* hit_record.texcoord */
    double camera[3];    /* camera->position */
    double emitter[3];   /* emitter->position */
    float intensity[3];  /* emitter->intensity */
    uint32_t object;     /* object index in scan order: its material shades */
} xt_shade_query;
int xt_shade(xt_dev *d, uint32_t scene_id, const xt_shade_query *q, uint64_t n, float *intensity, int32_t *result);

/* Plain one-job-per-launch render. rgba_out: width*height*4 floats, host. */
int xt_render(xt_dev *d, const xt_job *job, float *rgba_out, xt_job_stats *stats);

/* xt_render, then the render's ray graph written to ply_path as ASCII PLY:
 * for a depth job one two-vertex path per sample, camera origin to hit point
 * ((0, 0, 0) on a miss, the hit record's default), coloured with the pixel
 * after that sample, in the
 * reference's tile / pixel / sample order; other integrators record no paths.
 * Replaces render::assemble(raygraph_t&) + raygraph::write (context.cc:54-61,
 * util/raygraph.cc:8-78) as the GUI's export calls them (action.cc:65-85).
 * The samples are held on the host: about 56 B per sample. XT_E_IO when the
 * file cannot be written. */
int xt_render_raygraph(xt_dev *d, const xt_job *job, float *rgba_out, const char *ply_path);
/* Same with a device-resident output buffer. */
int xt_render_device(xt_dev *d, const xt_job *job, float *d_rgba_out, xt_job_stats *stats);

/* Serving: continuous batching of many jobs on one device, in place of the
 * reference's one blocking IIntegrator::render() per context.
 * xt_submit queues the job and returns a ticket; it is admitted (FIFO) once
 * its buffers (80 B per pixel in whole passes; on the regeneration engine
 * below 16 B per pixel and sample plus 16, or 40 per pixel) fit the HBM
 * budget, and refused with
 * XT_E_BUDGET when they never could. xt_serve_step runs the configured
 * query rounds of the regeneration engine and one wavefront pass shared by
 * the other admitted jobs, and reports the jobs not yet finished. xt_poll
 * sets *done once the job's image
 * is ready; with rgba_out (width*height*4 floats) it then copies the image
 * and forgets the ticket, with rgba_out NULL it only reports (the image is
 * kept for a later call); stats->latency_ms = completion - submission and
 * stats->queue_ms = admission - submission. A job served in whole passes is
 * bit-identical to the same job rendered with xt_render. xt_drain steps until every
 * admitted and waiting job has finished. */
int xt_submit(xt_dev *d, const xt_job *job, uint64_t *ticket);
int xt_poll(xt_dev *d, uint64_t ticket, int *done, float *rgba_out, xt_job_stats *stats);
int xt_serve_step(xt_dev *d, int *active_jobs);
int xt_drain(xt_dev *d);

typedef struct xt_serve_info {
    uint64_t queries;      /* closest-hit queries of all passes */
    uint64_t retests;
    uint64_t samples;      /* camera samples of all passes */
    uint64_t passes;
    uint32_t jobs_waiting, jobs_active, jobs_finished, pad;
    double ms_device;      /* device time of all passes, HIP events */
    uint64_t bytes_held;   /* job buffers held in HBM (in use + pooled) */
} xt_serve_info;
int xt_serve_get_info(const xt_dev *d, xt_serve_info *info);

/* Scheduling. XT_SERVE_SRPT (default): the jobs with the least work left go
 * first. XT_SERVE_FAIR: whole passes give the oldest job half of each pass and
 * share the rest equally; the regeneration engine gives every job an equal
 * share of each round's free path slots. XT_SERVE_FCFS: admission order.
 * XT_SERVE_HYBRID: jobs with less work left than half the path slots first
 * (shortest first), then the others in admission order. Images do not depend
 * on it. */
#define XT_SERVE_FAIR 0
#define XT_SERVE_SRPT 1
#define XT_SERVE_FCFS 2
#define XT_SERVE_HYBRID 3
int xt_serve_set_policy(xt_dev *d, int policy);

/* Engines. ao and pathtracer jobs with the default accumulation run on the
 * path-regeneration engine: a pool of path slots shared by all of them, in
 * which every query round refills the slots of finished paths with camera
 * rays of waiting samples (in priority order), so jobs join and leave between
 * rounds; each round is one fixed launch sequence, replayed from a hipGraph.
 * Each finished path's radiance is kept per (pixel, sample) and summed per
 * pixel in sample order when the job ends, so a served image is bit-identical
 * to xt_render's. A job whose sample buffer would exceed sample_buffer_mb sums
 * in 64-bit fixed point (2^-32) instead, independent of the order in which
 * paths end; its image differs from xt_render's by the rounding of the two
 * sums (a few units in the last place of the float result). Every other job
 * runs in whole passes (bit-identical to xt_render). In xt_serve_configure a
 * 0 field keeps its value. */
typedef struct xt_serve_config {
    uint32_t regen;           /* 2: the regeneration engine as above; 1: every job in whole passes (2) */
    uint32_t graph;           /* 2: rounds replayed from a hipGraph; 1: launched kernel by kernel (2) */
    uint32_t rounds_per_step; /* query rounds per xt_serve_step on the regeneration engine, even (4) */
    uint32_t paths;           /* path slots of the regeneration engine, 4096..2^28 (16M; a larger pool
                                 runs a long batch faster and makes new requests wait longer for a
                                 round to end); set while it holds no job */
    uint32_t sample_buffer_mb;/* a job keeps its samples' radiance (16 B each) up to this size in MiB and is
                                 bit-identical to xt_render; a larger one sums in fixed point (4096) */
    uint32_t lane_streams;    /* 2: the lanes of a round (trace, re-check, re-test, shading) run concurrently,
                                 on up to 4 streams (branches of the graph); 1: one after another (2) */
    uint32_t loop;            /* 2: device-fed: the host writes admitted jobs to a request ring in host memory;
                                 every round the device takes new requests and orders the jobs by the policy,
                                 and each replay of two rounds resolves the jobs it finishes and posts them
                                 to a completion ring; xt_serve_step keeps two replays queued and waits on the
                                 device only when nothing else moved. 1: host-stepped: each step sets the
                                 priorities, replays the rounds and synchronises (2); set while the engine
                                 holds no job */
} xt_serve_config;
int xt_serve_configure(xt_dev *d, const xt_serve_config *c);
int xt_serve_get_config(xt_dev *d, xt_serve_config *c);

#ifdef __cplusplus
}
#endif

#endif /* XTRACER_AMD_H */
