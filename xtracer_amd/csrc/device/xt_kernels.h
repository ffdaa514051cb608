// Wavefront kernels. One "pass" renders `spp_pass` samples of every pixel:
//   k_camera  -> queue of camera rays (compacted), path state, zeroed radiance
//   repeat per query index q = 1..rdepth:
//     k_trace   FP32 closest hit per queued ray, rays needing the FP64 re-test
//               appended to a list
//     k_retest  FP64 re-trace of the listed rays with the reference predicates
//     k_shade_* integrator rule; surviving paths appended (ballot + prefix
//               sum, one atomic per wave) to the next queue
//   k_accumulate  per-pixel sum over the pass's samples in sample order
// Every kernel reads its work count from device memory and loops over it, so
// the launch sequence is fixed and can be captured in a hipGraph.
#pragma once
#include <hip/hip_runtime.h>

#include "../../../include/xtracer_amd.h"
#include "xt_sample.h"
#include "xt_texture.h"
#include "xt_trace.h"

#define XT_BLOCK 256
#define XT_OCT_BLOCK 64
#define XT_TRACE_CHUNK 1  // 64-ray groups a trace wave takes per cursor fetch
#define XT_DEAD 0xfffffffeu  // origin word of a queue entry without a ray

// Indices into the device counter block.
enum {
    CNT_Q0 = 0,      // ray count of queue 0
    CNT_Q1 = 1,      // ray count of queue 1
    CNT_RETEST = 2,  // rays listed for the FP64 re-test
    CNT_CURSOR = 3,  // work cursor of the persistent trace kernel
    CNT_RECHECK = 4, // rays listed for the FP64 re-check of their deferred candidates
    CNT_WORDS = 8
};
#define XT_TRACE_SEGMENTS 8  // XCDs of an MI355X: one trace-queue segment each
struct Counters {
    uint32_t n[CNT_WORDS];
    uint32_t seg[XT_TRACE_SEGMENTS];  // work cursors of the trace queue's segments (reset with CNT_CURSOR)
    unsigned long long queries, retests, steps_exceeded, trace_fail;
};

struct Queue {
    float4 *r0;  // origin xyz, origin primitive bits
    float4 *r1;  // direction xyz, analytic self-hit t
    uint32_t *path;
};

struct JobDev {
    uint32_t W, H, tiles_x, NT;  // NT: pixel slots in 8x8 tile order (>= W*H)
    uint32_t spp, aa, aa_mode, n_aa, total;  // total = n_aa * spp samples per pixel
    uint32_t rdepth, integrator, seed, job_key, accum;
    float wf, hf, aa_sub, aa_off;
    double weight;  // per-sample weight: float(aa weight) * (1. / spp)
    float weight_f;
    xt_camera cam;
};

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Wave-aggregated append: returns this lane's slot (only valid if `want`).
__device__ __forceinline__ uint32_t wave_append(bool want, uint32_t *counter) {
    const unsigned long long m = __ballot(want);
    const uint32_t n = (uint32_t)__popcll(m);
    const uint32_t lane = lane_id();
    const uint32_t leader = m ? (uint32_t)__ffsll((long long)m) - 1 : 0;
    uint32_t base = 0;
    if (lane == leader && n) base = atomicAdd(counter, n);
    base = __shfl(base, (int)leader);
    const unsigned long long below = lane ? (m & ((~0ull) >> (64 - lane))) : 0ull;
    return base + (uint32_t)__popcll(below);
}

// Block-aggregated append (one atomic per block): every thread of the block
// must call it, with `lds` pointing at BS / 64 + 1 shared words.
template <int BS>
__device__ __forceinline__ uint32_t block_append(bool want, uint32_t *counter, uint32_t *lds) {
    constexpr int NW = BS / 64;
    const unsigned long long m = __ballot(want);
    const uint32_t lane = lane_id(), wid = threadIdx.x >> 6;
    if (lane == 0) lds[wid] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot = 0;
        for (int w = 0; w < NW; ++w) {
            const uint32_t c = lds[w];
            lds[w] = tot;
            tot += c;
        }
        lds[NW] = tot ? atomicAdd(counter, tot) : 0u;
    }
    __syncthreads();
    const uint32_t base = lds[NW] + lds[wid];
    __syncthreads();
    const unsigned long long below = lane ? (m & ((~0ull) >> (64 - lane))) : 0ull;
    return base + (uint32_t)__popcll(below);
}

// Slot -> pixel in 8x8 tile order, so that a wave holds one 8x8 block.
__device__ __forceinline__ bool slot_pixel(const JobDev &j, uint32_t q, uint32_t &x, uint32_t &y) {
    const uint32_t tile = q >> 6, within = q & 63u;
    x = (tile % j.tiles_x) * 8u + (within & 7u);
    y = (tile / j.tiles_x) * 8u + (within >> 3);
    return x < j.W && y < j.H;
}

static __global__ void k_reset(Counters *c, uint32_t mask) {
    const uint32_t i = threadIdx.x;
    if (i < CNT_WORDS && (mask >> i) & 1u) c->n[i] = 0;
    if (i < XT_TRACE_SEGMENTS && (mask >> CNT_CURSOR) & 1u) c->seg[i] = 0;
    if (i == 0 && (mask >> 31) & 1u) {
        c->queries = 0;
        c->retests = 0;
        c->steps_exceeded = 0;
        c->trace_fail = 0;
    }
}

// k_reset over `gridDim.x` consecutive counter blocks (one per lane).
static __global__ void k_reset_lanes(Counters *c, uint32_t mask) {
    Counters *b = c + blockIdx.x;
    const uint32_t i = threadIdx.x;
    if (i < CNT_WORDS && (mask >> i) & 1u) b->n[i] = 0;
    if (i < XT_TRACE_SEGMENTS && (mask >> CNT_CURSOR) & 1u) b->seg[i] = 0;
    if (i == 0 && (mask >> 31) & 1u) {
        b->queries = 0;
        b->retests = 0;
        b->steps_exceeded = 0;
        b->trace_fail = 0;
    }
}

static __global__ void k_set_count(Counters *c, uint32_t which, uint32_t value) { c->n[which] = value; }

static __global__ void k_fill_u32(uint32_t *p, uint32_t value, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = value;
}

// --------------------------------------------------------------- camera

#pragma clang fp contract(off)
// Perspective::get_primary_ray (src/xtcore/camera/perspective.cc:38-98) in
// FP64; the lens draws come from query stream 0 of the path.
__device__ __forceinline__ void perspective_ray(const JobDev &j, float x, float y, uint32_t key, double o[3],
                                                double d[3]) {
    const xt_camera &c = j.cam;
    const double aspect = (double)j.wf / (double)j.hf;
    double dir[3] = {(2.0 * (double)x / (double)j.wf) - 1.0, ((2.0 * (double)y / (double)j.hf) - 1.0) / aspect,
                     c.dir_z};
    double org[3] = {c.position[0], c.position[1], c.position[2]};
    if (c.thin_lens) {
        const double half = (double)((float)c.aperture / 2.f);
        const uint32_t st = xt_stream(key, 0);
        double fo[3] = {dir[0], dir[1], dir[2]};
        fo[0] += xt_unit64(xt_draw(st, 0)) * ((half - -half)) + -half;
        fo[1] += xt_unit64(xt_draw(st, 1)) * ((half - -half)) + -half;
        const double l = sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
        double nd[3] = {dir[0], dir[1], dir[2]};
        if (l != 0.0) {
            nd[0] = dir[0] / l;
            nd[1] = dir[1] / l;
            nd[2] = dir[2] / l;
        }
        const double fl = (double)(float)c.flength;
        const double fp[3] = {dir[0] + fl * nd[0], dir[1] + fl * nd[1], dir[2] + fl * nd[2]};
        dir[0] = fp[0] - fo[0];
        dir[1] = fp[1] - fo[1];
        dir[2] = fp[2] - fo[2];
        org[0] = fo[0];
        org[1] = fo[1];
        org[2] = fo[2];
    }
    const double *m = c.basis;
    double t0 = m[0] * dir[0] + m[1] * dir[1] + m[2] * dir[2] + 0.0;
    double t1 = m[3] * dir[0] + m[4] * dir[1] + m[5] * dir[2] + 0.0;
    double t2 = m[6] * dir[0] + m[7] * dir[1] + m[8] * dir[2] + 0.0;
    const double l = sqrt(t0 * t0 + t1 * t1 + t2 * t2);
    if (l != 0.0) {
        t0 /= l;
        t1 /= l;
        t2 /= l;
    }
    d[0] = t0;
    d[1] = t1;
    d[2] = t2;
    if (c.thin_lens) {
        const double a0 = m[0] * org[0] + m[1] * org[1] + m[2] * org[2] + 0.0;
        const double a1 = m[3] * org[0] + m[4] * org[1] + m[5] * org[2] + 0.0;
        const double a2 = m[6] * org[0] + m[7] * org[1] + m[8] * org[2] + 0.0;
        o[0] = a0 + c.position[0];
        o[1] = a1 + c.position[1];
        o[2] = a2 + c.position[2];
    } else {
        o[0] = org[0];
        o[1] = org[1];
        o[2] = org[2];
    }
}

#define XT_PI_D 3.14159265358979323846  // nmath::PI (lib/nmath/precision.h:101), a double

// v transformed by the orientation matrix m (Vector3f::transformed,
// lib/nmath/vector.inl:517-523; the translation column is 0).
__device__ __forceinline__ void orient(const double *m, double v[3]) {
    const double a = m[0] * v[0] + m[1] * v[1] + m[2] * v[2] + 0.0;
    const double b = m[3] * v[0] + m[4] * v[1] + m[5] * v[2] + 0.0;
    const double c = m[6] * v[0] + m[7] * v[1] + m[8] * v[2] + 0.0;
    v[0] = a;
    v[1] = b;
    v[2] = c;
}

// ERP::get_primary_ray (src/xtcore/camera/erp.cc:11-25): the angles are
// floats, their sines and cosines doubles; the direction is not normalised.
__device__ __forceinline__ void erp_ray(const JobDev &j, float x, float y, double o[3], double d[3]) {
    const xt_camera &c = j.cam;
    const float phi = (float)((double)(x / j.wf) * XT_PI_D * 2.0);
    const float theta = (float)((double)(y / j.hf) * XT_PI_D);
    const double st = sin((double)theta);
    d[0] = st * cos((double)phi);
    d[1] = cos((double)theta);
    d[2] = st * sin((double)phi);
    orient(c.basis, d);
    o[0] = c.position[0];
    o[1] = c.position[1];
    o[2] = c.position[2];
}

// ODS::get_primary_ray (src/xtcore/camera/ods.cc:15-33): the top half of the
// image is the left eye (origin offset -ipd/2 along the tangent), the
// bottom half the right eye; the direction is not normalised.
__device__ __forceinline__ void ods_ray(const JobDev &j, float x, float y, double o[3], double d[3]) {
    const xt_camera &c = j.cam;
    const float h = j.hf * .5f;
    const float scale = (float)c.ipd * .5f * ((y < h) ? -1.f : 1.f);
    if (y > h) y = y - h;
    const float theta = (float)((double)((x / j.wf) * 2.f) * XT_PI_D);
    const float phi = (float)(XT_PI_D * (double).5f - (double)(y / h) * XT_PI_D);
    const double ct = cos((double)theta), s_t = sin((double)theta);
    const double cp = cos((double)phi), sp = sin((double)phi);
    o[0] = ct * (double)scale + c.position[0];
    o[1] = 0.0 * (double)scale + c.position[1];
    o[2] = s_t * (double)scale + c.position[2];
    d[0] = s_t * cp;
    d[1] = sp;
    d[2] = -ct * cp;
    orient(c.basis, d);
}

// Cubemap::get_primary_ray (src/xtcore/camera/cubemap.cc:17-60): a vertical
// strip of six square faces +x -x +y -y +z -z, each face_size = H / 6 rows.
__device__ __forceinline__ void cubemap_ray(const JobDev &j, float x, float y, double o[3], double d[3]) {
    const xt_camera &c = j.cam;
    const uint64_t face_size = (uint64_t)j.hf / 6u;
    const uint32_t face = (uint32_t)((uint64_t)y / face_size);
    double cd[3] = {0, 0, 0}, up[3] = {0, 1, 0};
    switch (face) {
        case 0: cd[0] = 1; break;
        case 1: cd[0] = -1; break;
        case 2: cd[1] = 1; up[1] = 0; up[2] = -1; break;
        case 3: cd[1] = -1; up[1] = 0; up[2] = 1; break;
        case 4: cd[2] = 1; break;
        default: cd[2] = -1; break;
    }
    y -= (float)(face * face_size);
    double v[3] = {(2.f * (double)x / (double)j.wf) - 1.f, (2.f * (double)y / (double)face_size) - 1.f, 1.f};
    double l = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (l != 0.0) {
        v[0] /= l;
        v[1] /= l;
        v[2] /= l;
    }
    // rz = camdir, rx = normalize(up x rz), ry = normalize(rx x rz)
    double rx[3] = {up[1] * cd[2] - up[2] * cd[1], up[2] * cd[0] - up[0] * cd[2], up[0] * cd[1] - up[1] * cd[0]};
    l = sqrt(rx[0] * rx[0] + rx[1] * rx[1] + rx[2] * rx[2]);
    if (l != 0.0) {
        rx[0] /= l;
        rx[1] /= l;
        rx[2] /= l;
    }
    double ry[3] = {rx[1] * cd[2] - rx[2] * cd[1], rx[2] * cd[0] - rx[0] * cd[2], rx[0] * cd[1] - rx[1] * cd[0]};
    l = sqrt(ry[0] * ry[0] + ry[1] * ry[1] + ry[2] * ry[2]);
    if (l != 0.0) {
        ry[0] /= l;
        ry[1] /= l;
        ry[2] /= l;
    }
    const double m[9] = {rx[0], ry[0], cd[0], rx[1], ry[1], cd[1], rx[2], ry[2], cd[2]};
    orient(m, v);
    l = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (l != 0.0) {
        v[0] /= l;
        v[1] /= l;
        v[2] /= l;
    }
    d[0] = v[0];
    d[1] = v[1];
    d[2] = v[2];
    o[0] = c.position[0];
    o[1] = c.position[1];
    o[2] = c.position[2];
}

// ICamera::get_primary_ray of the job's camera kind.
__device__ __forceinline__ void camera_ray(const JobDev &j, float x, float y, uint32_t key, double o[3],
                                           double d[3]) {
    switch (j.cam.kind) {
        case XT_CAMERA_ODS: ods_ray(j, x, y, o, d); break;
        case XT_CAMERA_ERP: erp_ray(j, x, y, o, d); break;
        case XT_CAMERA_CUBEMAP: cubemap_ray(j, x, y, o, d); break;
        default: perspective_ray(j, x, y, key, o, d); break;
    }
}

// Sample coordinates of anti-aliasing sub-sample `a` of pixel (px, py)
// (src/xtcore/aa.cc:7-48).
__device__ __forceinline__ void aa_coords(const JobDev &j, uint32_t px, uint32_t py, uint32_t a, float &cx,
                                          float &cy) {
    if (j.aa_mode == XT_AA_RANDOM) {
        const uint32_t key = xt_path_key(j.seed, j.job_key, py * j.W + px, 0xfff00000u + a);
        const uint32_t st = xt_stream(key, 0);
        const float rx = (float)xt_unit64(xt_draw(st, 0));
        const float ry = (float)xt_unit64(xt_draw(st, 1));
        cx = (float)((double)px + (double)rx);
        cy = (float)((double)py + (double)ry);
    } else if (j.aa <= 1) {
        cx = (float)((double)px + 0.5);
        cy = (float)((double)py + 0.5);
    } else {
        const uint32_t ix = a % j.aa, iy = a / j.aa;
        cx = (float)((double)px + (double)ix * (double)j.aa_sub + (double)j.aa_off);
        cy = (float)((double)py + (double)iy * (double)j.aa_sub + (double)j.aa_off);
    }
}
#pragma clang fp contract(on)

// Slots [0, spp_pass * qn): path path_base + sl * qn + qi renders sample
// pass_base + sl of the pixel at tile position q0 + qi (a whole pass has
// q0 = 0, qn = NT); path state, radiance and FP64 rays are indexed by path id.
// The queue entry that carries it follows the ray order below. Slots without a sample (pixels past
// the image edge, samples past the job's count) get a dead ray, which the
// trace turns into a miss whose radiance is never accumulated. The path
// state word ps1.w holds the query count (low 16 bits) and the job's rdepth
// (high 16 bits).
static __global__ __launch_bounds__(XT_BLOCK) void k_camera(const JobDev j, uint32_t pass_base, uint32_t spp_pass,
                                                     uint32_t q0, uint32_t qn, uint32_t path_base, Queue out,
                                                     float4 *ps0, float4 *ps1, float4 *rad, double *ray64,
                                                     uint32_t order) {
    const uint32_t n = spp_pass * qn;
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t qword = 1u | (j.rdepth << 16);
    // order 2 (with whole tiles): queue entry i holds sample sl of slot w of
    // tile t, i = (t * spp_pass + sl) * 64 + w, so that a stretch of the
    // queue covers one part of the image; order 3: the pass's samples of one
    // pixel follow each other, i = qi * spp_pass + sl; otherwise the pass's
    // samples follow each other over the whole image. The path id (and so the
    // radiance slot and the accumulation order) is sl * qn + q - q0 either way.
    const bool tiles = order == 2 && (qn & 63u) == 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        uint32_t sl, qi;
        if (order == 3) {
            qi = i / spp_pass;
            sl = i - qi * spp_pass;
        } else if (tiles) {
            const uint32_t t = i / (64u * spp_pass), r = i - t * 64u * spp_pass;
            sl = r >> 6;
            qi = t * 64u + (r & 63u);
        } else {
            sl = i / qn;
            qi = i - sl * qn;
        }
        const uint32_t path = path_base + sl * qn + qi;
        const uint32_t q = q0 + qi;
        const uint32_t s = pass_base + sl;
        uint32_t x = 0, y = 0;
        const bool live = slot_pixel(j, q, x, y) && s < j.total;
        rad[path] = make_float4(0.f, 0.f, 0.f, 0.f);
        out.path[i] = path;
        if (!live) {
            out.r0[i] = make_float4(0.f, 0.f, 0.f, __uint_as_float(XT_DEAD));
            out.r1[i] = make_float4(0.f, 0.f, 1.f, 0.f);
            continue;
        }
        double o[3], d[3];
        const uint32_t pixel = y * j.W + x;
        const uint32_t key = xt_path_key(j.seed, j.job_key, pixel, s);
        float cx, cy;
        aa_coords(j, x, y, s / j.spp, cx, cy);
        camera_ray(j, cx, cy, key, o, d);
        ps0[path] = make_float4(1.f, 1.f, 1.f, __uint_as_float(key));
        ps1[path] = make_float4(1.f, 1.f, 1.f, __uint_as_float(qword));
        if (ray64) {
            double *r = ray64 + 6 * (size_t)path;
            r[0] = o[0];
            r[1] = o[1];
            r[2] = o[2];
            r[3] = d[0];
            r[4] = d[1];
            r[5] = d[2];
        }
        out.r0[i] = make_float4((float)o[0], (float)o[1], (float)o[2], __uint_as_float(XT_NONE));
        out.r1[i] = make_float4((float)d[0], (float)d[1], (float)d[2], 0.f);
    }
}

// --------------------------------------------------------------- trace

__device__ __forceinline__ float tsmall_of(f3 o) {
    return XT_TSMALL * (1.f + fmaxf(fabsf(o.x), fmaxf(fabsf(o.y), fabsf(o.z))));
}

__device__ __forceinline__ void seed_best(const RayF &r, Best &b) {
    b.t = INFINITY;
    b.prim = XT_NONE;
    b.mark_t = INFINITY;
    b.c0 = b.c1 = XT_NONE;
    if (r.self_t > 0.f && r.origin != XT_NONE) {
        b.t = r.self_t;
        b.prim = r.origin;
    }
}

// First ray of segment `s` of `nseg` over a queue of n rays; segments start
// on multiples of a wave's chunk, segment nseg starts at n.
__device__ __forceinline__ uint32_t seg_begin(uint32_t n, uint32_t s, uint32_t nseg) {
    if (s >= nseg) return n;
    constexpr uint32_t G = 64u * XT_TRACE_CHUNK;
    return (uint32_t)((uint64_t)n * s / nseg) / G * G;
}

// Persistent trace: each wave takes XT_TRACE_CHUNK x 64 consecutive rays at
// a time from a device cursor (one atomic per chunk) until the queue is
// drained, so waves that drew short traversals pick up more work instead of
// idling. With nseg = 8 the queue is cut into 8 contiguous segments with a
// cursor each; workgroups are dispatched to the 8 XCDs round-robin, so the
// waves of XCD (block % 8) drain that segment first and then help with the
// others. Rays next to each other in a queue come from neighbouring pixels,
// so each XCD's L2 then holds the nodes and leaves of one part of the image
// rather than of all of it.
template <int ACCEL, int STACK, int BS>
static __global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(6))) void k_trace(const SceneView S, const Queue in, const uint32_t *in_count,
                                                     float4 *hit, uint32_t *retest_list, uint4 *recheck_list,
                                                     Counters *cnt, uint32_t nseg) {
    __shared__ uint2 kstk[ACCEL == XT_ACCEL_KDTREE ? (STACK + 1) * BS : 1];
    __shared__ uint32_t ostk[ACCEL == XT_ACCEL_OCTREE ? XT_OCT_STACK * BS : 1];
    const uint32_t n = *in_count;
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&cnt->queries, (unsigned long long)n);
    const uint32_t lane = lane_id();
    uint32_t chunk = 0, part = XT_TRACE_CHUNK;
    uint32_t seg = blockIdx.x % nseg, tried = 0;
    for (;;) {
        if (part == XT_TRACE_CHUNK) {  // next chunk of XT_TRACE_CHUNK x 64 rays for this wave
            if (lane == 0) {
                chunk = n;
                // every segment is tried at most once after it ran dry, so
                // the loop ends
                while (tried < nseg) {
                    const uint32_t lo = seg_begin(n, seg, nseg), hi = seg_begin(n, seg + 1, nseg);
                    const uint32_t c = atomicAdd(&cnt->seg[seg], 64u * XT_TRACE_CHUNK);
                    if (c < hi - lo) {
                        chunk = lo + c;
                        break;
                    }
                    ++tried;
                    seg = seg + 1 == nseg ? 0 : seg + 1;
                }
            }
            chunk = __shfl(chunk, 0);
            part = 0;
        }
        const uint32_t base = chunk + 64u * part++;
        if (base >= n) {
            if (chunk >= n) break;
            part = XT_TRACE_CHUNK;
            continue;
        }
        const uint32_t i = base + lane;
        bool mark = false, full = false;
        uint32_t c0 = XT_NONE, c1 = XT_NONE;
        if (i < n && __float_as_uint(in.r0[i].w) == XT_DEAD) {
            hit[i] = make_float4(INFINITY, __uint_as_float(XT_NONE), 0.f, 0.f);
        } else if (i < n) {
            const float4 a = in.r0[i], b = in.r1[i];
            RayF r;
            r.o = mk(a.x, a.y, a.z);
            r.d = mk(b.x, b.y, b.z);
            r.origin = __float_as_uint(a.w);
            r.self_t = b.w;
            Tester32 tst{S, r, tsmall_of(r.o), {}, make_shear(r.d)};
            seed_best(r, tst.best);
            for (uint32_t k = 0; k < S.n_planes; ++k)
                if (XT_PLANE_BASE + k != r.origin) plane32(r.o, r.d, tst.tsmall, S.planes[k], XT_PLANE_BASE + k, tst.best);
            bool ok;
            if (ACCEL == XT_ACCEL_KDTREE) ok = walk_kd<STACK>(S, tst, r.o, r.d, kstk + threadIdx.x, BS, 1e-6f);
            else ok = walk_oct(S, tst, r.o, r.d, ostk + threadIdx.x, BS, 1e-6f);
            Best &bb = tst.best;
            // a deferred candidate matters only up to the closest hit (a miss
            // has t = INF, so compare only when something was deferred)
            if (!ok) atomicAdd(&cnt->trace_fail, 1ull);
            mark = !ok || (bb.mark_t != INFINITY && bb.mark_t <= bb.t * (1.f + XT_TIE));
            full = !ok || bb.c1 == XT_CAND_OVERFLOW;
            c0 = bb.c0;
            c1 = bb.c1;
            float ht = bb.t, hu = 0.f, hv = 0.f;
            if (!mark && bb.prim < S.n_prims && !(r.self_t > 0.f && bb.prim == r.origin)) {
                const float4 a = S.prims[3 * bb.prim], b = S.prims[3 * bb.prim + 1];
                // the closest triangle's t and barycentrics in FP64
                if (XT_PRIM_TYPE(__float_as_uint(a.w)) == XT_PRIM_TRIANGLE)
                    tri_hit64(r.o, r.d, a, b, S.prims[3 * bb.prim + 2], ht, hu, hv);
            }
            hit[i] = make_float4(ht, __uint_as_float(bb.prim), hu, hv);
        }
        // rays with a few deferred candidates get the FP64 re-check of just
        // those; rays with more, or whose traversal gave up, the full re-test
        const bool rf = mark && full, rc = mark && !full;
        const uint32_t slot = wave_append(rf, &cnt->n[CNT_RETEST]);
        if (rf) retest_list[slot] = i;
        const uint32_t cslot = wave_append(rc, &cnt->n[CNT_RECHECK]);
        if (rc) recheck_list[cslot] = make_uint4(i, c0, c1, XT_NONE);
    }
}

__device__ __forceinline__ uint32_t object_of(const SceneView &S, uint32_t p) {
    return p >= XT_PLANE_BASE ? S.planes[p - XT_PLANE_BASE].object : __float_as_uint(S.prims[3 * p + 1].w);
}

// FP64 re-check of a ray's FP32 closest hit and its deferred candidates: the
// reference's predicates (triangle.cc:82-123, sphere.cc:35-61,
// plane.cc:19-61) on those primitives alone, ordered by (t, object, soup
// index) like Scene::intersection. The FP32 traversal stopped at its closest
// hit, so when the FP64 winner lies beyond it (or the FP32 closest hit is
// rejected and nothing else hits) the ray goes on to the full re-test.
static __global__ __launch_bounds__(XT_BLOCK) void k_recheck(const SceneView S, const Queue in, float4 *hit,
                                                        const uint4 *list, uint32_t *retest_list, Counters *cnt,
                                                        const double *ray64, uint32_t *flags) {
    const uint32_t n = cnt->n[CNT_RECHECK];
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {
        const uint32_t k = base + threadIdx.x;
        bool need = false;
        uint32_t i = 0;
        if (k < n) {
            const uint4 e = list[k];
            i = e.x;
            const float4 a = in.r0[i], b = in.r1[i];
            const uint32_t origin = __float_as_uint(a.w);
            d3 o = {a.x, a.y, a.z}, d = {b.x, b.y, b.z};
            if (ray64) {
                const double *q = ray64 + 6 * (size_t)in.path[i];
                o = {q[0], q[1], q[2]};
                d = {q[3], q[4], q[5]};
            }
            const float4 h = hit[i];
            const uint32_t B = __float_as_uint(h.y);
            Tester64 tst{S, RayF{}, o, d, INFINITY, XT_NONE, XT_NONE, 0, 0, 0, mk(a.x, a.y, a.z), Shear{}};
            if (b.w > 0.f && origin != XT_NONE) tst.offer64(b.w, origin, object_of(S, origin), nullptr);
            const uint32_t cand[4] = {B, e.y, e.z, e.w};
            for (int c = 0; c < 4; ++c) {
                const uint32_t p = cand[c];
                if (p == XT_NONE || p == origin) continue;
                bool seen = false;
                for (int c2 = 0; c2 < c; ++c2) seen = seen || cand[c2] == p;
                if (seen) continue;
                double t, bc[3] = {0, 0, 0};
                bool tri = false;
                if (p >= XT_PLANE_BASE) {
                    t = plane64(o, d, S.planes[p - XT_PLANE_BASE]);
                } else if (XT_PRIM_TYPE(__float_as_uint(S.prims[3 * p].w)) == XT_PRIM_TRIANGLE) {
                    t = tri64(o, d, S.prims64[p], bc);
                    tri = true;
                } else {
                    t = sph64(o, d, S.prims64[p]);
                }
                if (t > 0.0) tst.offer64(t, p, object_of(S, p), tri ? bc : nullptr);
            }
            // the FP32 traversal explored everything up to its closest hit (all
            // of space when it found none)
            if (B != XT_NONE && !(tst.bt <= (double)h.x * (1.0 + 4.0 * XT_TIE))) {
                need = true;
            } else {
                const bool is_tri = tst.prim < S.n_prims;
                hit[i] = make_float4(tst.prim == XT_NONE ? INFINITY : (float)tst.bt, __uint_as_float(tst.prim),
                                     is_tri ? (float)tst.b1 : 0.f, is_tri ? (float)tst.b2 : 0.f);
                if (flags) flags[i] |= 1u;
            }
        }
        const uint32_t slot = wave_append(need, &cnt->n[CNT_RETEST]);
        if (need) retest_list[slot] = i;
    }
}

// FP64 re-test of listed rays. ray64 (optional) holds the FP64 rays by path
// id; otherwise the FP32 ray is widened.
template <int ACCEL>
static __global__ __launch_bounds__(ACCEL == XT_ACCEL_KDTREE ? XT_BLOCK : XT_OCT_BLOCK) void k_retest(
    const SceneView S, const Queue in, float4 *hit, const uint32_t *list, Counters *cnt, const double *ray64,
    uint32_t *flags) {
    constexpr int BS = ACCEL == XT_ACCEL_KDTREE ? XT_BLOCK : XT_OCT_BLOCK;
    __shared__ uint2 kstk[ACCEL == XT_ACCEL_KDTREE ? (XT_KD_STACK + 1) * XT_BLOCK : 1];
    __shared__ uint32_t ostk[ACCEL == XT_ACCEL_OCTREE ? XT_OCT_STACK * XT_OCT_BLOCK : 1];
    const uint32_t n = cnt->n[CNT_RETEST];
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&cnt->retests, (unsigned long long)n);
    const uint32_t stride = gridDim.x * BS;
    for (uint32_t base = blockIdx.x * BS; base < n; base += stride) {
        const uint32_t k = base + threadIdx.x;
        if (k >= n) continue;
        const uint32_t i = list[k];
        const float4 a = in.r0[i], b = in.r1[i];
        RayF r;
        r.o = mk(a.x, a.y, a.z);
        r.d = mk(b.x, b.y, b.z);
        r.origin = __float_as_uint(a.w);
        r.self_t = b.w;
        Tester64 tst{S, r, {r.o.x, r.o.y, r.o.z}, {r.d.x, r.d.y, r.d.z}, INFINITY, XT_NONE, XT_NONE, 0, 0, 0,
                     r.o, Shear{}};
        if (ray64) {
            const double *q = ray64 + 6 * (size_t)in.path[i];
            tst.o = {q[0], q[1], q[2]};
            tst.d = {q[3], q[4], q[5]};
        }
        if (r.self_t > 0.f && r.origin != XT_NONE) {
            tst.bt = r.self_t;
            tst.prim = r.origin;
            tst.obj = r.origin >= XT_PLANE_BASE ? S.planes[r.origin - XT_PLANE_BASE].object
                                                : __float_as_uint(S.prims[3 * r.origin + 1].w);
        }
        for (uint32_t p = 0; p < S.n_planes; ++p) {
            if (XT_PLANE_BASE + p == r.origin) continue;
            const double t = plane64(tst.o, tst.d, S.planes[p]);
            if (t > 0.0) tst.offer64(t, XT_PLANE_BASE + p, S.planes[p].object, nullptr);
        }
        const f3 o32 = mk((float)tst.o.x, (float)tst.o.y, (float)tst.o.z);
        const f3 d32 = mk((float)tst.d.x, (float)tst.d.y, (float)tst.d.z);
        tst.o32 = o32;
        tst.sh = make_shear(d32);
        bool ok;
        if (ACCEL == XT_ACCEL_KDTREE) ok = walk_kd<XT_KD_STACK>(S, tst, o32, d32, kstk + threadIdx.x, BS, 1e-4f);
        else ok = walk_oct(S, tst, o32, d32, ostk + threadIdx.x, BS, 1e-4f);
        if (!ok) atomicAdd(&cnt->steps_exceeded, 1ull);
        const bool is_tri = tst.prim < S.n_prims;
        hit[i] = make_float4(tst.prim == XT_NONE ? INFINITY : (float)tst.bt, __uint_as_float(tst.prim),
                             is_tri ? (float)tst.b1 : 0.f, is_tri ? (float)tst.b2 : 0.f);
        if (flags) flags[i] |= 1u;
    }
}

// --------------------------------------------------------------- shading

// Surface data of a hit: point, shading normal N (unnormalised, as the
// reference hands it on) and a geometric normal g of any length.
__device__ __forceinline__ void surface(const SceneView &S, uint32_t prim, f3 o, f3 d, float t, float u, float v,
                                        f3 &p, f3 &N, f3 &g, uint32_t &obj) {
    p = add(o, scl(d, t));
    if (prim >= XT_PLANE_BASE) {
        const XtPlane &pl = S.planes[prim - XT_PLANE_BASE];
        N = mk(pl.normal[0], pl.normal[1], pl.normal[2]);
        g = N;
        obj = pl.object;
        return;
    }
    const float4 a = S.prims[3 * prim], b = S.prims[3 * prim + 1];
    obj = __float_as_uint(b.w);
    const uint32_t tf = __float_as_uint(a.w);
    if (XT_PRIM_TYPE(tf) == XT_PRIM_SPHERE) {
        N = scl(sub(p, mk(a.x, a.y, a.z)), b.z);  // (p - c) / r
        g = N;
        return;
    }
    const float4 c = S.prims[3 * prim + 2];
    const f3 v0 = mk(a.x, a.y, a.z);
    g = cross3(sub(mk(b.x, b.y, b.z), v0), sub(mk(c.x, c.y, c.z), v0));
    N = norm3(g);
    if ((XT_PRIM_FLAGS(tf) & XT_PF_VNORMALS) && S.vnormals) {
        const float *vn = S.vnormals + 9 * (size_t)prim;
        const float w = 1.f - u - v;
        const f3 pn = mk(vn[0] * w + vn[3] * u + vn[6] * v, vn[1] * w + vn[4] * u + vn[7] * v,
                         vn[2] * w + vn[5] * u + vn[8] * v);
        if (dot3(pn, pn) != 0.f) N = pn;
    }
}

// t of the reference's immediate re-hit of the surface a ray leaves from:
// the origin is p + N * 1e-8, so a direction into the surface meets it again
// at t = -1e-8 (g.N) / (g.d) when that is >= 1e-8 (lambert.cc:35,
// ao/integrator.cc:38, and the t < EPSILON tests of the primitives).
__device__ __forceinline__ float self_hit_t(f3 g, f3 N, f3 d) {
    const float gd = dot3(g, d);
    if (gd == 0.f) return 0.f;
    const float r = -dot3(g, N) / gd;
    return r >= 1.f ? 1e-8f * r : 0.f;
}

__device__ __forceinline__ void push_ray(Queue &out, uint32_t slot, f3 o, f3 d, uint32_t origin, float self_t,
                                         uint32_t path) {
    out.r0[slot] = make_float4(o.x, o.y, o.z, __uint_as_float(origin));
    out.r1[slot] = make_float4(d.x, d.y, d.z, self_t);
    out.path[slot] = path;
}

#pragma clang fp contract(off)
// Texture coordinates of a hit in FP64 (hitrecord.h:33), as the reference's
// primitives set them: a triangle's corner texcoords weighted by its
// barycentrics (triangle.cc:110-114), a sphere's asin mapping of the normal
// (sphere.cc:53-54), a plane's projection on two tangents, scaled and
// wrapped (plane.cc:43-57). o, d: the ray; h: its hit record.
__device__ __forceinline__ void hit_texcoord(const SceneView &S, uint32_t prim, const d3 o, const d3 d, const float4 h,
                                             double &tu, double &tv) {
    const double PI = 3.14159265358979323846;  // nmath::PI
    tu = 0.0;
    tv = 0.0;
    if (prim >= XT_PLANE_BASE) {
        const XtPlane &pl = S.planes[prim - XT_PLANE_BASE];
        const double us = S.obj_uv[2 * pl.object], vs = S.obj_uv[2 * pl.object + 1];
        double t = plane64(o, d, pl);
        if (!(t > 0.0)) t = (double)h.x;
        const d3 pt = dadd(o, dscl(d, t));
        const d3 v = {pl.point64[0], pl.point64[1], pl.point64[2]};
        const d3 nn = dnorm(d3{pl.normal64[0], pl.normal64[1], pl.normal64[2]});
        const d3 uvec = {nn.y, nn.z, -nn.x};
        const d3 vvec = dcross(uvec, nn);
        const d3 vp = dadd(v, pt);
        tu = ddot(uvec, vp) * us;
        tv = ddot(vvec, vp) * vs;
        if (tu > 1.f) tu -= (float)(int)tu;
        if (tv > 1.f) tv -= (float)(int)tv;
        if (tu < -1.f) tu -= (float)(int)tu;
        if (tv < -1.f) tv -= (float)(int)tv;
        if (tu < 0.f) tu = 1.f + tu;
        if (tv < 0.f) tv = 1.f + tv;
    } else {
        const XtPrim64 &p = S.prims64[prim];
        if (XT_PRIM_TYPE(__float_as_uint(S.prims[3 * prim].w)) == XT_PRIM_TRIANGLE) {
            double bc[3];
            const double t = tri64(o, d, p, bc);
            if (!(t > 0.0)) {  // the FP64 predicate disagrees: FP32 barycentrics
                bc[1] = h.z;
                bc[2] = h.w;
                bc[0] = 1.0 - bc[1] - bc[2];
            }
            const double *tc = S.tc64 + 6 * (size_t)prim;
            tu = tc[0] * bc[0] + tc[2] * bc[1] + tc[4] * bc[2];
            tv = tc[1] * bc[0] + tc[3] * bc[1] + tc[5] * bc[2];
        } else {
            const uint32_t obj = __float_as_uint(S.prims[3 * prim + 1].w);
            const double us = S.obj_uv[2 * obj], vs = S.obj_uv[2 * obj + 1];
            const d3 c = {p.a[0], p.a[1], p.a[2]};
            const double r = p.b[0];
            const double b = 2 * ddot(dsub(o, c), d);
            const d3 nc = {-c.x, -c.y, -c.z};
            const double cq = ddot(c, c) + ddot(o, o) + 2 * ddot(nc, o) - r * r;
            const double disc = (b * b - 4 * cq);
            const double s = sqrt(disc > 0.0 ? disc : 0.0);
            const double t1 = (-b + s) / 2.0, t2 = (-b - s) / 2.0;
            double t = t1 < t2 ? t1 : t2;
            if (!(disc > 0.0 && t > XT_EPS64)) t = (double)h.x;
            const d3 pt = dadd(o, dscl(d, t));
            const double sg = t1 * t2 > 0. ? 1. : -1;
            const d3 nn = {(pt.x - c.x) / r * sg, (pt.y - c.y) / r * sg, (pt.z - c.z) / r * sg};
            tu = asin(nn.x / (us != 0.0f ? us : 1.0f)) / PI + 0.5;
            tv = asin(nn.y / (vs != 0.0f ? vs : 1.0f)) / PI + 0.5;
        }
    }
}
#pragma clang fp contract(on)

// Pathtracer (src/xtcore/integrator/pathtracer/integrator.cc:21-47). Path
// state: T (product so far) and F (factor of the last hit); L = T * X where
// X is the value of the current query: miss -> env, emissive -> F * Le,
// continue -> F * f * X'. Materials: lambert.cc:30-40, phong.cc:35-55,
// blinnphong.cc:34-61, emissive.cc:20-27. A dielectric (dielectric.cc:24-32)
// leaves its hit_result's intensity at ColorRGBf's default 0, so the value
// of the query is 0 whatever the refracted ray returns: the path ends with
// L = 0 (the reference traces on and multiplies by 0).
// Material channels of a sampled scene's pathtracer queue, ahead of
// k_shade_pt: for queue slot i, chan[3i .. 3i+2] = the hit material's kd, ks
// and ke (its samplers read at the hit's FP64 texture coordinates, or its
// constant colours), or for a miss chan[3i] = the environment along the ray.
// The samplers' FP64 code more than doubles a kernel's registers, so it runs
// here and k_shade_pt keeps the register budget of an unsampled scene.
static __global__ __launch_bounds__(XT_BLOCK) void k_channels(const SceneView S, const Queue in,
                                                       const uint32_t *in_count, const float4 *hit, float4 *chan) {
    const uint32_t n = *in_count;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const float4 h = hit[i];
        const float4 a = in.r0[i], b = in.r1[i];
        const uint32_t prim = __float_as_uint(h.y);
        if (prim == XT_NONE) {
            const f3 e = env_sample(S, mk(b.x, b.y, b.z));
            chan[3 * (size_t)i] = make_float4(e.x, e.y, e.z, 0.f);
            continue;
        }
        const uint32_t obj = prim >= XT_PLANE_BASE ? S.planes[prim - XT_PLANE_BASE].object
                                                   : __float_as_uint(S.prims[3 * prim + 1].w);
        const uint32_t mi = S.obj_material[obj];
        const XtMaterial &m = S.materials[mi == XT_NONE ? 0 : mi];
        f3 kd = mk(m.kd[0], m.kd[1], m.kd[2]), ks = mk(m.ks[0], m.ks[1], m.ks[2]);
        f3 ke = mk(m.ke[0], m.ke[1], m.ke[2]);
        const uint32_t *ms = S.mat_smp + 3 * (mi == XT_NONE ? 0 : mi);
        const uint32_t s0 = ms[0], s1 = ms[1], s2 = ms[2];
        if (s0 != XT_NONE || s1 != XT_NONE || s2 != XT_NONE) {
            double tu, tv;
            hit_texcoord(S, prim, d3{a.x, a.y, a.z}, d3{b.x, b.y, b.z}, h, tu, tv);
            // one copy of the sampler code, run once per sampled channel
#pragma unroll 1
            for (int c = 0; c < 3; ++c) {
                const uint32_t si = c == 0 ? s0 : (c == 1 ? s1 : s2);
                if (si == XT_NONE) continue;
                const f3 v = sampler_eval(S, si, tu, tv, 0.0);
                kd = c == 0 ? v : kd;
                ks = c == 1 ? v : ks;
                ke = c == 2 ? v : ke;
            }
        }
        chan[3 * (size_t)i] = make_float4(kd.x, kd.y, kd.z, 0.f);
        chan[3 * (size_t)i + 1] = make_float4(ks.x, ks.y, ks.z, 0.f);
        chan[3 * (size_t)i + 2] = make_float4(ke.x, ke.y, ke.z, 0.f);
    }
}

// One shading step of a pathtracer path at its hit h (ray a, b): the closed
// form of pathtracer/integrator.cc:21-47 with the path state s0 = (T, key),
// s1 = (F, query word). A miss ends the path with L = T env, an emissive hit
// with L = T F Le, a dielectric with 0, a continuing material at query
// rdepth with 0; otherwise the path continues (cont) with the next ray
// (no, nd, nprim, nself) and s0 / s1 updated. chan: the hit's three channel
// colours (SAMPLED only).
template <bool SAMPLED>
__device__ __forceinline__ void pt_step(const SceneView &S, const float4 h, const float4 a, const float4 b, float4 &s0,
                                        float4 &s1, const float4 *chan, f3 &L, bool &cont, f3 &no, f3 &nd,
                                        uint32_t &nprim, float &nself) {
    const f3 o = mk(a.x, a.y, a.z), d = mk(b.x, b.y, b.z);
    const f3 T = mk(s0.x, s0.y, s0.z), F = mk(s1.x, s1.y, s1.z);
    const uint32_t key = __float_as_uint(s0.w), qword = __float_as_uint(s1.w);
    const uint32_t q = qword & 0xffffu, rdepth = qword >> 16;
    const uint32_t prim = __float_as_uint(h.y);
    cont = false;
    L = mk(0.f, 0.f, 0.f);
    if (prim == XT_NONE) {
        f3 e;
        if (SAMPLED) {
            const float4 c = chan[0];
            e = mk(c.x, c.y, c.z);
        } else {
            e = env_constant(S.env, d);
        }
        L = cmul(T, e);
        return;
    }
    f3 p, N, g;
    uint32_t obj;
    surface(S, prim, o, d, h.x, h.z, h.w, p, N, g, obj);
    const uint32_t mi = S.obj_material[obj];
    const XtMaterial m = S.materials[mi == XT_NONE ? 0 : mi];
    // channel colours: the material's constants, or its samplers at the hit's
    // texture coordinates (material.cc:65-70)
    f3 kd = mk(m.kd[0], m.kd[1], m.kd[2]), ks = mk(m.ks[0], m.ks[1], m.ks[2]);
    f3 ke = mk(m.ke[0], m.ke[1], m.ke[2]);
    if (SAMPLED) {
        const float4 c0 = chan[0], c1 = chan[1], c2 = chan[2];
        kd = mk(c0.x, c0.y, c0.z);
        ks = mk(c1.x, c1.y, c1.z);
        ke = mk(c2.x, c2.y, c2.z);
    }
    if (m.type == XT_MAT_EMISSIVE) {
        L = cmul(cmul(T, F), ke);
    } else if (m.type == XT_MAT_DIELECTRIC) {
        // ends with zero intensity
    } else if (q < rdepth) {
        Draws r{xt_stream(key, q), 0};
        f3 f, dir;
        if (m.type == XT_MAT_LAMBERT) {
            dir = norm3(smp_diffuse(N, r));
            f = cscale(kd, dot3(N, dir));
        } else {
            const float k = r.next();
            if (k > m.reflectance) {
                f = kd;
                dir = smp_diffuse(N, r);
            } else {
                f = ks;
                dir = smp_lobe(N, scl(d, -1.f), m.exponent, r);
                if (m.type == XT_MAT_BLINNPHONG) dir = norm3(dir);
            }
            f = cscale(f, 1.f);
        }
        const f3 T2 = cmul(cmul(T, F), f);
        s0 = make_float4(T2.x, T2.y, T2.z, s0.w);
        s1 = make_float4(f.x, f.y, f.z, __uint_as_float(qword + 1));
        cont = true;
        no = p;
        nd = dir;
        nprim = prim;
        nself = self_hit_t(g, N, dir);
    }
    // q == rdepth with a continuing material: eval(0) == 0
}

// Pathtracer shading over a queue with path state by path id: a path that
// ends writes its radiance to rad[path]; continuing paths are appended to
// the next queue (one atomic per block).
template <bool SAMPLED>
static __global__ __launch_bounds__(XT_BLOCK) void k_shade_pt(const SceneView S, const Queue in,
                                                       const uint32_t *in_count, const float4 *hit, Queue out,
                                                       uint32_t *out_count, float4 *ps0, float4 *ps1, float4 *rad,
                                                       const float4 *chan) {
    __shared__ uint32_t lds[XT_BLOCK / 64 + 1];
    const uint32_t n = *in_count;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {
        const uint32_t i = base + threadIdx.x;
        bool cont = false;
        f3 no = mk(0, 0, 0), nd = mk(0, 0, 1);
        uint32_t nprim = XT_NONE, path = 0;
        float nself = 0.f;
        if (i < n) {
            path = in.path[i];
            float4 s0 = ps0[path], s1 = ps1[path];
            f3 L;
            pt_step<SAMPLED>(S, hit[i], in.r0[i], in.r1[i], s0, s1, SAMPLED ? chan + 3 * (size_t)i : nullptr, L,
                             cont, no, nd, nprim, nself);
            if (cont) {
                ps0[path] = s0;
                ps1[path] = s1;
            } else {
                rad[path] = make_float4(L.x, L.y, L.z, 0.f);
            }
        }
        const uint32_t slot = block_append<XT_BLOCK>(cont, out_count, lds);
        if (cont) push_ray(out, slot, no, nd, nprim, nself, path);
    }
}

// AO, first query (src/xtcore/integrator/ao/integrator.cc:12-53): a hit
// spawns one hemisphere ray; a miss scores w = 1.
static __global__ __launch_bounds__(XT_BLOCK) void k_shade_ao1(const SceneView S, const Queue in,
                                                        const uint32_t *in_count, const float4 *hit, Queue out,
                                                        uint32_t *out_count, const float4 *ps0, float4 *rad) {
    __shared__ uint32_t lds[XT_BLOCK / 64 + 1];
    const uint32_t n = *in_count;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {
        const uint32_t i = base + threadIdx.x;
        bool cont = false;
        f3 no = mk(0, 0, 0), nd = mk(0, 0, 1);
        uint32_t nprim = XT_NONE, path = 0;
        float nself = 0.f;
        if (i < n) {
            path = in.path[i];
            const float4 h = hit[i];
            const uint32_t prim = __float_as_uint(h.y);
            if (prim == XT_NONE) {
                rad[path] = make_float4(1.f, 1.f, 1.f, 0.f);
            } else {
                const float4 a = in.r0[i], b = in.r1[i];
                const f3 o = mk(a.x, a.y, a.z), d = mk(b.x, b.y, b.z);
                f3 p, N, g;
                uint32_t obj;
                surface(S, prim, o, d, h.x, h.z, h.w, p, N, g, obj);
                Draws r{xt_stream(__float_as_uint(ps0[path].w), 1), 0};
                const f3 dir = smp_hemisphere(N, N, r);
                cont = true;
                no = p;
                nd = dir;
                nprim = prim;
                nself = self_hit_t(g, N, dir);
            }
        }
        const uint32_t slot = block_append<XT_BLOCK>(cont, out_count, lds);
        if (cont) push_ray(out, slot, no, nd, nprim, nself, path);
    }
}

static __global__ __launch_bounds__(XT_BLOCK) void k_shade_ao2(const Queue in, const uint32_t *in_count,
                                                        const float4 *hit, float4 *rad) {
    const uint32_t n = *in_count;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint32_t path = in.path[i];
        const float4 h = hit[i];
        float w = 1.f;
        if (__float_as_uint(h.y) != XT_NONE) {
            const float4 b = in.r1[i];
            const float dist = h.x * sqrtf(b.x * b.x + b.y * b.y + b.z * b.z);
            w = fminf(dist, 100.0f) / 100.0f;
        }
        rad[path] = make_float4(w, w, w, 0.f);
    }
}

// Depth (src/xtcore/integrator/depth/integrator.cc:33-39): 1 / log of the
// FP64 hit distance, recomputed from the FP64 camera ray with the
// reference's predicate for the primitive found, rounded to float.
static __global__ __launch_bounds__(XT_BLOCK) void k_shade_depth(const SceneView S, const Queue in,
                                                          const uint32_t *in_count, const float4 *hit,
                                                          double *ray64, float4 *rad) {
    const uint32_t n = *in_count;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint32_t path = in.path[i];
        const float4 h = hit[i];
        const uint32_t prim = __float_as_uint(h.y);
        if (prim == XT_NONE) {
            rad[path] = make_float4(0.f, 0.f, 0.f, 0.f);
            continue;
        }
        const double *q = ray64 + 6 * (size_t)path;
        const d3 o = {q[0], q[1], q[2]}, d = {q[3], q[4], q[5]};
        double t;
        if (prim >= XT_PLANE_BASE) {
            t = plane64(o, d, S.planes[prim - XT_PLANE_BASE]);
        } else if (XT_PRIM_TYPE(__float_as_uint(S.prims[3 * prim].w)) == XT_PRIM_TRIANGLE) {
            double bc[3];
            t = tri64(o, d, S.prims64[prim], bc);
        } else {
            t = sph64(o, d, S.prims64[prim]);
        }
        if (!(t > 0.0)) t = (double)h.x;
        const d3 pt = dadd(o, dscl(d, t));
        const double dist = dlen(dsub(o, pt));
        const float depth = (float)(1. / log(dist));
        rad[path] = make_float4(depth, 1.f, 0.f, 0.f);  // y = 1 marks a hit
        // the hit point replaces the direction in the path's FP64 ray, for
        // the ray graph's sample (the camera ray is not read after this)
        double *w = ray64 + 6 * (size_t)path;
        w[3] = pt.x;
        w[4] = pt.y;
        w[5] = pt.z;
    }
}

#pragma clang fp contract(off)
// Normal (src/xtcore/integrator/normal/integrator.cc:26-36): the hit record's
// normal in FP64 as the reference forms it. Triangles interpolate their
// vertex normals with the area-sum barycentrics and fall back to the
// normalised geometric normal when that sum has zero length
// (triangle.cc:110-113); spheres give (p - c) / r (sphere.cc:55-58); planes
// their normal as given. The normal goes to the FP64 ray slot of the path
// (the camera ray is not needed after this query); rad.y marks a hit.
static __global__ __launch_bounds__(XT_BLOCK) void k_shade_normal(const SceneView S, const Queue in,
                                                           const uint32_t *in_count, const float4 *hit,
                                                           double *ray64, float4 *rad) {
    const uint32_t n = *in_count;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint32_t path = in.path[i];
        const float4 h = hit[i];
        const uint32_t prim = __float_as_uint(h.y);
        if (prim == XT_NONE) {
            rad[path] = make_float4(0.f, 0.f, 0.f, 0.f);
            continue;
        }
        double *q = ray64 + 6 * (size_t)path;
        const d3 o = {q[0], q[1], q[2]}, d = {q[3], q[4], q[5]};
        d3 nn;
        if (prim >= XT_PLANE_BASE) {
            const XtPlane &pl = S.planes[prim - XT_PLANE_BASE];
            nn = {pl.normal64[0], pl.normal64[1], pl.normal64[2]};
        } else {
            const XtPrim64 &p = S.prims64[prim];
            if (XT_PRIM_TYPE(__float_as_uint(S.prims[3 * prim].w)) == XT_PRIM_TRIANGLE) {
                double bc[3];
                double t = tri64(o, d, p, bc);
                if (!(t > 0.0)) {  // the FP64 predicate disagrees: FP32 barycentrics
                    bc[1] = h.z;
                    bc[2] = h.w;
                    bc[0] = 1.0 - bc[1] - bc[2];
                }
                const d3 A = {p.a[0], p.a[1], p.a[2]};
                const d3 geo = dnorm(dcross(dsub(d3{p.b[0], p.b[1], p.b[2]}, A), dsub(d3{p.c[0], p.c[1], p.c[2]}, A)));
                nn = geo;
                if (S.vnormals) {
                    const float *vn = S.vnormals + 9 * (size_t)prim;
                    const d3 pn = dadd(dadd(dscl(d3{vn[0], vn[1], vn[2]}, bc[0]), dscl(d3{vn[3], vn[4], vn[5]}, bc[1])),
                                       dscl(d3{vn[6], vn[7], vn[8]}, bc[2]));
                    if (dlen(pn) != 0.0) nn = pn;
                }
            } else {
                double t = sph64(o, d, p);
                if (!(t > 0.0)) t = (double)h.x;
                const d3 pt = dadd(o, dscl(d, t));
                const double r = p.b[0];
                nn = {(pt.x - p.a[0]) / r, (pt.y - p.a[1]) / r, (pt.z - p.a[2]) / r};
            }
        }
        q[0] = nn.x;
        q[1] = nn.y;
        q[2] = nn.z;
        rad[path] = make_float4(0.f, 1.f, 0.f, 0.f);
    }
}

// The normal integrator's per-sample update of the stored pixel
// (normal/integrator.cc:24-37): acc = rgb * 2 - 1; on a hit
// acc = (normalize((n + acc) * 0.5) + 1) * 0.5; stored as ColorRGBAf(acc, 1),
// whose constructor clamps negative channels to 0.
__device__ __forceinline__ float4 normal_update(float4 a, bool hit, d3 n) {
    d3 acc = {(double)a.x * 2.0 - 1.0, (double)a.y * 2.0 - 1.0, (double)a.z * 2.0 - 1.0};
    if (hit) {
        const d3 m = dscl(dadd(n, acc), 0.5);
        const d3 u = dnorm(m);
        acc = {(u.x + 1.0) * 0.5, (u.y + 1.0) * 0.5, (u.z + 1.0) * 0.5};
    }
    return make_float4(fmaxf((float)acc.x, 0.f), fmaxf((float)acc.y, 0.f), fmaxf((float)acc.z, 0.f), 1.f);
}

// The uv integrator's shading: the hit's texture coordinates (hit_texcoord)
// into the path's ray64 slot.
static __global__ __launch_bounds__(XT_BLOCK) void k_shade_uv(const SceneView S, const Queue in,
                                                       const uint32_t *in_count, const float4 *hit, double *ray64,
                                                       float4 *rad) {
    const uint32_t n = *in_count;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint32_t path = in.path[i];
        const float4 h = hit[i];
        const uint32_t prim = __float_as_uint(h.y);
        if (prim == XT_NONE) {
            rad[path] = make_float4(0.f, 0.f, 0.f, 0.f);
            continue;
        }
        double *q = ray64 + 6 * (size_t)path;
        double tu, tv;
        hit_texcoord(S, prim, d3{q[0], q[1], q[2]}, d3{q[3], q[4], q[5]}, h, tu, tv);
        q[0] = tu;
        q[1] = tv;
        rad[path] = make_float4(0.f, 1.f, 0.f, 0.f);
    }
}

// The uv integrator's per-sample update of the stored pixel
// (uv/integrator.cc:24-40): acc = rgb; on a hit acc += (texcoord * 0.5 +
// 0.5) * weight with texcoord.z = 0, and the alpha becomes
// alpha + weight * (hit ? weight : 0); stored as ColorRGBAf(acc, 0) (negative
// channels to 0) with the alpha setter's clamp to [0, 1].
__device__ __forceinline__ float4 uv_update(float4 a, bool hit, double tu, double tv, double w) {
    double ax = (double)a.x, ay = (double)a.y, az = (double)a.z;
    const double alpha = (double)a.w;
    double alpha_sample = 0.0;
    if (hit) {
        ax = ax + (tu * 0.5 + 0.5) * w;
        ay = ay + (tv * 0.5 + 0.5) * w;
        az = az + (0.0 * 0.5 + 0.5) * w;
        alpha_sample += w;
    }
    const float fx = (float)ax, fy = (float)ay, fz = (float)az, fa = (float)(alpha + w * alpha_sample);
    return make_float4(fx < 0.f ? 0.f : fx, fy < 0.f ? 0.f : fy, fz < 0.f ? 0.f : fz,
                       fa < 0.f ? 0.f : (fa > 1.f ? 1.f : fa));
}
#pragma clang fp contract(on)

// Emission (src/xtcore/integrator/emission/integrator.cc:26-35): the hit
// object's emissive colour; rad.w marks a hit on an object with a material.
static __global__ __launch_bounds__(XT_BLOCK) void k_shade_emission(const SceneView S, const Queue in,
                                                             const uint32_t *in_count, const float4 *hit,
                                                             const double *ray64, float4 *rad) {
    const uint32_t n = *in_count;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint32_t path = in.path[i];
        const uint32_t prim = __float_as_uint(hit[i].y);
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (prim != XT_NONE) {
            const uint32_t obj = prim >= XT_PLANE_BASE ? S.planes[prim - XT_PLANE_BASE].object
                                                       : __float_as_uint(S.prims[3 * prim + 1].w);
            const uint32_t mi = S.obj_material[obj];
            if (mi != XT_NONE) {
                const XtMaterial &m = S.materials[mi];
                v = make_float4(m.ke[0], m.ke[1], m.ke[2], 1.f);
                const uint32_t sm = S.mat_smp[3 * mi + 2];
                if (sm != XT_NONE) {  // a sampled emissive channel, at the hit's texture coordinates
                    const double *q = ray64 + 6 * (size_t)path;
                    double tu, tv;
                    hit_texcoord(S, prim, d3{q[0], q[1], q[2]}, d3{q[3], q[4], q[5]}, hit[i], tu, tv);
                    const f3 c = sampler_eval(S, sm, tu, tv, 0.0);
                    v = make_float4(c.x, c.y, c.z, 1.f);
                }
            }
        }
        rad[path] = v;
    }
}

// Stencil (src/xtcore/integrator/stencil/integrator.cc:28-30).
static __global__ __launch_bounds__(XT_BLOCK) void k_shade_stencil(const Queue in, const uint32_t *in_count,
                                                            const float4 *hit, float4 *rad) {
    const uint32_t n = *in_count;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint32_t path = in.path[i];
        const bool h = __float_as_uint(hit[i].y) != XT_NONE;
        rad[path] = make_float4(h ? 1.f : 0.f, h ? 1.f : 0.f, 0.f, 0.f);
    }
}

// --------------------------------------------------------------- accumulation

#pragma clang fp contract(off)
// A sample's value times its weight in FP64, and its addition to a
// Neumaier-compensated sum (sum, compensation), without contraction so that
// every kernel that accumulates samples rounds them alike.
__device__ __forceinline__ double weighted(float v, double w) { return (double)v * w; }
__device__ __forceinline__ void neumaier_add(double2 &a, double x) {
    const double t = a.x + x;
    if (fabs(a.x) >= fabs(x)) a.y += (a.x - t) + x;
    else a.y += (x - t) + a.x;
    a.x = t;
}
#pragma clang fp contract(on)

// Per pixel of tile positions [q0, q0 + qn), samples of this pass in sample
// order (radiance of sample pass_base + sl at rad[sl * qn + q - q0]).
// REFERENCE reproduces the float read-modify-write of each integrator
// (tile_t::read/write, lib/nimg/color.cc:35-40 clamps); FP64 keeps a
// Neumaier sum per channel.
static __global__ __launch_bounds__(XT_BLOCK) void k_accumulate(const JobDev j, uint32_t pass_base, uint32_t spp_pass,
                                                         uint32_t q0, uint32_t qn, const float4 *rad,
                                                         const double *nrm64, float4 *accf, double2 *acc64) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t qi = blockIdx.x * blockDim.x + threadIdx.x; qi < qn; qi += stride) {
        const uint32_t q = q0 + qi;
        uint32_t x, y;
        if (!slot_pixel(j, q, x, y)) continue;
        const uint32_t pix = y * j.W + x;
        float4 a = accf[pix];
        double2 s0 = acc64[3 * (size_t)pix], s1 = acc64[3 * (size_t)pix + 1], s2 = acc64[3 * (size_t)pix + 2];
        for (uint32_t sl = 0; sl < spp_pass; ++sl) {
            if (pass_base + sl >= j.total) break;
            const float4 v = rad[sl * qn + qi];
            if (j.accum == XT_ACCUM_REFERENCE) {
                if (j.integrator == XT_INTEGRATOR_NORMAL) {
                    const double *q = nrm64 + 6 * ((size_t)sl * qn + qi);
                    a = normal_update(a, v.y != 0.f, d3{q[0], q[1], q[2]});
                } else if (j.integrator == XT_INTEGRATOR_UV) {
                    const double *q = nrm64 + 6 * ((size_t)sl * qn + qi);
                    a = uv_update(a, v.y != 0.f, q[0], q[1], j.weight);
                } else if (j.integrator == XT_INTEGRATOR_EMISSION) {
                    // acc += emissive * float(weight) (ColorRGBf operators clamp at 0); alpha 1
                    if (v.w != 0.f) {
                        a.x = fmaxf(a.x + fmaxf(v.x * j.weight_f, 0.f), 0.f);
                        a.y = fmaxf(a.y + fmaxf(v.y * j.weight_f, 0.f), 0.f);
                        a.z = fmaxf(a.z + fmaxf(v.z * j.weight_f, 0.f), 0.f);
                    }
                    a.w = 1.f;
                } else if (j.integrator == XT_INTEGRATOR_STENCIL || j.integrator == XT_INTEGRATOR_WIREFRAME) {
                    if (v.x != 0.f) {
                        const float na = (float)((double)a.w + j.weight);
                        a = make_float4(1.f, 1.f, 1.f, fminf(fmaxf(na, 0.f), 1.f));
                    }
                } else if (j.integrator == XT_INTEGRATOR_DEPTH) {
                    if (v.y != 0.f) {
                        const float r = fmaxf((float)((double)a.x + (double)v.x * j.weight), 0.f);
                        a = make_float4(r, r, r, 1.f);
                    } else {
                        a.w = 1.f;
                    }
                } else if (j.integrator == XT_INTEGRATOR_AO) {
                    a.x = fmaxf((float)((double)a.x + (double)v.x * j.weight), 0.f);
                    a.y = fmaxf((float)((double)a.y + (double)v.y * j.weight), 0.f);
                    a.z = fmaxf((float)((double)a.z + (double)v.z * j.weight), 0.f);
                    a.w = 1.f;
                } else {
                    a.x = fmaxf(a.x + fmaxf(v.x * j.weight_f, 0.f), 0.f);
                    a.y = fmaxf(a.y + fmaxf(v.y * j.weight_f, 0.f), 0.f);
                    a.z = fmaxf(a.z + fmaxf(v.z * j.weight_f, 0.f), 0.f);
                    a.w = 1.f;
                }
            } else {
                double xs[3] = {weighted(v.x, j.weight), weighted(v.y, j.weight), weighted(v.z, j.weight)};
                if (j.integrator == XT_INTEGRATOR_STENCIL || j.integrator == XT_INTEGRATOR_WIREFRAME) xs[1] = xs[2] = xs[0];
                if (j.integrator == XT_INTEGRATOR_EMISSION && v.w == 0.f) xs[0] = xs[1] = xs[2] = 0.0;
                if (j.integrator == XT_INTEGRATOR_DEPTH && v.y == 0.f) xs[0] = xs[1] = xs[2] = 0.0;
                if (j.integrator == XT_INTEGRATOR_DEPTH) xs[1] = xs[2] = xs[0];
                neumaier_add(s0, xs[0]);
                neumaier_add(s1, xs[1]);
                neumaier_add(s2, xs[2]);
            }
        }
        accf[pix] = a;
        acc64[3 * (size_t)pix] = s0;
        acc64[3 * (size_t)pix + 1] = s1;
        acc64[3 * (size_t)pix + 2] = s2;
    }
}

static __global__ __launch_bounds__(XT_BLOCK) void k_resolve(const JobDev j, const float4 *accf, const double2 *acc64,
                                                      float4 *out) {
    const uint32_t n = j.W * j.H;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += stride) {
        if (j.accum == XT_ACCUM_REFERENCE) {
            out[p] = accf[p];
            continue;
        }
        float c[3];
        for (int k = 0; k < 3; ++k) {
            const double2 s = acc64[3 * (size_t)p + k];
            c[k] = fmaxf((float)(s.x + s.y), 0.f);
        }
        if (j.integrator == XT_INTEGRATOR_STENCIL || j.integrator == XT_INTEGRATOR_WIREFRAME) {
            const float a = fminf(c[0], 1.f);
            const float rgb = c[0] > 0.f ? 1.f : 0.f;
            out[p] = make_float4(rgb, rgb, rgb, a);
        } else {
            out[p] = make_float4(c[0], c[1], c[2], 1.f);
        }
    }
}

#pragma clang fp contract(off)
// ColorRGBf arithmetic (lib/nimg/color.inl): float components, every result
// clamped at 0 by the setters.
__device__ __forceinline__ float cclamp(float v) { return v < 0.f ? 0.f : v; }
__device__ __forceinline__ f3 cmulc(f3 a, f3 b) { return mk(cclamp(a.x * b.x), cclamp(a.y * b.y), cclamp(a.z * b.z)); }
__device__ __forceinline__ f3 cmuls(f3 a, float s) { return mk(cclamp(a.x * s), cclamp(a.y * s), cclamp(a.z * s)); }
__device__ __forceinline__ f3 caddc(f3 a, f3 b) { return mk(cclamp(a.x + b.x), cclamp(a.y + b.y), cclamp(a.z + b.z)); }
// nmath::max (lib/nmath/mutil.inl:41): x > y ? x : y
__device__ __forceinline__ double nmax(double x, double y) { return x > y ? x : y; }

// IMaterial::shade of one query (material/lambert.cc:9-28, phong.cc:11-33,
// blinnphong.cc:9-32, emissive.cc:8-18, dielectric.cc:9-21): FP64 vectors,
// float colours, the exponent's power in FP64 (the reference's long double
// power rounded to float). Channel colours are the material's constants or
// its samplers at the texture coordinate.
static __global__ __launch_bounds__(XT_BLOCK) void k_shade_direct(const SceneView S, const xt_shade_query *qs,
                                                           uint32_t n, float *inout, int32_t *ret) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const xt_shade_query q = qs[i];
        const uint32_t mi = q.object < S.n_objects ? S.obj_material[q.object] : XT_NONE;
        if (mi == XT_NONE) {
            ret[i] = -1;
            continue;
        }
        const XtMaterial m = S.materials[mi];
        const uint32_t *ms = S.mat_smp + 3 * mi;
        f3 ch[3] = {mk(m.kd[0], m.kd[1], m.kd[2]), mk(m.ks[0], m.ks[1], m.ks[2]), mk(m.ke[0], m.ke[1], m.ke[2])};
        for (int c = 0; c < 3; ++c)
            if (ms[c] != XT_NONE) ch[c] = sampler_eval(S, ms[c], q.texcoord[0], q.texcoord[1], q.texcoord[2]);
        // both colours are ColorRGBf in the reference: clamped at 0 already
        f3 I = mk(cclamp(inout[3 * i]), cclamp(inout[3 * i + 1]), cclamp(inout[3 * i + 2]));
        const f3 E = mk(cclamp(q.intensity[0]), cclamp(q.intensity[1]), cclamp(q.intensity[2]));
        const d3 P = {q.point[0], q.point[1], q.point[2]}, N = {q.normal[0], q.normal[1], q.normal[2]};
        const d3 C = {q.camera[0], q.camera[1], q.camera[2]}, L = {q.emitter[0], q.emitter[1], q.emitter[2]};
        int r = 1;
        if (m.type == XT_MAT_LAMBERT) {
            const d3 ld = dnorm(dsub(L, P));
            const double dd = ddot(ld, N);
            if (dd > 0) I = caddc(I, cmulc(cmuls(E, (float)dd), ch[0]));
        } else if (m.type == XT_MAT_PHONG || m.type == XT_MAT_BLINNPHONG) {
            const d3 ld = dnorm(dsub(L, P));
            const d3 ray = dnorm(dsub(C, P));
            double dd, rmv;
            if (m.type == XT_MAT_PHONG) {
                dd = nmax(0, ddot(ld, N));
                // light_dir.reflected(normal) (vector.inl:455-460), normalised
                const d3 ii = dnorm(ld), nn = dnorm(N);
                const d3 rf = dnorm(dsub(dscl(nn, 2 * ddot(ii, nn)), ii));
                rmv = nmax(0, ddot(rf, ray));
            } else {
                dd = nmax(ddot(ld, N), 0);
                const d3 h = dnorm(dadd(ld, ray));
                rmv = nmax(ddot(h, N), 0);
            }
            const float pw = (float)pow(rmv, (double)m.exponent);
            const f3 term = caddc(cmuls(ch[0], (float)dd), cmuls(ch[1], pw));
            if (m.type == XT_MAT_PHONG) I = caddc(I, cmulc(E, term));
            else I = cmulc(E, term);
        } else if (m.type == XT_MAT_EMISSIVE) {
            I = ch[2];
            r = 0;
        } else {  // dielectric
            I = mk(0.f, 0.f, 0.f);
        }
        inout[3 * i] = I.x;
        inout[3 * i + 1] = I.y;
        inout[3 * i + 2] = I.z;
        ret[i] = r;
    }
}
#pragma clang fp contract(on)

// Generic-query helpers (xt_trace): pack SoA rays into the queue layout and
// unpack hits.
static __global__ void k_pack_rays(const float *ox, const float *oy, const float *oz, const float *dx, const float *dy,
                            const float *dz, Queue q, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        q.r0[i] = make_float4(ox[i], oy[i], oz[i], __uint_as_float(XT_NONE));
        q.r1[i] = make_float4(dx[i], dy[i], dz[i], 0.f);
        q.path[i] = i;
    }
}

static __global__ void k_unpack_hits(const float4 *hit, float *t, uint32_t *prim, float *u, float *v, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float4 h = hit[i];
        t[i] = h.x;
        prim[i] = __float_as_uint(h.y);
        u[i] = h.z;
        v[i] = h.w;
    }
}
