// Device half of the C ABI: resident scenes, closest-hit queries and the
// wavefront render loop on one HIP device and stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "capi_internal.h"
#include "device/xt_kernels.h"
#include "xt_device.h"

#define XT_TRY(expr)                                                                  \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess) {                                                       \
            d->err = std::string(#expr) + ": " + hipGetErrorString(e_);               \
            return XT_E_DEVICE;                                                       \
        }                                                                             \
    } while (0)

#define XT_LAUNCHED()                                                                 \
    do {                                                                              \
        hipError_t e_ = hipGetLastError();                                            \
        if (e_ != hipSuccess) {                                                       \
            d->err = std::string("kernel launch: ") + hipGetErrorString(e_);          \
            return XT_E_DEVICE;                                                       \
        }                                                                             \
    } while (0)

namespace {

template <class T>
int dev_alloc(xt_dev *d, T **p, size_t count, std::vector<void *> *track, uint64_t *bytes) {
    *p = nullptr;
    if (count == 0) count = 1;
    const size_t b = count * sizeof(T);
    if (d->budget && d->used + b > d->budget) {
        d->err = "HBM budget exceeded";
        return XT_E_BUDGET;
    }
    hipError_t e = hipMalloc((void **)p, b);
    if (e != hipSuccess) {
        d->err = std::string("hipMalloc: ") + hipGetErrorString(e);
        return XT_E_DEVICE;
    }
    d->used += b;
    if (bytes) *bytes += b;
    if (track) track->push_back(*p);
    return XT_OK;
}

template <class T>
int upload(xt_dev *d, T **p, const std::vector<T> &v, std::vector<void *> *track, uint64_t *bytes) {
    int rc = dev_alloc(d, p, v.size(), track, bytes);
    if (rc) return rc;
    if (!v.empty()) XT_TRY(hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return XT_OK;
}


}  // namespace

void xt_dev_free_wave(xt_dev *d) {
    for (void *p : d->wave.allocs) hipFree(p);
    d->wave.allocs.clear();
    d->used -= std::min<uint64_t>(d->used, d->wave.bytes);
    d->wave = WaveBuffers();
}

void xt_dev_free_accum(xt_dev *d) {
    for (void *p : d->accum.allocs) hipFree(p);
    d->accum.allocs.clear();
    d->used -= std::min<uint64_t>(d->used, d->accum.bytes);
    d->accum = AccumBuffers();
}

int xt_dev_ensure_wave(xt_dev *d, size_t cap, bool need64, bool need_chan) {
    if (d->wave.cap >= cap && (!need64 || d->wave.ray64) && (!need_chan || d->wave.chan)) return XT_OK;
    need64 = need64 || d->wave.ray64;  // keep what the buffers had
    need_chan = need_chan || d->wave.chan;
    cap = std::max(cap, d->wave.cap);
    xt_dev_free_wave(d);
    WaveBuffers &w = d->wave;
    std::vector<void *> *t = &w.allocs;
    uint64_t *b = &w.bytes;
    int rc = 0;
    for (int k = 0; k < 2 && !rc; ++k) {
        rc = rc ? rc : dev_alloc(d, &w.q[k].r0, cap, t, b);
        rc = rc ? rc : dev_alloc(d, &w.q[k].r1, cap, t, b);
        rc = rc ? rc : dev_alloc(d, &w.q[k].path, cap, t, b);
    }
    rc = rc ? rc : dev_alloc(d, &w.hit, cap, t, b);
    rc = rc ? rc : dev_alloc(d, &w.retest, cap, t, b);
    rc = rc ? rc : dev_alloc(d, &w.recheck, cap, t, b);
    rc = rc ? rc : dev_alloc(d, &w.ps0, cap, t, b);
    rc = rc ? rc : dev_alloc(d, &w.ps1, cap, t, b);
    rc = rc ? rc : dev_alloc(d, &w.rad, cap, t, b);
    rc = rc ? rc : dev_alloc(d, &w.flags, cap, t, b);
    rc = rc ? rc : dev_alloc(d, &w.cnt, XT_MAX_LANES, t, b);
    if (!rc && need64) rc = dev_alloc(d, &w.ray64, 6 * cap, t, b);
    if (!rc && need_chan) rc = dev_alloc(d, &w.chan, 3 * cap, t, b);
    if (rc) {
        xt_dev_free_wave(d);
        return rc;
    }
    w.cap = cap;
    return XT_OK;
}

int xt_dev_ensure_accum(xt_dev *d, size_t npix) {
    if (d->accum.npix >= npix) return XT_OK;
    xt_dev_free_accum(d);
    AccumBuffers &a = d->accum;
    int rc = dev_alloc(d, &a.accf, npix, &a.allocs, &a.bytes);
    rc = rc ? rc : dev_alloc(d, &a.acc64, 3 * npix, &a.allocs, &a.bytes);
    rc = rc ? rc : dev_alloc(d, &a.out, npix, &a.allocs, &a.bytes);
    if (rc) {
        xt_dev_free_accum(d);
        return rc;
    }
    a.npix = npix;
    return XT_OK;
}

extern "C" {

// Scratch of xt_trace_device_static: counters, then the re-check and re-test
// lists of XT_TRACE_STATIC_MAX_RAYS rays. In host code the array's name is not
// a device address: xt_dev_create looks its address up per device.
struct StaticTraceScratch {
    Counters cnt;
    uint32_t pad[64 - sizeof(Counters) / 4];
    uint4 recheck[XT_TRACE_STATIC_MAX_RAYS];
    uint32_t retest[XT_TRACE_STATIC_MAX_RAYS];
};
static_assert(sizeof(Counters) % 4 == 0 && sizeof(Counters) < 256, "scratch header");
__device__ __align__(256) StaticTraceScratch scratch;

int xt_dev_create(xt_dev **out, int hip_device, uint64_t hbm_budget_bytes) {
    if (!out) return XT_E_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return XT_E_DEVICE;
    if (hip_device < 0 || hip_device >= n) return XT_E_INVALID;
    xt_dev *d = new xt_dev;
    d->device = hip_device;
    d->budget = hbm_budget_bytes;
    if (hipSetDevice(hip_device) != hipSuccess || hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking) != hipSuccess) {
        delete d;
        return XT_E_DEVICE;
    }
    hipEventCreate(&d->ev0);
    hipEventCreate(&d->ev1);
    for (int i = 0; i < XT_TRACE_EVENTS; ++i) hipEventCreate(&d->tev[i]);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, hip_device) == hipSuccess) {
        d->n_cu = prop.multiProcessorCount;
        // default budget: 90% of the device memory
        if (!d->budget) d->budget = (uint64_t)(0.9 * (double)prop.totalGlobalMem);
    }
    // this device's copy of the static scratch (with the device current)
    void *sp = nullptr;
    if (hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)) == hipSuccess) d->scratch = sp;
    *out = d;
    return XT_OK;
}

void xt_dev_destroy(xt_dev *d) {
    if (!d) return;
    hipSetDevice(d->device);
    hipStreamSynchronize(d->stream);
    xt_serve_shutdown(d);
    for (auto &kv : d->scenes)
        for (void *p : kv.second.allocs) hipFree(p);
    xt_dev_free_wave(d);
    xt_dev_free_accum(d);
    if (d->host_counters) hipHostFree(d->host_counters);
    if (d->host_snap) hipHostFree(d->host_snap);
    for (auto &row : d->snap_ev)
        for (hipEvent_t e : row)
            if (e) hipEventDestroy(e);
    for (int k = 0; k < XT_LANE_STREAMS; ++k) {
        if (d->lane_stream[k]) hipStreamDestroy(d->lane_stream[k]);
        if (d->lane_ev[k]) hipEventDestroy(d->lane_ev[k]);
    }
    if (d->fork_ev) hipEventDestroy(d->fork_ev);
    hipEventDestroy(d->ev0);
    hipEventDestroy(d->ev1);
    for (int i = 0; i < XT_TRACE_EVENTS; ++i) hipEventDestroy(d->tev[i]);
    hipStreamDestroy(d->stream);
    delete d;
}

const char *xt_last_error(const xt_dev *d) { return d ? d->err.c_str() : "null device"; }

int xt_dev_set_tuning(xt_dev *d, const xt_tuning *t) {
    if (!d || !t) return XT_E_INVALID;
    xt_tuning n = d->tuning;
    if (t->kd_stack) {
        if (t->kd_stack != 8 && t->kd_stack != 12 && t->kd_stack != 16 && t->kd_stack != 24) {
            d->err = "kd_stack must be 8, 12, 16 or 24";
            return XT_E_INVALID;
        }
        n.kd_stack = t->kd_stack;
    }
    if (t->trace_block) {
        if (t->trace_block != 128 && t->trace_block != 256) {
            d->err = "trace_block must be 128 or 256";
            return XT_E_INVALID;
        }
        n.trace_block = t->trace_block;
    }
    if (t->wave_paths) {
        if (t->wave_paths < 4096 || t->wave_paths > (1u << 28)) {
            d->err = "wave_paths must lie in [4096, 2^28]";
            return XT_E_INVALID;
        }
        n.wave_paths = t->wave_paths;
    }
    if (t->trace_blocks_per_cu) {
        if (t->trace_blocks_per_cu == XT_TUNING_AUTO) {
            n.trace_blocks_per_cu = 0;
        } else if (t->trace_blocks_per_cu > 64) {
            d->err = "trace_blocks_per_cu must be at most 64";
            return XT_E_INVALID;
        } else {
            n.trace_blocks_per_cu = t->trace_blocks_per_cu;
        }
    }
    if (t->trace_segments) {
        if (t->trace_segments != 1 && t->trace_segments != XT_TRACE_SEGMENTS) {
            d->err = "trace_segments must be 1 or 8";
            return XT_E_INVALID;
        }
        n.trace_segments = t->trace_segments;
    }
    if (t->ray_order) {
        if (t->ray_order > 3) {
            d->err = "ray_order must be 1, 2 or 3";
            return XT_E_INVALID;
        }
        n.ray_order = t->ray_order;
    }
    if (t->kd_leaf) {
        if (t->kd_leaf > 2) {
            d->err = "kd_leaf must be 1 or 2";
            return XT_E_INVALID;
        }
        n.kd_leaf = t->kd_leaf;
    }
    if (t->kd_layout) {
        if (t->kd_layout > 2) {
            d->err = "kd_layout must be 1 or 2";
            return XT_E_INVALID;
        }
        n.kd_layout = t->kd_layout;
    }
    if (t->frame_lanes) {
        if (t->frame_lanes > XT_LANE_STREAMS) {
            d->err = "frame_lanes must be 1, 2, 3 or 4";
            return XT_E_INVALID;
        }
        n.frame_lanes = t->frame_lanes;
    }
    d->tuning = n;
    return XT_OK;
}

int xt_dev_get_tuning(const xt_dev *d, xt_tuning *t) {
    if (!d || !t) return XT_E_INVALID;
    *t = d->tuning;
    return XT_OK;
}

int xt_scene_upload(xt_dev *d, const xt_scene *s, uint32_t *scene_id) {
    if (!d || !s || !scene_id) return XT_E_INVALID;
    if (!s->built) {
        d->err = "scene not built";
        return XT_E_STATE;
    }
    XT_TRY(hipSetDevice(d->device));
    const xt::FlatScene &f = s->flat;
    if (f.planes.size() >= 0xfffdu || f.prims.size() >= 0xffff0000u) {
        d->err = "too many primitives for the hit id space";
        return XT_E_SCENE;
    }
    DevScene ds;
    std::vector<void *> *t = &ds.allocs;
    uint64_t *b = &ds.bytes;
    std::vector<uint32_t> obj_mat(f.objects.size());
    ds.missing_material = false;
    for (size_t i = 0; i < f.objects.size(); ++i) {
        obj_mat[i] = f.objects[i].material;
        if (f.objects[i].material == XT_NO_MATERIAL && f.objects[i].geom_type != 0xffffffffu) ds.missing_material = true;
    }
    ds.pt_supported = f.env_supported && !ds.missing_material;
    for (size_t i = 0; i < f.objects.size(); ++i) {
        uint32_t m = f.objects[i].material;
        if (m != XT_NO_MATERIAL && !f.material_supported[m]) ds.pt_supported = ds.materials_supported = false;
    }
    // kd leaves' primitives copied in leaf order, the object word replaced
    // by the soup index, so the FP32 traversal reads a leaf's records
    // without going through the reference list (a tree left to the device
    // build is made after the other arrays are resident)
    std::vector<XtPrim> leaf(f.kd_deferred ? 0 : f.kd.refs.size());
    for (size_t r = 0; r < f.kd.refs.size(); ++r) {
        leaf[r] = f.prims[f.kd.refs[r]];
        leaf[r].object = f.kd.refs[r];
    }
    XtPrim *prims = nullptr, *kdl = nullptr;
    XtPrim64 *p64 = nullptr;
    XtPlane *planes = nullptr;
    XtKdNode *kd = nullptr;
    uint32_t *kdr = nullptr, *octr = nullptr, *om = nullptr;
    XtOctNode *oct = nullptr;
    XtMaterial *mats = nullptr;
    float *vn = nullptr;
    double *tc64 = nullptr, *obj_uv = nullptr;
    float *texels = nullptr;
    XtTexture *texs = nullptr;
    XtSampler *smps = nullptr;
    uint32_t *msmp = nullptr;
    std::vector<XtMaterial> mv = f.materials;
    if (mv.empty()) mv.push_back(XtMaterial{});
    std::vector<uint32_t> ms = f.mat_smp;
    ms.resize(3 * mv.size(), XT_NONE);  // the stand-in material samples nothing
    int rc = upload(d, &prims, f.prims, t, b);
    rc = rc ? rc : upload(d, &p64, f.prims64, t, b);
    rc = rc ? rc : upload(d, &planes, f.planes, t, b);
    if (!f.kd_deferred) {
        const std::vector<XtKdNode> kdn = d->tuning.kd_layout == 2 ? xt::kd_line_layout(f.kd.nodes) : f.kd.nodes;
        rc = rc ? rc : upload(d, &kd, kdn, t, b);
        rc = rc ? rc : upload(d, &kdr, f.kd.refs, t, b);
        rc = rc ? rc : upload(d, &kdl, leaf, t, b);
        const size_t nr = std::max<size_t>(1, f.kd.refs.size());
        ds.kd_bytes = sizeof(XtKdNode) * std::max<size_t>(1, kdn.size()) + 4 * nr + sizeof(XtPrim) * nr;
        ds.n_kd_nodes = (uint32_t)std::max<size_t>(1, kdn.size());
        ds.n_kd_refs = (uint32_t)f.kd.refs.size();
    }
    rc = rc ? rc : upload(d, &oct, f.oct_nodes, t, b);
    rc = rc ? rc : upload(d, &octr, f.oct_refs, t, b);
    rc = rc ? rc : upload(d, &om, obj_mat, t, b);
    rc = rc ? rc : upload(d, &mats, mv, t, b);
    if (!rc && (f.tc64.size() != 6 * f.prims.size() || f.obj_uv_scale.size() != 2 * f.objects.size())) {
        d->err = "texture coordinate arrays do not match the primitives";
        rc = XT_E_SCENE;
    }
    rc = rc ? rc : upload(d, &tc64, f.tc64, t, b);
    rc = rc ? rc : upload(d, &obj_uv, f.obj_uv_scale, t, b);
    rc = rc ? rc : upload(d, &texels, f.texels, t, b);
    rc = rc ? rc : upload(d, &texs, f.textures, t, b);
    rc = rc ? rc : upload(d, &smps, f.samplers, t, b);
    rc = rc ? rc : upload(d, &msmp, ms, t, b);
    if (!rc && !f.vnormals.empty()) rc = upload(d, &vn, f.vnormals, t, b);
    if (rc) {
        for (void *p : ds.allocs) hipFree(p);
        d->used -= std::min<uint64_t>(d->used, ds.bytes);
        return rc;
    }
    SceneView &v = ds.view;
    std::memset(&v, 0, sizeof v);
    v.prims = (const float4 *)prims;
    v.prims64 = p64;
    v.planes = planes;
    v.kd = kd;
    v.kd_refs = kdr;
    v.kd_leaf = (const float4 *)kdl;
    v.oct = oct;
    v.oct_refs = octr;
    v.obj_material = om;
    v.materials = mats;
    v.vnormals = vn;
    v.tc64 = tc64;
    v.obj_uv = obj_uv;
    v.texels = (const float4 *)texels;
    v.textures = texs;
    v.samplers = smps;
    v.mat_smp = msmp;
    v.env = f.env;
    for (int a = 0; a < 3; ++a) {
        v.kb_lo[a] = f.kd.bounds.lo[a];
        v.kb_hi[a] = f.kd.bounds.hi[a];
    }
    v.n_prims = (uint32_t)f.prims.size();
    v.n_planes = (uint32_t)f.planes.size();
    v.n_oct = (uint32_t)f.oct_nodes.size();
    v.n_objects = (uint32_t)f.objects.size();
    v.sampled = f.samplers.empty() ? 0u : 1u;
    if (!f.kd_deferred && f.kd.nodes.empty()) {  // empty soup: a root leaf with no refs
        XtKdNode leaf;
        leaf.word = 3u;
        leaf.count = 0;
        XT_TRY(hipMemcpy(kd, &leaf, sizeof leaf, hipMemcpyHostToDevice));
    }
    if (f.kd_deferred) {
        xt_build_params bp{f.kd_params.max_depth, f.kd_params.leaf_prims, f.kd_params.cost_intersect, 0, 0};
        rc = xt_kd_build_into(d, ds, &bp);
        if (rc) {
            (void)hipStreamSynchronize(d->stream);
            for (void *p : ds.allocs) hipFree(p);
            d->used -= std::min<uint64_t>(d->used, ds.bytes);
            return rc;
        }
    }
    const uint32_t id = d->next_scene++;
    d->scenes[id] = std::move(ds);
    *scene_id = id;
    return XT_OK;
}

int xt_scene_release(xt_dev *d, uint32_t scene_id) {
    if (!d) return XT_E_INVALID;
    auto it = d->scenes.find(scene_id);
    if (it == d->scenes.end()) return XT_E_NOTFOUND;
    if (xt_serve_uses_scene(d, scene_id)) {
        d->err = "scene is in use by served jobs";
        return XT_E_STATE;
    }
    hipSetDevice(d->device);
    hipStreamSynchronize(d->stream);
    xt_serve_forget_scene(d, scene_id);  // idle serving lanes set up for it
    for (void *p : it->second.allocs) hipFree(p);
    d->used -= std::min<uint64_t>(d->used, it->second.bytes);
    d->scenes.erase(it);
    return XT_OK;
}

}  // extern "C"

// Instantiations of the kd-tree trace kernel over the tuning space. The
// persistent grid is `per_cu` blocks per CU, or (per_cu == 0) exactly the
// blocks the CU keeps resident, from the occupancy calculator.
template <int STACK, int BS>
static void launch_kd_trace_sb(unsigned per_cu, int n_cu, size_t cap, hipStream_t st, const SceneView &v,
                               const Queue &q, const uint32_t *count, float4 *hit, uint32_t *retest, uint4 *recheck, Counters *cnt, uint32_t nseg) {
    static int resident = 0;
    if (per_cu == 0) {
        if (resident == 0) {
            int nb = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_trace<XT_ACCEL_KDTREE, STACK, BS>, BS, 0) !=
                    hipSuccess ||
                nb < 1)
                nb = 1;
            resident = nb;
        }
        per_cu = (unsigned)resident;
    }
    const size_t need = (cap + BS - 1) / BS;
    const unsigned g = (unsigned)std::max<size_t>(1, std::min<size_t>(need, (size_t)n_cu * per_cu));
    hipLaunchKernelGGL((k_trace<XT_ACCEL_KDTREE, STACK, BS>), dim3(g), dim3(BS), 0, st, v, q, count, hit, retest, recheck,
                       cnt, nseg);
}

template <int STACK>
static void launch_kd_trace_s(unsigned bs, unsigned per_cu, int n_cu, size_t cap, hipStream_t st, const SceneView &v,
                              const Queue &q, const uint32_t *count, float4 *hit, uint32_t *retest, uint4 *recheck, Counters *cnt, uint32_t nseg) {
    if (bs == 128) launch_kd_trace_sb<STACK, 128>(per_cu, n_cu, cap, st, v, q, count, hit, retest, recheck, cnt, nseg);
    else launch_kd_trace_sb<STACK, 256>(per_cu, n_cu, cap, st, v, q, count, hit, retest, recheck, cnt, nseg);
}

static void launch_kd_trace(const xt_tuning &tn, int n_cu, size_t cap, hipStream_t st, const SceneView &view,
                            const Queue &q, const uint32_t *count, float4 *hit, uint32_t *retest, uint4 *recheck, Counters *cnt) {
    const uint32_t nseg = tn.trace_segments == XT_TRACE_SEGMENTS ? XT_TRACE_SEGMENTS : 1u;
    SceneView v = view;
    v.kd_indexed = tn.kd_leaf == 2;
    const unsigned bs = tn.trace_block, pc = tn.trace_blocks_per_cu;
    switch (tn.kd_stack) {
        case 8: launch_kd_trace_s<8>(bs, pc, n_cu, cap, st, v, q, count, hit, retest, recheck, cnt, nseg); break;
        case 12: launch_kd_trace_s<12>(bs, pc, n_cu, cap, st, v, q, count, hit, retest, recheck, cnt, nseg); break;
        case 24: launch_kd_trace_s<24>(bs, pc, n_cu, cap, st, v, q, count, hit, retest, recheck, cnt, nseg); break;
        default: launch_kd_trace_s<16>(bs, pc, n_cu, cap, st, v, q, count, hit, retest, recheck, cnt, nseg); break;
    }
}

// Launches trace + FP64 re-check / re-test over queue `q` (count at `count`)
// into explicit buffers.
int xt_launch_trace_bufs(xt_dev *d, const DevScene &ds, int accel, const TraceBufs &B, const double *ray64,
                         uint32_t *flags, bool time_it, hipStream_t st) {
    const bool timed = time_it && d->trace_launches < XT_TRACE_EVENTS - 1;
    if (timed) hipEventRecord(d->tev[d->trace_launches], st);
    const xt_tuning &tn = d->tuning;
    if (accel == XT_ACCEL_OCTREE) {
        // 16 resident 64-lane blocks per CU (LDS: 20 KB each)
        const unsigned g = (unsigned)std::min<size_t>(B.cap / XT_OCT_BLOCK + 1, (size_t)d->n_cu * 16);
        hipLaunchKernelGGL((k_trace<XT_ACCEL_OCTREE, 1, XT_OCT_BLOCK>), dim3(g), dim3(XT_OCT_BLOCK), 0, st,
                           ds.view, B.q, B.count, B.hit, B.retest, B.recheck, B.cnt,
                           tn.trace_segments == XT_TRACE_SEGMENTS ? XT_TRACE_SEGMENTS : 1u);
    } else {
        launch_kd_trace(tn, d->n_cu, B.cap, st, ds.view, B.q, B.count, B.hit, B.retest, B.recheck, B.cnt);
    }
    if (timed) {
        hipEventRecord(d->tev[d->trace_launches + 1], st);
        d->timed_pairs.push_back(d->trace_launches);
        d->trace_launches += 2;
    }
    XT_LAUNCHED();
    hipLaunchKernelGGL(k_recheck, dim3(grid_for(B.cap / 64 + 1, XT_BLOCK, 256)), dim3(XT_BLOCK), 0, st,
                       ds.view, B.q, B.hit, B.recheck, B.retest, B.cnt, ray64, flags);
    XT_LAUNCHED();
    if (accel == XT_ACCEL_OCTREE) {
        hipLaunchKernelGGL(k_retest<XT_ACCEL_OCTREE>, dim3(grid_for(B.cap / 64 + 1, XT_OCT_BLOCK, 2048)),
                           dim3(XT_OCT_BLOCK), 0, st, ds.view, B.q, B.hit, B.retest, B.cnt, ray64, flags);
    } else {
        hipLaunchKernelGGL(k_retest<XT_ACCEL_KDTREE>, dim3(grid_for(B.cap / 64 + 1, XT_BLOCK, 1024)), dim3(XT_BLOCK),
                           0, st, ds.view, B.q, B.hit, B.retest, B.cnt, ray64, flags);
    }
    XT_LAUNCHED();
    return XT_OK;
}

// Launches trace + FP64 re-test over queue `qi` of a lane (count in
// lane.cnt->n[qi]).
int xt_launch_trace(xt_dev *d, const DevScene &ds, int accel, int qi, const WaveLane &L, const double *ray64,
                    uint32_t *flags, bool time_it, hipStream_t st) {
    WaveBuffers &w = d->wave;
    TraceBufs B;
    B.q = Queue{w.q[qi].r0 + L.off, w.q[qi].r1 + L.off, w.q[qi].path + L.off};
    B.count = &L.cnt->n[qi];
    B.hit = w.hit + L.off;
    B.retest = w.retest + L.off;
    B.recheck = w.recheck + L.off;
    B.cnt = L.cnt;
    B.cap = L.cap;
    return xt_launch_trace_bufs(d, ds, accel, B, ray64, flags, time_it, st);
}

extern "C" int xt_trace_device(xt_dev *d, uint32_t scene_id, int accel, const void *ray0, const void *ray1,
                               void *hit, uint32_t *flags, uint64_t n, float *kernel_ms) {
    if (!d || !ray0 || !ray1 || !hit) return XT_E_INVALID;
    if (n >= (1ull << 31)) return XT_E_INVALID;
    auto it = d->scenes.find(scene_id);
    if (it == d->scenes.end()) return XT_E_NOTFOUND;
    if (accel != XT_ACCEL_KDTREE && accel != XT_ACCEL_OCTREE) return XT_E_INVALID;
    XT_TRY(hipSetDevice(d->device));
    int rc = xt_dev_ensure_wave(d, std::max<size_t>(n, 1), false);
    if (rc) return rc;
    WaveBuffers &w = d->wave;
    // the caller's buffers are used in place as queue 0 / the hit array
    float4 *save0 = w.q[0].r0, *save1 = w.q[0].r1, *saveh = w.hit;
    w.q[0].r0 = (float4 *)ray0;
    w.q[0].r1 = (float4 *)ray1;
    w.hit = (float4 *)hit;
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, d->stream, w.cnt, 0x800000ffu);
    // paths are only read by the FP64 re-test when FP64 rays are given; set
    // the count with a one-thread kernel instead of a memset (graph-safe)
    hipLaunchKernelGGL(k_set_count, dim3(1), dim3(1), 0, d->stream, w.cnt, 0u, (uint32_t)n);
    if (flags) hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n, 256, 4096)), dim3(256), 0, d->stream, flags, 0u, (size_t)n);
    d->trace_launches = 0;
    d->timed_pairs.clear();
    rc = xt_launch_trace(d, it->second, accel, 0, WaveLane{0, (size_t)n, w.cnt}, nullptr, flags, kernel_ms != nullptr, d->stream);
    w.q[0].r0 = save0;
    w.q[0].r1 = save1;
    w.hit = saveh;
    if (rc) return rc;
    XT_TRY(hipStreamSynchronize(d->stream));
    if (kernel_ms && !d->timed_pairs.empty()) {
        float ms = 0.f;
        hipEventElapsedTime(&ms, d->tev[d->timed_pairs[0]], d->tev[d->timed_pairs[0] + 1]);
        *kernel_ms = ms;
    }
    return XT_OK;
}

// xt_trace_device with the scratch in the per-device static array: the caller's
// rays and hits in place, counters and re-test lists in `scratch`. No
// allocation: the scene is resident, the counters are cleared by a kernel.
extern "C" int xt_trace_device_static(xt_dev *d, uint32_t scene_id, int accel, const void *ray0, const void *ray1,
                                      void *hit, uint32_t *flags, uint64_t n) {
    if (!d || !ray0 || !ray1 || !hit) return XT_E_INVALID;
    if (n > XT_TRACE_STATIC_MAX_RAYS) {
        d->err = "more rays than the static trace scratch holds (XT_TRACE_STATIC_MAX_RAYS)";
        return XT_E_INVALID;
    }
    auto it = d->scenes.find(scene_id);
    if (it == d->scenes.end()) return XT_E_NOTFOUND;
    if (accel != XT_ACCEL_KDTREE && accel != XT_ACCEL_OCTREE) return XT_E_INVALID;
    if (!d->scratch) {
        d->err = "the static trace scratch has no device address";
        return XT_E_DEVICE;
    }
    XT_TRY(hipSetDevice(d->device));
    StaticTraceScratch *sc = (StaticTraceScratch *)d->scratch;
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, d->stream, &sc->cnt, 0x800000ffu);
    hipLaunchKernelGGL(k_set_count, dim3(1), dim3(1), 0, d->stream, &sc->cnt, 0u, (uint32_t)n);
    if (flags) hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n, 256, 4096)), dim3(256), 0, d->stream, flags, 0u, (size_t)n);
    XT_LAUNCHED();
    TraceBufs B;
    B.q = Queue{(float4 *)ray0, (float4 *)ray1, nullptr};
    B.count = &sc->cnt.n[0];
    B.hit = (float4 *)hit;
    B.retest = sc->retest;
    B.recheck = sc->recheck;
    B.cnt = &sc->cnt;
    B.cap = (size_t)n;
    int rc = xt_launch_trace_bufs(d, it->second, accel, B, nullptr, flags, false, d->stream);
    if (rc) return rc;
    XT_TRY(hipStreamSynchronize(d->stream));
    return XT_OK;
}

extern "C" int xt_shade(xt_dev *d, uint32_t scene_id, const xt_shade_query *q, uint64_t n, float *intensity,
                        int32_t *result) {
    if (!d || !q || !intensity) return XT_E_INVALID;
    if (n == 0) return XT_OK;
    if (n >= (1ull << 31)) return XT_E_INVALID;
    auto it = d->scenes.find(scene_id);
    if (it == d->scenes.end()) return XT_E_NOTFOUND;
    XT_TRY(hipSetDevice(d->device));
    struct Temp {
        xt_dev *d;
        std::vector<void *> p;
        uint64_t bytes = 0;
        ~Temp() {
            (void)hipStreamSynchronize(d->stream);
            for (void *x : p) (void)hipFree(x);
            d->used -= std::min<uint64_t>(d->used, bytes);
        }
    } tmp{d, {}};
    xt_shade_query *dq = nullptr;
    float *di = nullptr;
    int32_t *dr = nullptr;
    int rc = dev_alloc(d, &dq, n, &tmp.p, &tmp.bytes);
    if (!rc) rc = dev_alloc(d, &di, 3 * n, &tmp.p, &tmp.bytes);
    if (!rc) rc = dev_alloc(d, &dr, n, &tmp.p, &tmp.bytes);
    if (rc) return rc;
    XT_TRY(hipMemcpyAsync(dq, q, n * sizeof(xt_shade_query), hipMemcpyHostToDevice, d->stream));
    XT_TRY(hipMemcpyAsync(di, intensity, 3 * n * sizeof(float), hipMemcpyHostToDevice, d->stream));
    hipLaunchKernelGGL(k_shade_direct, dim3(grid_for(n, XT_BLOCK, 4096)), dim3(XT_BLOCK), 0, d->stream,
                       it->second.view, (const xt_shade_query *)dq, (uint32_t)n, di, dr);
    XT_LAUNCHED();
    XT_TRY(hipMemcpyAsync(intensity, di, 3 * n * sizeof(float), hipMemcpyDeviceToHost, d->stream));
    if (result) XT_TRY(hipMemcpyAsync(result, dr, n * sizeof(int32_t), hipMemcpyDeviceToHost, d->stream));
    XT_TRY(hipStreamSynchronize(d->stream));
    return XT_OK;
}

extern "C" int xt_trace(xt_dev *d, uint32_t scene_id, int accel, const xt_rays *rays, xt_hits *hits, uint64_t n) {
    if (!d || !rays || !hits) return XT_E_INVALID;
    if (n == 0) return XT_OK;
    if (n >= (1ull << 31)) return XT_E_INVALID;
    auto it = d->scenes.find(scene_id);
    if (it == d->scenes.end()) return XT_E_NOTFOUND;
    XT_TRY(hipSetDevice(d->device));
    int rc = xt_dev_ensure_wave(d, n, false);
    if (rc) return rc;
    const float *src[6] = {rays->ox, rays->oy, rays->oz, rays->dx, rays->dy, rays->dz};
    for (int k = 0; k < 6; ++k)
        if (!src[k]) return XT_E_INVALID;
    // temporary SoA buffers, given back on every return below
    struct Temp {
        xt_dev *d;
        uint64_t bytes;
        std::vector<void *> p;
        ~Temp() {
            (void)hipStreamSynchronize(d->stream);
            for (void *q : p) (void)hipFree(q);
            d->used -= std::min<uint64_t>(d->used, (uint64_t)p.size() * bytes);
        }
    } tmp{d, n * sizeof(float), {}};
    float *soa[6];
    for (int k = 0; k < 6; ++k) {
        rc = dev_alloc(d, &soa[k], n, &tmp.p, nullptr);
        if (rc) return rc;
        XT_TRY(hipMemcpyAsync(soa[k], src[k], n * sizeof(float), hipMemcpyHostToDevice, d->stream));
    }
    WaveBuffers &w = d->wave;
    Queue q{w.q[0].r0, w.q[0].r1, w.q[0].path};
    hipLaunchKernelGGL(k_pack_rays, dim3(grid_for(n, 256, 4096)), dim3(256), 0, d->stream, soa[0], soa[1], soa[2],
                       soa[3], soa[4], soa[5], q, (uint32_t)n);
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, d->stream, w.cnt, 0x800000ffu);
    hipLaunchKernelGGL(k_set_count, dim3(1), dim3(1), 0, d->stream, w.cnt, 0u, (uint32_t)n);
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n, 256, 4096)), dim3(256), 0, d->stream, w.flags, 0u, (size_t)n);
    d->trace_launches = 0;
    d->timed_pairs.clear();
    rc = xt_launch_trace(d, it->second, accel, 0, WaveLane{0, (size_t)n, w.cnt}, nullptr, w.flags, false, d->stream);
    if (rc) return rc;
    hipLaunchKernelGGL(k_unpack_hits, dim3(grid_for(n, 256, 4096)), dim3(256), 0, d->stream, w.hit, soa[0],
                       (uint32_t *)soa[1], soa[2], soa[3], (uint32_t)n);
    XT_LAUNCHED();
    float *dst[4] = {hits->t, (float *)hits->prim, hits->u, hits->v};
    for (int k = 0; k < 4; ++k)
        if (dst[k]) XT_TRY(hipMemcpyAsync(dst[k], soa[k], n * sizeof(float), hipMemcpyDeviceToHost, d->stream));
    if (hits->flags) XT_TRY(hipMemcpyAsync(hits->flags, w.flags, n * sizeof(uint32_t), hipMemcpyDeviceToHost, d->stream));
    XT_TRY(hipStreamSynchronize(d->stream));
    return XT_OK;
}
