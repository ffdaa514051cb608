// Internal state behind an xt_dev handle.
#pragma once
#include <hip/hip_runtime.h>

#include <deque>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "device/xt_kernels.h"

#define XT_TRACE_EVENTS 512

struct DevScene {
    SceneView view;
    std::vector<void *> allocs;
    uint64_t bytes = 0;
    uint64_t kd_bytes = 0;            // of the kd nodes, references and leaf copies (in `allocs`)
    uint32_t n_kd_nodes = 0, n_kd_refs = 0;
    xt_kd_build_stats kd_stats{};     // of the last device build (zero for a host-built tree)
    bool pt_supported = true;
    bool materials_supported = true;
    bool missing_material = false;
};

struct WaveQueue {
    float4 *r0 = nullptr, *r1 = nullptr;
    uint32_t *path = nullptr;
};

#define XT_MAX_LANES 16
#define XT_LANE_STREAMS 4  // streams the groups of a served pass spread over

struct WaveBuffers {
    size_t cap = 0;
    WaveQueue q[2];
    float4 *hit = nullptr;
    uint32_t *retest = nullptr;
    uint4 *recheck = nullptr;
    float4 *ps0 = nullptr, *ps1 = nullptr, *rad = nullptr;
    uint32_t *flags = nullptr;
    Counters *cnt = nullptr;  // XT_MAX_LANES counter blocks
    double *ray64 = nullptr;
    float4 *chan = nullptr;  // 3 per slot: sampled material colours (k_channels)
    std::vector<void *> allocs;
    uint64_t bytes = 0;
};

struct AccumBuffers {
    size_t npix = 0;
    float4 *accf = nullptr;
    double2 *acc64 = nullptr;
    float4 *out = nullptr;
    std::vector<void *> allocs;
    uint64_t bytes = 0;
};

struct ServeState;  // continuous-batching server (xt_serve.hip)

struct xt_dev {
    int device = 0;
    int n_cu = 256;
    hipStream_t stream = nullptr;
    hipStream_t lane_stream[XT_LANE_STREAMS] = {};  // default priority, like `stream`
    hipEvent_t lane_ev[XT_LANE_STREAMS] = {}, fork_ev = nullptr;
    uint64_t budget = 0, used = 0;
    std::map<uint32_t, DevScene> scenes;
    uint32_t next_scene = 1;
    WaveBuffers wave;
    AccumBuffers accum;
    Counters *host_counters = nullptr;  // pinned, XT_MAX_LANES blocks
    uint32_t *host_snap = nullptr;      // pinned, 2 x XT_MAX_LANES queue counts of the early-stop check
    hipEvent_t snap_ev[2][XT_MAX_LANES] = {};
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t tev[XT_TRACE_EVENTS];
    int trace_launches = 0;
    std::vector<int> timed_pairs;
    ServeState *serve = nullptr;
    xt_tuning tuning{12, 256, 64u << 20, 0, 1, 1, 1, 1, 2};
    void *scratch = nullptr;  // this device's address of xt_trace_device_static's scratch array
    std::string err;
};

// Blocks for a grid-stride launch over n items, at least 1, at most cap.
static inline unsigned grid_for(size_t n, unsigned block, unsigned cap) {
    size_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

int xt_dev_ensure_wave(xt_dev *d, size_t cap, bool need64, bool need_chan = false);
int xt_dev_ensure_accum(xt_dev *d, size_t npix);
void xt_dev_free_wave(xt_dev *d);
void xt_dev_free_accum(xt_dev *d);
// A sub-queue of the wavefront buffers: slots [off, off + cap) of the ray
// queues, the hit and the re-test arrays, counted by its own counter block.
// A one-job render uses one lane over the whole buffers; the server gives
// each (scene, integrator, accel) group of a pass a lane of its own.
struct WaveLane {
    size_t off = 0, cap = 0;
    Counters *cnt = nullptr;
};
int xt_launch_trace(xt_dev *d, const DevScene &ds, int accel, int qi, const WaveLane &lane, const double *ray64,
                    uint32_t *flags, bool time_it, hipStream_t st);
// The buffers one trace launch reads and writes.
struct TraceBufs {
    Queue q;
    const uint32_t *count = nullptr;  // rays in q
    float4 *hit = nullptr;
    uint32_t *retest = nullptr;
    uint4 *recheck = nullptr;
    Counters *cnt = nullptr;
    size_t cap = 0;  // queue capacity (sizes the grids)
};
int xt_launch_trace_bufs(xt_dev *d, const DevScene &ds, int accel, const TraceBufs &B, const double *ray64,
                         uint32_t *flags, bool time_it, hipStream_t st);
// Builds the kd-tree of a resident scene on the device (xt_kdbuild.hip) and
// replaces the scene's kd arrays with it.
int xt_kd_build_into(xt_dev *d, DevScene &ds, const xt_build_params *p);
void xt_serve_shutdown(xt_dev *d);
bool xt_serve_uses_scene(const xt_dev *d, uint32_t scene_id);
void xt_serve_forget_scene(xt_dev *d, uint32_t scene_id);
