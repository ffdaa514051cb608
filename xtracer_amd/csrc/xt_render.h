// Shared pieces of the render loop used by one-job renders and the server.
#pragma once
#include <functional>

#include "xt_device.h"

int xt_job_check(xt_dev *d, const xt_job *job, JobDev &j, const DevScene **ds_out);
uint32_t xt_rounds(const JobDev &j);
uint32_t xt_auto_spp_pass(const JobDev &j, uint32_t requested, uint32_t target_paths);

// One (scene, integrator, accel) group of a pass and the lane its rays use.
struct PassGroup {
    const DevScene *ds = nullptr;
    uint32_t integrator = 0, accel = 0, rounds = 0;
    WaveLane lane;
    bool live = true;
};
int xt_launch_shade(xt_dev *d, const DevScene &ds, uint32_t integrator, uint32_t r, int qi, const WaveLane &lane,
                    hipStream_t st);
// `before(k, stream)` / `after(k, stream)`, when given, queue group k's work
// ahead of its first round / after its last, on the group's stream.
using GroupHook = std::function<int(int, hipStream_t)>;
int xt_run_rounds(xt_dev *d, PassGroup *groups, int n_groups, bool time_trace, const GroupHook *before = nullptr,
                  const GroupHook *after = nullptr);

// Integrators whose shading reads the camera ray in FP64 (the ray64 slot):
// depth's distance, normal, uv and emission's texture coordinates.
// The pathtracer over a scene with samplers shades from k_channels' output.
inline bool xt_needs_chan(uint32_t integrator, const SceneView &v) {
    return integrator == XT_INTEGRATOR_PATHTRACER && v.sampled;
}

inline bool xt_needs_ray64(uint32_t integrator) {
    return integrator == XT_INTEGRATOR_DEPTH || integrator == XT_INTEGRATOR_NORMAL ||
           integrator == XT_INTEGRATOR_UV || integrator == XT_INTEGRATOR_EMISSION;
}
int xt_ensure_host_counters(xt_dev *d);

