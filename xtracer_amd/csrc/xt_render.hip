// One-job-per-launch rendering (xt_render / xt_render_device): the wavefront
// loop that replaces IIntegrator::render()'s OpenMP tile loop
// (src/xtcore/integrator.cc:29-53) for the stencil, depth, ao and
// pathtracer integrators.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <functional>
#include <new>
#include <vector>

#include "xt_device.h"
#include "xt_render.h"

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
}  // namespace

int xt_job_check(xt_dev *d, const xt_job *job, JobDev &j, const DevScene **ds_out) {
    if (!job) return XT_E_INVALID;
    auto it = d->scenes.find(job->scene_id);
    if (it == d->scenes.end()) {
        d->err = "unknown scene id";
        return XT_E_NOTFOUND;
    }
    if (job->width == 0 || job->height == 0 || job->width > 16384 || job->height > 16384) {
        d->err = "bad resolution";
        return XT_E_INVALID;
    }
    if (job->integrator > XT_INTEGRATOR_WIREFRAME || job->accel > XT_ACCEL_OCTREE || job->aa_mode > XT_AA_RANDOM ||
        job->accum > XT_ACCUM_FP64) {
        d->err = "bad job parameter";
        return XT_E_INVALID;
    }
    if (job->rdepth > 0xffffu) {
        d->err = "rdepth above 65535";
        return XT_E_INVALID;
    }
    if (job->camera.kind > XT_CAMERA_CUBEMAP) {
        d->err = "unknown camera kind";
        return XT_E_INVALID;
    }
    if (job->camera.kind == XT_CAMERA_CUBEMAP && job->height % 6u != 0) {
        // rows past the sixth face have no face in the reference (cubemap.cc:26-36)
        d->err = "cubemap camera: the image height must be a multiple of 6";
        return XT_E_INVALID;
    }
    if (job->integrator == XT_INTEGRATOR_EMISSION && !it->second.materials_supported) {
        d->err = "scene has material samplers the GPU path does not evaluate";
        return XT_E_SCENE;
    }
    if (job->integrator == XT_INTEGRATOR_PATHTRACER && !it->second.pt_supported) {
        d->err = "scene has materials or an environment the GPU pathtracer does not evaluate";
        return XT_E_SCENE;
    }
    std::memset(&j, 0, sizeof j);
    j.W = job->width;
    j.H = job->height;
    j.tiles_x = (j.W + 7) / 8;
    const uint32_t tiles_y = (j.H + 7) / 8;
    j.NT = j.tiles_x * tiles_y * 64;
    j.spp = job->samples;
    j.aa = job->aa;
    j.aa_mode = job->aa_mode;
    // antialiasing::produce (aa.cc:51-77): GRID level L gives L*L positions
    // of weight 1/L^2 (level 0: none), RANDOM gives L of weight 1/L; each is
    // repeated `samples` times with its weight divided by `samples`.
    float aw;
    if (j.aa_mode == XT_AA_GRID) {
        j.n_aa = j.aa == 0 ? 0 : j.aa * j.aa;
        aw = j.aa <= 1 ? 1.f : 1.f / (float)(j.aa * j.aa);
        j.aa_sub = j.aa ? 1.f / (float)j.aa : 1.f;
        j.aa_off = j.aa_sub * 0.5f;
    } else {
        j.n_aa = j.aa;
        aw = j.aa ? 1.f / (float)j.aa : 1.f;
    }
    j.total = j.n_aa * j.spp;
    j.weight = (double)aw;
    if (j.spp) j.weight *= 1. / (double)j.spp;
    j.weight_f = (float)j.weight;
    j.rdepth = job->rdepth;
    j.integrator = job->integrator;
    j.seed = job->seed;
    j.job_key = job->job_key;
    j.accum = job->accum;
    if (j.accum == XT_ACCUM_DEFAULT)
        j.accum = (j.integrator == XT_INTEGRATOR_AO || j.integrator == XT_INTEGRATOR_PATHTRACER) ? XT_ACCUM_FP64
                                                                                              : XT_ACCUM_REFERENCE;
    // the normal integrator blends into the stored pixel, and the uv
    // integrator's alpha adds weight squared: neither is a plain sum
    if (j.integrator == XT_INTEGRATOR_NORMAL || j.integrator == XT_INTEGRATOR_UV) j.accum = XT_ACCUM_REFERENCE;
    j.wf = (float)j.W;
    j.hf = (float)j.H;
    j.cam = job->camera;
    *ds_out = &it->second;
    return XT_OK;
}

uint32_t xt_rounds(const JobDev &j) {
    switch (j.integrator) {
        case XT_INTEGRATOR_AO: return 2;
        case XT_INTEGRATOR_PATHTRACER: return j.rdepth;
        default: return 1;
    }
}

uint32_t xt_auto_spp_pass(const JobDev &j, uint32_t requested, uint32_t target) {
    if (j.total == 0) return 1;
    if (requested) return std::min(requested, j.total);
    uint64_t s = std::max<uint64_t>(1, target / j.NT);
    return (uint32_t)std::min<uint64_t>(s, j.total);
}

// Launches one shading step of a lane for query round r over queue qi into
// queue qi^1.
int xt_launch_shade(xt_dev *d, const DevScene &ds, uint32_t integrator, uint32_t r, int qi, const WaveLane &L,
                    hipStream_t st) {
    WaveBuffers &w = d->wave;
    Queue in{w.q[qi].r0 + L.off, w.q[qi].r1 + L.off, w.q[qi].path + L.off};
    Queue out{w.q[qi ^ 1].r0 + L.off, w.q[qi ^ 1].r1 + L.off, w.q[qi ^ 1].path + L.off};
    const float4 *hit = w.hit + L.off;
    const uint32_t *n_in = &L.cnt->n[qi];
    uint32_t *n_out = &L.cnt->n[qi ^ 1];
    const unsigned g = grid_for(L.cap, XT_BLOCK, 8192);
    switch (integrator) {
        case XT_INTEGRATOR_STENCIL:
        case XT_INTEGRATOR_WIREFRAME:  // wireframe/integrator.cc is the stencil rule
            hipLaunchKernelGGL(k_shade_stencil, dim3(g), dim3(XT_BLOCK), 0, st, in, n_in, hit, w.rad);
            break;
        case XT_INTEGRATOR_NORMAL:
            hipLaunchKernelGGL(k_shade_normal, dim3(g), dim3(XT_BLOCK), 0, st, ds.view, in, n_in, hit, w.ray64, w.rad);
            break;
        case XT_INTEGRATOR_UV:
            hipLaunchKernelGGL(k_shade_uv, dim3(g), dim3(XT_BLOCK), 0, st, ds.view, in, n_in, hit, w.ray64, w.rad);
            break;
        case XT_INTEGRATOR_EMISSION:
            hipLaunchKernelGGL(k_shade_emission, dim3(g), dim3(XT_BLOCK), 0, st, ds.view, in, n_in, hit, w.ray64,
                               w.rad);
            break;
        case XT_INTEGRATOR_DEPTH:
            hipLaunchKernelGGL(k_shade_depth, dim3(g), dim3(XT_BLOCK), 0, st, ds.view, in, n_in, hit, w.ray64,
                               w.rad);
            break;
        case XT_INTEGRATOR_AO:
            if (r == 0)
                hipLaunchKernelGGL(k_shade_ao1, dim3(g), dim3(XT_BLOCK), 0, st, ds.view, in, n_in, hit, out,
                                   n_out, w.ps0, w.rad);
            else
                hipLaunchKernelGGL(k_shade_ao2, dim3(g), dim3(XT_BLOCK), 0, st, in, n_in, hit, w.rad);
            break;
        default:
            // the sampler code more than doubles the kernel's registers:
            // scenes without samplers take the variant that leaves it out
            if (ds.view.sampled) {
                if (!w.chan) {
                    d->err = "sampled scene without channel buffers";
                    return XT_E_INVALID;
                }
                float4 *chan = w.chan + 3 * L.off;
                hipLaunchKernelGGL(k_channels, dim3(g), dim3(XT_BLOCK), 0, st, ds.view, in, n_in, hit, chan);
                XT_LAUNCHED();
                hipLaunchKernelGGL(k_shade_pt<true>, dim3(g), dim3(XT_BLOCK), 0, st, ds.view, in, n_in, hit, out,
                                   n_out, w.ps0, w.ps1, w.rad, chan);
            } else {
                hipLaunchKernelGGL(k_shade_pt<false>, dim3(g), dim3(XT_BLOCK), 0, st, ds.view, in, n_in, hit, out,
                                   n_out, w.ps0, w.ps1, w.rad, (const float4 *)nullptr);
            }
    }
    XT_LAUNCHED();
    return XT_OK;
}

int xt_ensure_host_counters(xt_dev *d) {
    if (!d->host_counters)
        XT_TRY(hipHostMalloc((void **)&d->host_counters, XT_MAX_LANES * sizeof(Counters), hipHostMallocDefault));
    return XT_OK;
}

// Query rounds of one pass over every group: trace + shade per round, queue
// 0 holding the camera rays. With several groups, each group's kernels run
// on a stream of its own so that one group's launch tails overlap another's
// work; the device stream forks to them and joins them again at the end.
// Every second round past the second, each live group's next queue count is
// copied to pinned memory on the group's stream. The host then reads the
// copies of the check before, two rounds back, so that the device still has
// two rounds queued while the host waits and the groups never wait for each
// other, and issues no more rounds to groups whose paths have all ended.
int xt_run_rounds(xt_dev *d, PassGroup *groups, int n_groups, bool time_trace, const GroupHook *before,
                  const GroupHook *after) {
    WaveBuffers &w = d->wave;
    if (n_groups < 1 || n_groups > XT_MAX_LANES) {
        d->err = "pass groups out of range";
        return XT_E_INVALID;
    }
    uint32_t max_rounds = 0;
    for (int k = 0; k < n_groups; ++k) {
        groups[k].live = true;
        max_rounds = std::max(max_rounds, groups[k].rounds);
    }
    const int n_st = n_groups > 1 ? std::min(n_groups, XT_LANE_STREAMS) : 0;
    for (int k = 0; k < n_st; ++k) {
        if (!d->lane_stream[k]) {
            XT_TRY(hipStreamCreateWithFlags(&d->lane_stream[k], hipStreamNonBlocking));
            XT_TRY(hipEventCreateWithFlags(&d->lane_ev[k], hipEventDisableTiming));
        }
    }
    if (n_st && !d->fork_ev) XT_TRY(hipEventCreateWithFlags(&d->fork_ev, hipEventDisableTiming));
    if (!d->host_snap)
        XT_TRY(hipHostMalloc((void **)&d->host_snap, 2 * XT_MAX_LANES * sizeof(uint32_t), hipHostMallocDefault));
    for (int s = 0; s < 2; ++s)
        for (int k = 0; k < n_groups; ++k)
            if (!d->snap_ev[s][k]) XT_TRY(hipEventCreateWithFlags(&d->snap_ev[s][k], hipEventDisableTiming));
    auto stream_of = [&](int k) { return n_st ? d->lane_stream[k % n_st] : d->stream; };
    if (n_st) {
        XT_TRY(hipEventRecord(d->fork_ev, d->stream));
        for (int k = 0; k < n_st; ++k) XT_TRY(hipStreamWaitEvent(d->lane_stream[k], d->fork_ev, 0));
    }
    // the device stream waits for the group streams on every return from
    // here, errors included, so that later work on it cannot overlap this
    // call's kernels
    struct Join {
        xt_dev *d;
        int n_st;
        ~Join() {
            for (int k = 0; k < n_st; ++k) {
                if (hipEventRecord(d->lane_ev[k], d->lane_stream[k]) != hipSuccess ||
                    hipStreamWaitEvent(d->stream, d->lane_ev[k], 0) != hipSuccess)
                    (void)hipStreamSynchronize(d->lane_stream[k]);
            }
        }
    } join{d, n_st};
    for (int k = 0; before && k < n_groups; ++k) {
        const int rc = (*before)(k, stream_of(k));
        if (rc) return rc;
    }
    int qi = 0, slot = 0;
    bool pending = false;
    for (uint32_t r = 0; r < max_rounds; ++r) {
        for (int k = 0; k < n_groups; ++k) {
            PassGroup &G = groups[k];
            if (!G.live || r >= G.rounds) continue;
            const hipStream_t st = stream_of(k);
            hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, st, G.lane.cnt,
                               (1u << (qi ^ 1)) | (1u << CNT_RETEST) | (1u << CNT_CURSOR) | (1u << CNT_RECHECK));
            const bool fp64_rays = xt_needs_ray64(G.integrator) && r == 0;
            int rc = xt_launch_trace(d, *G.ds, (int)G.accel, qi, G.lane, fp64_rays ? w.ray64 : nullptr, nullptr,
                                     time_trace, st);
            if (rc) return rc;
            rc = xt_launch_shade(d, *G.ds, G.integrator, r, qi, G.lane, st);
            if (rc) return rc;
        }
        qi ^= 1;
        if (max_rounds > 2 && r + 1 < max_rounds && (r % 2 == 1)) {
            if (pending) {  // the previous check: paths left after round r - 2
                bool any = false;
                for (int k = 0; k < n_groups; ++k) {
                    PassGroup &G = groups[k];
                    if (!G.live) continue;
                    XT_TRY(hipEventSynchronize(d->snap_ev[slot ^ 1][k]));
                    if (r + 1 >= G.rounds || d->host_snap[(slot ^ 1) * XT_MAX_LANES + k] == 0) G.live = false;
                    any = any || G.live;
                }
                if (!any) break;
            }
            for (int k = 0; k < n_groups; ++k) {
                PassGroup &G = groups[k];
                if (!G.live) continue;
                XT_TRY(hipMemcpyAsync(d->host_snap + slot * XT_MAX_LANES + k, &G.lane.cnt->n[qi], sizeof(uint32_t),
                                      hipMemcpyDeviceToHost, stream_of(k)));
                XT_TRY(hipEventRecord(d->snap_ev[slot][k], stream_of(k)));
            }
            slot ^= 1;
            pending = true;
        }
    }
    for (int k = 0; after && k < n_groups; ++k) {
        const int rc = (*after)(k, stream_of(k));
        if (rc) return rc;
    }
    return XT_OK;
}

// `after_pass(j, base, spp_pass)`, when given, runs on the host after each
// pass's accumulation has been queued.
static int render_core(xt_dev *d, const xt_job *job, float *d_out, xt_job_stats *stats,
                       const std::function<int(const JobDev &, uint32_t, uint32_t)> *after_pass = nullptr) {
    JobDev j;
    const DevScene *ds = nullptr;
    int rc = xt_job_check(d, job, j, &ds);
    if (rc) return rc;
    XT_TRY(hipSetDevice(d->device));
    const uint32_t spp_pass = xt_auto_spp_pass(j, job->samples_per_pass, d->tuning.wave_paths);
    const size_t cap = (size_t)spp_pass * j.NT;
    if (cap >= (1ull << 31)) {
        d->err = "pass too large";
        return XT_E_INVALID;
    }
    const bool need64 = xt_needs_ray64(j.integrator);
    rc = xt_dev_ensure_wave(d, cap, need64, xt_needs_chan(j.integrator, ds->view));
    if (rc) return rc;
    rc = xt_dev_ensure_accum(d, (size_t)j.W * j.H);
    if (rc) return rc;
    rc = xt_ensure_host_counters(d);
    if (rc) return rc;
    WaveBuffers &w = d->wave;
    AccumBuffers &a = d->accum;
    const size_t npix = (size_t)j.W * j.H;
    const unsigned gfill = grid_for(npix * 4, 256, 8192);
    hipLaunchKernelGGL(k_fill_u32, dim3(gfill), dim3(256), 0, d->stream, (uint32_t *)a.accf, 0u, npix * 4);
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(npix * 12, 256, 8192)), dim3(256), 0, d->stream,
                       (uint32_t *)a.acc64, 0u, npix * 12);
    // the pixel slots of each group: the slots cut at 8x8 tiles into
    // frame_lanes nearly equal parts (every sample of a pixel stays in its
    // group); one group when a pass hook reads the paths, which it does in
    // the one-group order (path = sample * NT + slot)
    const uint32_t tiles = j.NT / 64;
    const int n_groups = after_pass ? 1 : (int)std::max(1u, std::min(d->tuning.frame_lanes, tiles));
    uint32_t q_at[XT_LANE_STREAMS + 1];
    for (int g = 0; g <= n_groups; ++g) q_at[g] = (uint32_t)((uint64_t)tiles * g / n_groups) * 64u;
    hipLaunchKernelGGL(k_reset_lanes, dim3(n_groups), dim3(64), 0, d->stream, w.cnt, 0x800000ffu);
    XT_LAUNCHED();
    d->trace_launches = 0;
    d->timed_pairs.clear();
    const bool time_trace = stats != nullptr;
    XT_TRY(hipEventRecord(d->ev0, d->stream));
    PassGroup G[XT_LANE_STREAMS];
    for (int g = 0; g < n_groups; ++g) {
        G[g].ds = ds;
        G[g].integrator = j.integrator;
        G[g].accel = job->accel;
        G[g].rounds = xt_rounds(j);
        G[g].lane = WaveLane{(size_t)spp_pass * q_at[g], (size_t)spp_pass * (q_at[g + 1] - q_at[g]), w.cnt + g};
    }
    uint32_t passes = 0;
    for (uint32_t base = 0; base < j.total; base += spp_pass) {
        ++passes;
        hipLaunchKernelGGL(k_reset_lanes, dim3(n_groups), dim3(64), 0, d->stream, w.cnt, 0x7u);
        // each group's camera rays and accumulation run on the group's own
        // stream, so that they overlap the other groups' rounds (the groups
        // write disjoint pixels)
        const GroupHook camera = [&](int g, hipStream_t st) -> int {
            const WaveLane &L = G[g].lane;
            hipLaunchKernelGGL(k_set_count, dim3(1), dim3(1), 0, st, L.cnt, 0u, (uint32_t)L.cap);
            Queue q0{w.q[0].r0 + L.off, w.q[0].r1 + L.off, w.q[0].path + L.off};
            hipLaunchKernelGGL(k_camera, dim3(grid_for(L.cap, XT_BLOCK, 8192)), dim3(XT_BLOCK), 0, st, j, base,
                               spp_pass, q_at[g], q_at[g + 1] - q_at[g], (uint32_t)L.off, q0, w.ps0, w.ps1, w.rad,
                               need64 ? w.ray64 : nullptr, d->tuning.ray_order);
            XT_LAUNCHED();
            return XT_OK;
        };
        const GroupHook accumulate = [&](int g, hipStream_t st) -> int {
            const WaveLane &L = G[g].lane;
            hipLaunchKernelGGL(k_accumulate, dim3(grid_for(q_at[g + 1] - q_at[g], XT_BLOCK, 8192)), dim3(XT_BLOCK), 0,
                               st, j, base, spp_pass, q_at[g], q_at[g + 1] - q_at[g], w.rad + L.off,
                               need64 ? w.ray64 + 6 * L.off : nullptr, a.accf, a.acc64);
            XT_LAUNCHED();
            return XT_OK;
        };
        rc = xt_run_rounds(d, G, n_groups, time_trace, &camera, &accumulate);
        if (rc) return rc;
        if (after_pass) {
            rc = (*after_pass)(j, base, spp_pass);
            if (rc) return rc;
        }
    }
    hipLaunchKernelGGL(k_resolve, dim3(grid_for(npix, XT_BLOCK, 8192)), dim3(XT_BLOCK), 0, d->stream, j, a.accf,
                       a.acc64, (float4 *)d_out);
    XT_LAUNCHED();
    XT_TRY(hipEventRecord(d->ev1, d->stream));
    XT_TRY(hipMemcpyAsync(d->host_counters, w.cnt, n_groups * sizeof(Counters), hipMemcpyDeviceToHost, d->stream));
    XT_TRY(hipStreamSynchronize(d->stream));
    for (int g = 1; g < n_groups; ++g) {
        d->host_counters->queries += d->host_counters[g].queries;
        d->host_counters->retests += d->host_counters[g].retests;
        d->host_counters->trace_fail += d->host_counters[g].trace_fail;
        d->host_counters->steps_exceeded += d->host_counters[g].steps_exceeded;
    }
    if (stats) {
        std::memset(stats, 0, sizeof *stats);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, d->ev0, d->ev1);
        stats->ms_total = ms;
        stats->queries = d->host_counters->queries;
        stats->retests = d->host_counters->retests;
        stats->trace_fail = d->host_counters->trace_fail;
        stats->retest_fail = d->host_counters->steps_exceeded;
        stats->samples = (uint64_t)j.total * j.W * j.H;
        stats->passes = passes;
        // each launch's own time, and the union of the launches' intervals
        // (launches of different groups run at the same time)
        double tsum = 0;
        std::vector<std::pair<float, float>> spans;
        for (int k : d->timed_pairs) {
            float t = 0.f, t0 = 0.f, t1 = 0.f;
            (void)hipEventElapsedTime(&t, d->tev[k], d->tev[k + 1]);
            (void)hipEventElapsedTime(&t0, d->ev0, d->tev[k]);
            (void)hipEventElapsedTime(&t1, d->ev0, d->tev[k + 1]);
            tsum += t;
            spans.emplace_back(t0, t1);
        }
        std::sort(spans.begin(), spans.end());
        double busy = 0, at = -1e30;
        for (const auto &sp : spans) {
            const double a = std::max<double>(sp.first, at), b = sp.second;
            if (b > a) busy += b - a;
            at = std::max<double>(at, b);
        }
        stats->ms_trace = tsum;
        stats->ms_trace_busy = busy;
        stats->trace_launches = (uint32_t)d->timed_pairs.size();
    }
    return XT_OK;
}

extern "C" int xt_render_device(xt_dev *d, const xt_job *job, float *d_rgba_out, xt_job_stats *stats) {
    if (!d || !job || !d_rgba_out) return XT_E_INVALID;
    return render_core(d, job, d_rgba_out, stats);
}

extern "C" int xt_render(xt_dev *d, const xt_job *job, float *rgba_out, xt_job_stats *stats) {
    if (!d || !job || !rgba_out) return XT_E_INVALID;
    const size_t npix = (size_t)job->width * job->height;
    int rc = xt_dev_ensure_accum(d, npix);
    if (rc) return rc;
    rc = render_core(d, job, (float *)d->accum.out, stats);
    if (rc) return rc;
    XT_TRY(hipMemcpy(rgba_out, d->accum.out, npix * sizeof(float4), hipMemcpyDeviceToHost));
    return XT_OK;
}

// Ray graph of a depth render (integrator/depth/integrator.cc:43-51): per
// sample a two-vertex path from the camera ray's origin (colour 0) to its hit
// point ((0, 0, 0) on a miss) coloured with the pixel's value after that
// sample. Samples are gathered per pass from the FP64 rays (origin, and the
// hit point k_shade_depth leaves in the direction slots) and the per-sample
// depth, then ordered as the reference's tiles hold them: 32-pixel tiles in
// row-major order (segment_framebuffer, tile.cc:67-93, the context's default
// tile size), pixels x-major within a tile (aa.cc:55-56), samples in FIFO
// order; raygraph::write (util/raygraph.cc:8-78) writes them as ASCII PLY,
// numbers as an ostream prints them (%g).
extern "C" int xt_render_raygraph(xt_dev *d, const xt_job *job, float *rgba_out, const char *ply_path) {
    if (!d || !job || !rgba_out || !ply_path) return XT_E_INVALID;
    const bool depth = job->integrator == XT_INTEGRATOR_DEPTH;
    const size_t npix = (size_t)job->width * job->height;
    std::vector<double> org, pnt;
    std::vector<float> val;
    std::vector<uint8_t> hits;
    uint32_t total = 0, W = job->width, H = job->height;
    double weight = 0;
    std::function<int(const JobDev &, uint32_t, uint32_t)> grab = [&](const JobDev &j, uint32_t base,
                                                                         uint32_t spp_pass) -> int {
        if (org.empty()) {
            total = j.total;
            weight = j.weight;
            const size_t n = npix * total;
            org.resize(3 * n);
            pnt.resize(3 * n);
            val.resize(n);
            hits.resize(n);
        }
        const size_t cap = (size_t)spp_pass * j.NT;
        std::vector<double> r64(6 * cap);
        std::vector<float4> rad(cap);
        XT_TRY(hipStreamSynchronize(d->stream));
        XT_TRY(hipMemcpy(r64.data(), d->wave.ray64, r64.size() * sizeof(double), hipMemcpyDeviceToHost));
        XT_TRY(hipMemcpy(rad.data(), d->wave.rad, cap * sizeof(float4), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < cap; ++i) {
            const uint32_t s = base + (uint32_t)(i / j.NT), q = (uint32_t)(i % j.NT);
            const uint32_t tile = q >> 6, within = q & 63u;
            const uint32_t x = (tile % j.tiles_x) * 8u + (within & 7u), y = (tile / j.tiles_x) * 8u + (within >> 3);
            if (x >= j.W || y >= j.H || s >= j.total) continue;
            const size_t k = ((size_t)y * j.W + x) * total + s;
            const bool hit = rad[i].y != 0.f;
            for (int c = 0; c < 3; ++c) {
                org[3 * k + c] = r64[6 * i + c];
                pnt[3 * k + c] = hit ? r64[6 * i + 3 + c] : 0.0;
            }
            val[k] = rad[i].x;
            hits[k] = hit;
        }
        return XT_OK;
    };
    int rc = xt_dev_ensure_accum(d, npix);
    if (rc) return rc;
    try {
        rc = render_core(d, job, (float *)d->accum.out, nullptr, depth ? &grab : nullptr);
    } catch (const std::bad_alloc &) {
        d->err = "ray graph too large for host memory";
        return XT_E_INVALID;
    }
    if (rc) return rc;
    XT_TRY(hipMemcpy(rgba_out, d->accum.out, npix * sizeof(float4), hipMemcpyDeviceToHost));
    FILE *f = std::fopen(ply_path, "w");
    if (!f) {
        d->err = std::string("cannot write ") + ply_path;
        return XT_E_IO;
    }
    const size_t n = depth ? npix * total : 0;
    std::fprintf(f, "ply\nformat ascii 1.0\ncomment object: A single line\nelement vertex %zu\n", 2 * n);
    std::fprintf(f, "property float x\nproperty float y\nproperty float z\nproperty float red\n"
                    "property float green\nproperty float blue\nelement edge %zu\nproperty int vertex1\n"
                    "property int vertex2\nend_header\n", n);
    const uint32_t TS = 32;
    auto each = [&](const std::function<void(size_t)> &fn) {
        for (uint32_t ty = 0; ty < (H + TS - 1) / TS; ++ty)
            for (uint32_t tx = 0; tx < (W + TS - 1) / TS; ++tx)
                for (uint32_t x = tx * TS; x < std::min(W, tx * TS + TS); ++x)
                    for (uint32_t y = ty * TS; y < std::min(H, ty * TS + TS); ++y) fn((size_t)y * W + x);
    };
    if (depth) {
        each([&](size_t pix) {
            float r = 0.f;  // the tile starts at zero
            for (uint32_t s = 0; s < total; ++s) {
                const size_t k = pix * total + s;
                // depth = color.r() + depth * weight, then ColorRGBAf's clamp at 0
                if (hits[k]) r = std::max((float)((double)r + (double)val[k] * weight), 0.f);
                std::fprintf(f, "%g %g %g 0 0 0\n", org[3 * k], org[3 * k + 1], org[3 * k + 2]);
                std::fprintf(f, "%g %g %g %g %g %g\n", pnt[3 * k], pnt[3 * k + 1], pnt[3 * k + 2], (double)r,
                             (double)r, (double)r);
            }
        });
        size_t at = 0;
        each([&](size_t) {
            for (uint32_t s = 0; s < total; ++s, at += 2) std::fprintf(f, "%zu %zu\n", at, at + 1);
        });
    }
    const bool ok = std::fclose(f) == 0;
    if (!ok) {
        d->err = std::string("cannot write ") + ply_path;
        return XT_E_IO;
    }
    return XT_OK;
}
