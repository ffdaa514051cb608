// Continuous-batching render server (xt_submit / xt_poll / xt_serve_step /
// xt_drain). The reference renders one context at a time with a blocking
// IIntegrator::render() (src/xtcore/integrator.cc:29-53); here many jobs
// share every wavefront pass.
//
// Admission: a submitted job waits in FIFO order until its per-job buffers
// (FP32 + FP64 accumulators and the resolved image, 80 B per pixel) fit the
// device's HBM budget; a job that could never fit is refused at submission.
//
// Scheduling: each step runs one pass. XT_SERVE_SRPT (the default) fills it
// with the jobs that have the least work left. XT_SERVE_FAIR lets the oldest unfinished job take
// half of the pass's path slots (the wavefront size), which bounds how much
// sharing slows a large job down; the other jobs with samples left get equal
// shares of the rest, at least one 8x8 tile each, in round-robin order so
// that a job skipped for lack of room goes first in the next pass; leftover
// room goes to the same jobs in the same order. Jobs join at any pass
// boundary (iteration-level batching).
//
// Layout: the pass's slots are grouped by (scene, integrator, accel). Each
// group is a lane of the wavefront buffers with its own counters, so one
// trace launch per group and round covers all of the group's jobs. A job's
// segment covers consecutive samples and is accumulated in sample order into
// the job's own buffers, so a served image is bit-identical to the same job
// rendered alone with xt_render.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <tuple>
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

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

constexpr uint64_t kJobBytesPerPixel = sizeof(float4) + 3 * sizeof(double2) + sizeof(float4);

uint64_t job_bytes(size_t npix) { return ((uint64_t)npix * kJobBytesPerPixel + 255) & ~255ull; }

struct DevBlock {
    void *p = nullptr;
    uint64_t bytes = 0;
};

struct HostBlock {
    float *p = nullptr;
    size_t bytes = 0;
};

enum JobState { WAITING = 0, ACTIVE = 1, FINISHED = 2 };

struct SJob {
    xt_job job;
    JobDev jd;
    int state = WAITING;
    uint32_t issued = 0;  // whole samples handed to passes
    uint32_t qc = 0;      // tile positions of sample `issued` handed out so far
    uint32_t passes = 0;
    size_t npix = 0;
    DevBlock blk;
    HostBlock host;
    float4 *accf = nullptr, *out = nullptr;
    double2 *acc64 = nullptr;
    hipEvent_t done_ev = nullptr;
    double t_submit = 0, t_admit = 0, t_done = 0;
};

// Samples [first, first + samples) of tile positions [q0, q0 + qn) of one
// job: whole samples (q0 = 0, qn = NT) or a slice of one sample.
struct Segment {
    uint64_t ticket;
    uint32_t samples, q0, qn;
    size_t path_base;
};

}  // namespace

struct ServeState {
    std::map<uint64_t, SJob> jobs;
    std::deque<uint64_t> waiting;
    std::vector<uint64_t> active;  // admission order
    uint64_t next_ticket = 1;
    size_t rr = 0;  // round-robin start among active jobs
    std::vector<DevBlock> free_dev;
    std::vector<HostBlock> free_host;
    std::vector<hipEvent_t> free_events;
    uint64_t held = 0;  // bytes of job blocks (in use + pooled)
    hipEvent_t epoch = nullptr, p0 = nullptr, p1 = nullptr;
    double epoch_ms = 0;
    xt_serve_info info{};
    int policy = XT_SERVE_SRPT;
};

namespace {

int serve_init(xt_dev *d) {
    if (d->serve) return XT_OK;
    XT_TRY(hipSetDevice(d->device));
    ServeState *S = new ServeState();
    d->serve = S;
    XT_TRY(hipEventCreate(&S->epoch));
    XT_TRY(hipEventCreate(&S->p0));
    XT_TRY(hipEventCreate(&S->p1));
    XT_TRY(hipEventRecord(S->epoch, d->stream));
    XT_TRY(hipEventSynchronize(S->epoch));
    S->epoch_ms = now_ms();
    int rc = xt_ensure_host_counters(d);
    if (rc) return rc;
    // wavefront buffers at the pass size up front, so that job buffers
    // admitted later cannot crowd them out of the budget
    return xt_dev_ensure_wave(d, d->tuning.wave_paths, true, true);
}

// Takes a pooled block of at least `need` bytes, or allocates one within the
// budget (releasing pooled blocks that are too small when that makes room).
// Returns false when the budget has no room now.
int take_dev_block(xt_dev *d, ServeState &S, uint64_t need, DevBlock &out, bool &ok) {
    ok = false;
    int best = -1;
    for (size_t k = 0; k < S.free_dev.size(); ++k)
        if (S.free_dev[k].bytes >= need && S.free_dev[k].bytes <= 2 * need &&
            (best < 0 || S.free_dev[k].bytes < S.free_dev[best].bytes))
            best = (int)k;
    if (best >= 0) {
        out = S.free_dev[best];
        S.free_dev.erase(S.free_dev.begin() + best);
        ok = true;
        return XT_OK;
    }
    while (d->used + need > d->budget && !S.free_dev.empty()) {
        const DevBlock b = S.free_dev.back();
        S.free_dev.pop_back();
        XT_TRY(hipFree(b.p));
        d->used -= std::min(d->used, b.bytes);
        S.held -= std::min(S.held, b.bytes);
    }
    if (d->used + need > d->budget) return XT_OK;
    void *p = nullptr;
    XT_TRY(hipMalloc(&p, need));
    d->used += need;
    S.held += need;
    out.p = p;
    out.bytes = need;
    ok = true;
    return XT_OK;
}

int take_host_block(xt_dev *d, ServeState &S, size_t need, HostBlock &out) {
    for (size_t k = 0; k < S.free_host.size(); ++k) {
        if (S.free_host[k].bytes >= need && S.free_host[k].bytes <= 2 * need) {
            out = S.free_host[k];
            S.free_host.erase(S.free_host.begin() + k);
            return XT_OK;
        }
    }
    void *p = nullptr;
    XT_TRY(hipHostMalloc(&p, need, hipHostMallocDefault));
    out.p = (float *)p;
    out.bytes = need;
    return XT_OK;
}

// Admits waiting jobs in FIFO order while their buffers fit.
int admit(xt_dev *d, ServeState &S) {
    while (!S.waiting.empty()) {
        SJob &J = S.jobs[S.waiting.front()];
        bool ok = false;
        int rc = take_dev_block(d, S, job_bytes(J.npix), J.blk, ok);
        if (rc) return rc;
        if (!ok) break;
        char *base = (char *)J.blk.p;
        J.accf = (float4 *)base;
        J.acc64 = (double2 *)(base + J.npix * sizeof(float4));
        J.out = (float4 *)(base + J.npix * (sizeof(float4) + 3 * sizeof(double2)));
        hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(J.npix * 16, 256, 8192)), dim3(256), 0, d->stream,
                           (uint32_t *)J.blk.p, 0u, J.npix * 16);
        XT_LAUNCHED();
        if (!S.free_events.empty()) {
            J.done_ev = S.free_events.back();
            S.free_events.pop_back();
        } else {
            XT_TRY(hipEventCreate(&J.done_ev));
        }
        J.state = ACTIVE;
        J.t_admit = now_ms();
        S.active.push_back(S.waiting.front());
        S.waiting.pop_front();
    }
    return XT_OK;
}

// Shortest remaining work first: jobs sorted by the path slots they still
// need, each given all the room left (whole samples, or a slice of its
// current sample), so small jobs finish in the pass after they arrive.
void plan_srpt(xt_dev *d, ServeState &S, std::vector<Segment> &segs) {
    std::vector<std::pair<uint64_t, uint64_t>> order;  // (remaining slots, ticket)
    for (uint64_t t : S.active) {
        const SJob &J = S.jobs[t];
        if (J.issued < J.jd.total)
            order.push_back({(uint64_t)(J.jd.total - J.issued) * J.jd.NT - J.qc, t});
    }
    std::stable_sort(order.begin(), order.end(),
                     [](const std::pair<uint64_t, uint64_t> &a, const std::pair<uint64_t, uint64_t> &b) {
                         return a.first < b.first;
                     });
    int64_t room = std::max<int64_t>(64, d->tuning.wave_paths & ~63u);
    std::vector<std::tuple<uint32_t, uint32_t, uint32_t>> keys;
    for (const auto &e : order) {
        if (room < 64) break;
        const SJob &J = S.jobs[e.second];
        const auto key = std::make_tuple(J.job.scene_id, J.jd.integrator, J.job.accel);
        if (std::find(keys.begin(), keys.end(), key) == keys.end()) {
            if (keys.size() >= XT_MAX_LANES) continue;
            keys.push_back(key);
        }
        const int64_t nt = J.jd.NT;
        if (J.qc == 0 && nt <= room) {
            const int64_t k = std::min<int64_t>(J.jd.total - J.issued, room / nt);
            segs.push_back({e.second, (uint32_t)k, 0u, (uint32_t)nt, 0});
            room -= k * nt;
        } else {
            const int64_t len = std::min<int64_t>(nt - J.qc, room & ~63ll);
            segs.push_back({e.second, 1u, J.qc, (uint32_t)len, 0});
            room -= len;
        }
    }
}

// Picks the segments of the next pass (see the scheduling note above). A
// pass never exceeds the wavefront size: a job whose image has more tile
// positions than its share gets a slice of its current sample (a multiple
// of 64 positions, one 8x8 tile), and finishes that sample in later passes
// before starting the next, so every pixel still sees its samples in order.
void plan_pass(xt_dev *d, ServeState &S, std::vector<Segment> &segs) {
    segs.clear();
    if (S.policy == XT_SERVE_SRPT) {
        plan_srpt(d, S, segs);
        return;
    }
    // the oldest unfinished job first (it may take half of the pass), then
    // the others in round-robin order sharing the rest
    std::vector<uint64_t> order;
    uint64_t oldest = 0;
    for (uint64_t t : S.active) {
        const SJob &J = S.jobs[t];
        if (J.issued < J.jd.total) {
            oldest = t;
            break;
        }
    }
    if (!oldest) return;
    order.push_back(oldest);
    for (size_t k = 0; k < S.active.size(); ++k) {
        const uint64_t t = S.active[(S.rr + k) % S.active.size()];
        const SJob &J = S.jobs[t];
        if (t != oldest && J.issued < J.jd.total) order.push_back(t);
    }
    const int64_t C = std::max<int64_t>(64, d->tuning.wave_paths & ~63u);
    const int64_t share = std::max<int64_t>(64, (C / (int64_t)order.size()) & ~63ll);
    const int64_t first_share = std::max<int64_t>(share, (C / 2) & ~63ll);
    int64_t room = C;
    std::vector<std::tuple<uint32_t, uint32_t, uint32_t>> keys;  // groups in use
    for (size_t oi = 0; oi < order.size(); ++oi) {
        const uint64_t t = order[oi];
        if (room < 64) break;
        const SJob &J = S.jobs[t];
        const auto key = std::make_tuple(J.job.scene_id, J.jd.integrator, J.job.accel);
        if (std::find(keys.begin(), keys.end(), key) == keys.end()) {
            if (keys.size() >= XT_MAX_LANES) continue;
            keys.push_back(key);
        }
        const int64_t nt = J.jd.NT, give = std::min(oi == 0 ? first_share : share, room);
        if (J.qc == 0 && nt <= give) {
            const int64_t k = std::min<int64_t>(J.jd.total - J.issued, give / nt);
            segs.push_back({t, (uint32_t)k, 0u, (uint32_t)nt, 0});
            room -= k * nt;
        } else {
            const int64_t len = std::min<int64_t>(nt - J.qc, give & ~63ll);
            segs.push_back({t, 1u, J.qc, (uint32_t)len, 0});
            room -= len;
        }
    }
    for (Segment &g : segs) {  // leftover room, same order
        if (room < 64) break;
        const SJob &J = S.jobs[g.ticket];
        const int64_t nt = J.jd.NT;
        if (g.q0 == 0 && g.qn == nt) {
            const int64_t extra = std::min<int64_t>((int64_t)J.jd.total - J.issued - g.samples, room / nt);
            if (extra > 0) {
                g.samples += (uint32_t)extra;
                room -= extra * nt;
            }
        } else {
            const int64_t extra = std::min<int64_t>(nt - (g.q0 + g.qn), room & ~63ll);
            if (extra > 0) {
                g.qn += (uint32_t)extra;
                room -= extra;
            }
        }
    }
    S.rr = S.active.empty() ? 0 : (S.rr + 1) % S.active.size();
}

int run_pass(xt_dev *d, ServeState &S, const std::vector<Segment> &segs, std::vector<uint64_t> &finished) {
    // group segments by (scene, integrator, accel), lanes contiguous
    struct Grp {
        uint32_t scene, integrator, accel;
        std::vector<size_t> segs;
    };
    std::vector<Grp> grps;
    for (size_t k = 0; k < segs.size(); ++k) {
        const SJob &J = S.jobs[segs[k].ticket];
        size_t g = 0;
        while (g < grps.size() && !(grps[g].scene == J.job.scene_id && grps[g].integrator == J.jd.integrator &&
                                    grps[g].accel == J.job.accel))
            ++g;
        if (g == grps.size()) grps.push_back({J.job.scene_id, J.jd.integrator, J.job.accel, {}});
        grps[g].segs.push_back(k);
    }
    std::vector<Segment> placed(segs);
    std::vector<PassGroup> groups(grps.size());
    size_t total = 0;
    bool need64 = false, need_chan = false;
    for (size_t g = 0; g < grps.size(); ++g) {
        groups[g].lane.off = total;
        uint32_t rounds = 0;
        for (size_t k : grps[g].segs) {
            const SJob &J = S.jobs[segs[k].ticket];
            placed[k].path_base = total;
            total += (size_t)segs[k].samples * segs[k].qn;
            rounds = std::max(rounds, xt_rounds(J.jd));
        }
        groups[g].lane.cap = total - groups[g].lane.off;
        groups[g].integrator = grps[g].integrator;
        groups[g].accel = grps[g].accel;
        groups[g].rounds = rounds;
        groups[g].ds = &d->scenes.at(grps[g].scene);
        need64 = need64 || xt_needs_ray64(grps[g].integrator);
        need_chan = need_chan || xt_needs_chan(grps[g].integrator, groups[g].ds->view);
    }
    if (total >= (1ull << 31)) {
        d->err = "pass too large";
        return XT_E_INVALID;
    }
    int rc = xt_dev_ensure_wave(d, total, need64, need_chan);
    if (rc) return rc;
    WaveBuffers &w = d->wave;
    for (size_t g = 0; g < grps.size(); ++g) groups[g].lane.cnt = w.cnt + g;
    XT_TRY(hipEventRecord(S.p0, d->stream));
    hipLaunchKernelGGL(k_reset_lanes, dim3((unsigned)grps.size()), dim3(64), 0, d->stream, w.cnt, 0x800000ffu);
    XT_LAUNCHED();
    for (size_t g = 0; g < grps.size(); ++g) {
        const WaveLane &L = groups[g].lane;
        // camera rays fill the lane's queue 0 in slot order
        hipLaunchKernelGGL(k_set_count, dim3(1), dim3(1), 0, d->stream, L.cnt, 0u, (uint32_t)L.cap);
        for (size_t k : grps[g].segs) {
            const SJob &J = S.jobs[segs[k].ticket];
            const size_t n = (size_t)segs[k].samples * segs[k].qn;
            const size_t at = placed[k].path_base;
            Queue q0{w.q[0].r0 + at, w.q[0].r1 + at, w.q[0].path + at};
            hipLaunchKernelGGL(k_camera, dim3(grid_for(n, XT_BLOCK, 8192)), dim3(XT_BLOCK), 0, d->stream, J.jd,
                               J.issued, segs[k].samples, segs[k].q0, segs[k].qn, (uint32_t)at, q0, w.ps0, w.ps1,
                               w.rad, need64 ? w.ray64 : nullptr, (uint32_t)(d->tuning.ray_order == 2));
            XT_LAUNCHED();
        }
    }
    rc = xt_run_rounds(d, groups.data(), (int)groups.size(), false);
    if (rc) return rc;
    for (size_t k = 0; k < segs.size(); ++k) {
        SJob &J = S.jobs[segs[k].ticket];
        hipLaunchKernelGGL(k_accumulate, dim3(grid_for(segs[k].qn, XT_BLOCK, 8192)), dim3(XT_BLOCK), 0, d->stream,
                           J.jd, J.issued, segs[k].samples, segs[k].q0, segs[k].qn, w.rad + placed[k].path_base,
                           need64 ? w.ray64 + 6 * placed[k].path_base : nullptr, J.accf, J.acc64);
        XT_LAUNCHED();
        if (segs[k].qn == J.jd.NT) {
            J.issued += segs[k].samples;
        } else {
            J.qc += segs[k].qn;
            if (J.qc >= J.jd.NT) {
                J.qc = 0;
                J.issued += 1;
            }
        }
        J.passes += 1;
        if (J.issued >= J.jd.total) {
            hipLaunchKernelGGL(k_resolve, dim3(grid_for(J.npix, XT_BLOCK, 8192)), dim3(XT_BLOCK), 0, d->stream, J.jd,
                               J.accf, J.acc64, J.out);
            XT_LAUNCHED();
            rc = take_host_block(d, S, J.npix * sizeof(float4), J.host);
            if (rc) return rc;
            XT_TRY(hipMemcpyAsync(J.host.p, J.out, J.npix * sizeof(float4), hipMemcpyDeviceToHost, d->stream));
            XT_TRY(hipEventRecord(J.done_ev, d->stream));
            finished.push_back(segs[k].ticket);
        }
        S.info.samples += (uint64_t)segs[k].samples * J.npix * segs[k].qn / J.jd.NT;
    }
    XT_TRY(hipEventRecord(S.p1, d->stream));
    XT_TRY(hipMemcpyAsync(d->host_counters, w.cnt, grps.size() * sizeof(Counters), hipMemcpyDeviceToHost, d->stream));
    XT_TRY(hipStreamSynchronize(d->stream));
    for (size_t g = 0; g < grps.size(); ++g) {
        S.info.queries += d->host_counters[g].queries;
        S.info.retests += d->host_counters[g].retests;
    }
    float ms = 0.f;
    hipEventElapsedTime(&ms, S.p0, S.p1);
    S.info.ms_device += ms;
    S.info.passes += 1;
    return XT_OK;
}

void release_block(ServeState &S, SJob &J) {
    if (J.blk.p) S.free_dev.push_back(J.blk);
    J.blk = DevBlock();
    J.accf = J.out = nullptr;
    J.acc64 = nullptr;
}

}  // namespace

bool xt_serve_uses_scene(const xt_dev *d, uint32_t scene_id) {
    if (!d->serve) return false;
    for (const auto &kv : d->serve->jobs)
        if (kv.second.state != FINISHED && kv.second.job.scene_id == scene_id) return true;
    return false;
}

void xt_serve_shutdown(xt_dev *d) {
    ServeState *S = d->serve;
    if (!S) return;
    hipSetDevice(d->device);
    hipStreamSynchronize(d->stream);
    for (auto &kv : S->jobs) {
        if (kv.second.blk.p) {
            hipFree(kv.second.blk.p);
            d->used -= std::min(d->used, kv.second.blk.bytes);
        }
        if (kv.second.host.p) hipHostFree(kv.second.host.p);
        if (kv.second.done_ev) hipEventDestroy(kv.second.done_ev);
    }
    for (const DevBlock &b : S->free_dev) {
        hipFree(b.p);
        d->used -= std::min(d->used, b.bytes);
    }
    for (const HostBlock &b : S->free_host) hipHostFree(b.p);
    for (hipEvent_t e : S->free_events) hipEventDestroy(e);
    if (S->epoch) hipEventDestroy(S->epoch);
    if (S->p0) hipEventDestroy(S->p0);
    if (S->p1) hipEventDestroy(S->p1);
    delete S;
    d->serve = nullptr;
}

extern "C" int xt_submit(xt_dev *d, const xt_job *job, uint64_t *ticket) {
    if (!d || !job || !ticket) return XT_E_INVALID;
    JobDev jd;
    const DevScene *ds = nullptr;
    int rc = xt_job_check(d, job, jd, &ds);
    if (rc) return rc;
    if (jd.total == 0) {
        d->err = "job has no samples";
        return XT_E_INVALID;
    }
    rc = serve_init(d);
    if (rc) return rc;
    ServeState &S = *d->serve;
    const size_t npix = (size_t)jd.W * jd.H;
    // memory that job buffers can never use: scenes and wavefront buffers
    const uint64_t fixed = d->used - std::min(d->used, S.held);
    if (fixed + job_bytes(npix) > d->budget) {
        d->err = "job buffers exceed the HBM budget";
        return XT_E_BUDGET;
    }
    const uint64_t t = S.next_ticket++;
    SJob &J = S.jobs[t];
    J.job = *job;
    J.jd = jd;
    J.npix = npix;
    J.t_submit = now_ms();
    S.waiting.push_back(t);
    *ticket = t;
    return XT_OK;
}

extern "C" int xt_poll(xt_dev *d, uint64_t ticket, int *done, float *rgba_out, xt_job_stats *stats) {
    if (!d || !done) return XT_E_INVALID;
    *done = 0;
    if (!d->serve) return XT_E_NOTFOUND;
    ServeState &S = *d->serve;
    auto it = S.jobs.find(ticket);
    if (it == S.jobs.end()) {
        d->err = "unknown ticket";
        return XT_E_NOTFOUND;
    }
    SJob &J = it->second;
    if (stats) {
        std::memset(stats, 0, sizeof *stats);
        stats->samples = (uint64_t)J.issued * J.npix;
        stats->passes = J.passes;
        stats->queue_ms = J.state == WAITING ? now_ms() - J.t_submit : J.t_admit - J.t_submit;
    }
    if (J.state != FINISHED) return XT_OK;
    if (stats) {
        stats->latency_ms = J.t_done - J.t_submit;
        stats->ms_total = J.t_done - J.t_admit;
    }
    *done = 1;
    if (!rgba_out) return XT_OK;  // report only: the image stays until it is fetched
    std::memcpy(rgba_out, J.host.p, J.npix * sizeof(float4));
    S.free_host.push_back(J.host);
    S.free_events.push_back(J.done_ev);
    S.jobs.erase(it);
    return XT_OK;
}

extern "C" int xt_serve_step(xt_dev *d, int *active_jobs) {
    if (!d) return XT_E_INVALID;
    if (active_jobs) *active_jobs = 0;
    if (!d->serve) return XT_OK;
    ServeState &S = *d->serve;
    XT_TRY(hipSetDevice(d->device));
    int rc = admit(d, S);
    if (rc) return rc;
    std::vector<Segment> segs;
    plan_pass(d, S, segs);
    if (!segs.empty()) {
        std::vector<uint64_t> finished;
        rc = run_pass(d, S, segs, finished);
        if (rc) return rc;
        for (uint64_t t : finished) {
            SJob &J = S.jobs[t];
            float ms = 0.f;
            hipEventElapsedTime(&ms, S.epoch, J.done_ev);
            J.t_done = S.epoch_ms + ms;
            J.state = FINISHED;
            release_block(S, J);
            S.active.erase(std::find(S.active.begin(), S.active.end(), t));
            S.info.jobs_finished += 1;
        }
        if (!S.active.empty()) S.rr %= S.active.size();
        // room freed by finished jobs admits waiting ones for the next pass
        rc = admit(d, S);
        if (rc) return rc;
    }
    if (S.active.empty() && !S.waiting.empty()) {
        // nothing runs and the first waiting job still does not fit: the
        // memory outside the job pool grew after it was submitted
        d->err = "waiting job does not fit the HBM budget";
        return XT_E_BUDGET;
    }
    if (active_jobs) *active_jobs = (int)(S.active.size() + S.waiting.size());
    return XT_OK;
}

extern "C" int xt_drain(xt_dev *d) {
    if (!d) return XT_E_INVALID;
    for (;;) {
        int n = 0;
        int rc = xt_serve_step(d, &n);
        if (rc) return rc;
        if (n == 0) return XT_OK;
    }
}

extern "C" int xt_serve_set_policy(xt_dev *d, int policy) {
    if (!d || (policy != XT_SERVE_FAIR && policy != XT_SERVE_SRPT)) return XT_E_INVALID;
    int rc = serve_init(d);
    if (rc) return rc;
    d->serve->policy = policy;
    return XT_OK;
}

extern "C" int xt_serve_get_info(const xt_dev *d, xt_serve_info *info) {
    if (!d || !info) return XT_E_INVALID;
    std::memset(info, 0, sizeof *info);
    if (!d->serve) return XT_OK;
    const ServeState &S = *d->serve;
    *info = S.info;
    info->jobs_waiting = (uint32_t)S.waiting.size();
    info->jobs_active = (uint32_t)S.active.size();
    info->bytes_held = S.held;
    return XT_OK;
}
