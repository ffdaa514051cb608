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

#include "device/xt_regen.h"
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

// regeneration engine: per pixel, the resolved image and either a radiance
// record per sample or 3 fixed-point sums (jobs whose sample buffer would
// exceed the configured size)
uint64_t regen_sample_bytes(size_t npix, const JobDev &jd) { return (uint64_t)npix * jd.total * sizeof(float4); }

bool regen_ordered(const xt_serve_config &cfg, size_t npix, const JobDev &jd) {
    return regen_sample_bytes(npix, jd) <= ((uint64_t)cfg.sample_buffer_mb << 20);
}

uint64_t regen_job_bytes(const xt_serve_config &cfg, size_t npix, const JobDev &jd) {
    const uint64_t per = regen_ordered(cfg, npix, jd) ? regen_sample_bytes(npix, jd) : (uint64_t)npix * 3 * sizeof(long long);
    return (per + (uint64_t)npix * sizeof(float4) + 511) & ~255ull;
}

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
    bool regen = false;    // runs on the path-regeneration engine
    int entry = -1;        // its device job-table entry once started there
    uint64_t paths_done = 0;
    long long *acc = nullptr;
    float4 *sbuf = nullptr;  // regeneration engine: radiance per (pixel, sample), or fixed-point sums in acc
    uint32_t issued = 0;  // whole samples handed to passes
    uint32_t qc = 0;      // tile positions of sample `issued` handed out so far
    uint32_t passes = 0;
    size_t npix = 0;
    DevBlock blk;
    HostBlock host;
    float4 *accf = nullptr, *out = nullptr;
    double2 *acc64 = nullptr;
    hipEvent_t done_ev = nullptr;
    bool copying = false;  // device-fed serving: resolved, its image on the way to host memory
    bool zeroing = false;  // its fixed-point sums are being cleared (done_ev marks the end)
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

// Host side of the path-regeneration engine (see device/xt_regen.h).
struct RegenLaneH {
    bool used = false;
    uint32_t scene = 0, integrator = 0, accel = 0;
    void *blk = nullptr;
    uint64_t bytes = 0;
    float4 *hit = nullptr, *chan = nullptr;
    uint32_t *retest = nullptr;
    uint4 *recheck = nullptr;
};

struct RegenHost {
    bool ready = false;
    uint32_t P = 0;
    void *blk = nullptr;
    uint64_t bytes = 0;
    RegenView V{};
    Queue *lq = nullptr;        // device copy of lq_host
    RegenState *ls = nullptr;   // device copy of ls_host
    Queue lq_host[2 * XT_REGEN_MAX_LANES] = {};
    RegenState ls_host[2 * XT_REGEN_MAX_LANES] = {};
    RegenLaneH lanes[XT_REGEN_MAX_LANES];
    // fork / join events of a round's concurrent lanes, per queue parity
    hipEvent_t fork_ev[2] = {}, join_ev[2][XT_LANE_STREAMS] = {};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool dirty = true;
    xt_tuning captured{};
    RegenJob *h_job = nullptr;             // pinned staging
    unsigned long long *h_done = nullptr;  // pinned, per table entry
    uint32_t *h_prio = nullptr;            // pinned: n_prio, cap_per_job, prio[]
    Counters *h_cnt = nullptr;             // pinned, per lane
    std::vector<int> free_entries;
    unsigned long long queries_seen[XT_REGEN_MAX_LANES] = {}, retests_seen[XT_REGEN_MAX_LANES] = {};
    // device-fed serving (xt_serve_config.loop = 2)
    RegenRing *ring = nullptr;      // fine-grained host memory
    RegenRing *ring_dev = nullptr;  // its device address
    uint32_t req_head = 0, fin_tail = 0;
    unsigned long long seq = 0;
    uint64_t entry_ticket[XT_REGEN_MAX_JOBS] = {};
    std::deque<std::pair<hipEvent_t, int>> inflight;  // replays on the stream: end event, counter snapshot slot
    std::vector<hipEvent_t> free_ev;
    Counters *h_snap = nullptr;     // pinned: lane counters after a replay, XT_REGEN_INFLIGHT slots
    uint32_t *h_policy = nullptr;   // pinned
    uint32_t policy_sent = 0xffffffffu;
    uint64_t launches = 0;
    std::vector<uint64_t> resolving;  // tickets whose image is being copied to host memory
};

#define XT_REGEN_INFLIGHT 2  // device-fed serving: replays kept queued on the stream

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
    xt_serve_config cfg{2, 2, 4, 0, 4096, 2, 2};
    RegenHost R;
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

bool regen_eligible(const ServeState &S, const JobDev &jd) {
    return S.cfg.regen == 2 && jd.accum == XT_ACCUM_FP64 &&
           (jd.integrator == XT_INTEGRATOR_AO || jd.integrator == XT_INTEGRATOR_PATHTRACER);
}

// Admits waiting jobs in FIFO order while their buffers fit.
int admit(xt_dev *d, ServeState &S) {
    while (!S.waiting.empty()) {
        SJob &J = S.jobs[S.waiting.front()];
        bool ok = false;
        J.regen = regen_eligible(S, J.jd);
        int rc = take_dev_block(d, S, J.regen ? regen_job_bytes(S.cfg, J.npix, J.jd) : job_bytes(J.npix), J.blk, ok);
        if (rc) return rc;
        if (!ok) break;
        char *base = (char *)J.blk.p;
        if (J.regen) {
            // the resolved image, then the sample buffer or 3 fixed-point sums
            J.out = (float4 *)base;
            char *rest = (char *)(((uintptr_t)(base + J.npix * sizeof(float4)) + 255) & ~(uintptr_t)255);
            J.accf = nullptr;
            J.acc64 = nullptr;
            J.entry = -1;
            if (regen_ordered(S.cfg, J.npix, J.jd)) {
                J.sbuf = (float4 *)rest;
                J.acc = nullptr;
            } else {
                J.sbuf = nullptr;
                J.acc = (long long *)rest;
                hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(J.npix * 6, 256, 8192)), dim3(256), 0, d->stream,
                                   (uint32_t *)J.acc, 0u, J.npix * 6);
            }
        } else {
            J.accf = (float4 *)base;
            J.acc64 = (double2 *)(base + J.npix * sizeof(float4));
            J.out = (float4 *)(base + J.npix * (sizeof(float4) + 3 * sizeof(double2)));
            hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(J.npix * 16, 256, 8192)), dim3(256), 0, d->stream,
                               (uint32_t *)J.blk.p, 0u, J.npix * 16);
        }
        XT_LAUNCHED();
        if (!S.free_events.empty()) {
            J.done_ev = S.free_events.back();
            S.free_events.pop_back();
        } else {
            XT_TRY(hipEventCreate(&J.done_ev));
        }
        if (J.regen && J.acc) {
            // device-fed serving hands the job to the device only once its
            // sums are zero (a replay already queued could take it earlier)
            XT_TRY(hipEventRecord(J.done_ev, d->stream));
            J.zeroing = true;
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
        if (!J.regen && J.issued < J.jd.total)
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
        if (!J.regen && J.issued < J.jd.total) {
            oldest = t;
            break;
        }
    }
    if (!oldest) return;
    order.push_back(oldest);
    for (size_t k = 0; k < S.active.size(); ++k) {
        const uint64_t t = S.active[(S.rr + k) % S.active.size()];
        const SJob &J = S.jobs[t];
        if (t != oldest && !J.regen && J.issued < J.jd.total) order.push_back(t);
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
                               w.rad, need64 ? w.ray64 : nullptr, d->tuning.ray_order);
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


// ------------------------------------------------- path-regeneration engine


int regen_malloc(xt_dev *d, void **p, uint64_t bytes) {
    if (d->used + bytes > d->budget) {
        d->err = "regeneration engine buffers exceed the HBM budget";
        return XT_E_BUDGET;
    }
    XT_TRY(hipMalloc(p, bytes));
    d->used += bytes;
    return XT_OK;
}

void regen_free(xt_dev *d, void *&p, uint64_t &bytes) {
    if (p) {
        (void)hipFree(p);
        d->used -= std::min(d->used, bytes);
    }
    p = nullptr;
    bytes = 0;
}

inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Shared buffers of the engine: retire list, job table, done counts,
// control block, segments, lane counters and the lane queue tables.
int regen_init(xt_dev *d, ServeState &S) {
    RegenHost &R = S.R;
    if (R.ready) return XT_OK;
    const uint32_t P = S.cfg.paths ? S.cfg.paths : XT_REGEN_PATHS;
    const size_t sz[] = {16ull * P, 8ull * P, sizeof(RegenJob) * XT_REGEN_MAX_JOBS, 8ull * XT_REGEN_MAX_JOBS,
                         sizeof(RegenCtl), sizeof(RegenSeg) * XT_REGEN_MAX_JOBS,
                         sizeof(Counters) * XT_REGEN_MAX_LANES, sizeof(Queue) * 2 * XT_REGEN_MAX_LANES,
                         sizeof(RegenState) * 2 * XT_REGEN_MAX_LANES,
                         8ull * XT_REGEN_MAX_JOBS * XT_REGEN_SHARDS};
    size_t off[sizeof(sz) / sizeof(sz[0])], total = 0;
    for (size_t k = 0; k < sizeof(sz) / sizeof(sz[0]); ++k) {
        off[k] = total;
        total += align256(sz[k]);
    }
    int rc = regen_malloc(d, &R.blk, total);
    if (rc) return rc;
    R.bytes = total;
    char *b = (char *)R.blk;
    RegenView &V = R.V;
    V.ret = (float4 *)(b + off[0]);
    V.ret_js = (uint2 *)(b + off[1]);
    V.jobs = (RegenJob *)(b + off[2]);
    V.done = (unsigned long long *)(b + off[3]);
    V.ctl = (RegenCtl *)(b + off[4]);
    V.segs = (RegenSeg *)(b + off[5]);
    V.cnt = (Counters *)(b + off[6]);
    R.lq = (Queue *)(b + off[7]);
    R.ls = (RegenState *)(b + off[8]);
    V.done_shard = (unsigned long long *)(b + off[9]);
    V.lq = R.lq;
    V.ls = R.ls;
    for (int k = 0; k < XT_LANE_STREAMS; ++k)
        if (!d->lane_stream[k]) {
            XT_TRY(hipStreamCreateWithFlags(&d->lane_stream[k], hipStreamNonBlocking));
            XT_TRY(hipEventCreateWithFlags(&d->lane_ev[k], hipEventDisableTiming));
        }
    for (int q = 0; q < 2; ++q) {
        XT_TRY(hipEventCreateWithFlags(&R.fork_ev[q], hipEventDisableTiming));
        for (int k = 0; k < XT_LANE_STREAMS; ++k) XT_TRY(hipEventCreateWithFlags(&R.join_ev[q][k], hipEventDisableTiming));
    }
    XT_TRY(hipHostMalloc((void **)&R.ring, sizeof(RegenRing), hipHostMallocCoherent | hipHostMallocMapped));
    std::memset(R.ring, 0, sizeof(RegenRing));
    XT_TRY(hipHostGetDevicePointer((void **)&R.ring_dev, R.ring, 0));
    XT_TRY(hipHostMalloc((void **)&R.h_snap, sizeof(Counters) * XT_REGEN_MAX_LANES * XT_REGEN_INFLIGHT,
                         hipHostMallocDefault));
    XT_TRY(hipHostMalloc((void **)&R.h_policy, 4, hipHostMallocDefault));
    R.req_head = R.fin_tail = 0;
    R.policy_sent = 0xffffffffu;
    XT_TRY(hipHostMalloc((void **)&R.h_job, sizeof(RegenJob) * XT_REGEN_MAX_JOBS, hipHostMallocDefault));
    XT_TRY(hipHostMalloc((void **)&R.h_done, 8ull * XT_REGEN_MAX_JOBS, hipHostMallocDefault));
    XT_TRY(hipHostMalloc((void **)&R.h_prio, 4ull * (XT_REGEN_MAX_JOBS + 2), hipHostMallocDefault));
    XT_TRY(hipHostMalloc((void **)&R.h_cnt, sizeof(Counters) * XT_REGEN_MAX_LANES, hipHostMallocDefault));
    std::memset(R.lq_host, 0, sizeof R.lq_host);
    std::memset(R.ls_host, 0, sizeof R.ls_host);
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(align256(sz[4]) / 4, 256, 64)), dim3(256), 0, d->stream,
                       (uint32_t *)V.ctl, 0u, align256(sz[4]) / 4);
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(2ull * XT_REGEN_MAX_JOBS, 256, 64)), dim3(256), 0, d->stream,
                       (uint32_t *)V.done, 0u, 2ull * XT_REGEN_MAX_JOBS);
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(2ull * XT_REGEN_MAX_JOBS * XT_REGEN_SHARDS, 256, 512)), dim3(256), 0,
                       d->stream, (uint32_t *)V.done_shard, 0u, 2ull * XT_REGEN_MAX_JOBS * XT_REGEN_SHARDS);
    hipLaunchKernelGGL(k_regen_init, dim3(1), dim3(1), 0, d->stream, V.ctl, P);
    hipLaunchKernelGGL(k_reset_lanes, dim3(XT_REGEN_MAX_LANES), dim3(64), 0, d->stream, V.cnt, 0x800000ffu);
    XT_LAUNCHED();
    XT_TRY(hipMemcpyAsync(R.lq, R.lq_host, sizeof R.lq_host, hipMemcpyHostToDevice, d->stream));
    XT_TRY(hipMemcpyAsync(R.ls, R.ls_host, sizeof R.ls_host, hipMemcpyHostToDevice, d->stream));
    XT_TRY(hipStreamSynchronize(d->stream));
    R.P = P;
    R.free_entries.clear();
    for (int e = XT_REGEN_MAX_JOBS - 1; e >= 0; --e) R.free_entries.push_back(e);
    for (auto &q : R.queries_seen) q = 0;
    for (auto &q : R.retests_seen) q = 0;
    R.ready = true;
    R.dirty = true;
    return XT_OK;
}

void regen_drop_graph(RegenHost &R) {
    if (R.exec) (void)hipGraphExecDestroy(R.exec);
    if (R.graph) (void)hipGraphDestroy(R.graph);
    R.exec = nullptr;
    R.graph = nullptr;
    R.dirty = true;
}

void regen_shutdown(xt_dev *d, ServeState &S) {
    RegenHost &R = S.R;
    if (R.ready) (void)hipStreamSynchronize(d->stream);  // replays still queued use the buffers
    regen_drop_graph(R);
    for (RegenLaneH &L : R.lanes) regen_free(d, L.blk, L.bytes);
    regen_free(d, R.blk, R.bytes);
    if (R.h_job) (void)hipHostFree(R.h_job);
    if (R.h_done) (void)hipHostFree(R.h_done);
    if (R.h_prio) (void)hipHostFree(R.h_prio);
    if (R.h_cnt) (void)hipHostFree(R.h_cnt);
    if (R.ring) (void)hipHostFree(R.ring);
    if (R.h_snap) (void)hipHostFree(R.h_snap);
    if (R.h_policy) (void)hipHostFree(R.h_policy);
    for (auto &f : R.inflight) (void)hipEventDestroy(f.first);
    for (hipEvent_t e : R.free_ev) (void)hipEventDestroy(e);
    for (int q = 0; q < 2; ++q) {
        if (R.fork_ev[q]) (void)hipEventDestroy(R.fork_ev[q]);
        for (int k = 0; k < XT_LANE_STREAMS; ++k)
            if (R.join_ev[q][k]) (void)hipEventDestroy(R.join_ev[q][k]);
    }
    R = RegenHost();
}

// True when an unfinished started job runs on lane l.
bool regen_lane_busy(const ServeState &S, const RegenLaneH &L) {
    for (uint64_t t : S.active) {
        const SJob &J = S.jobs.at(t);
        if (J.regen && J.entry >= 0 && J.job.scene_id == L.scene && J.jd.integrator == L.integrator &&
            J.job.accel == L.accel)
            return true;
    }
    return false;
}

// Gives lane l's buffers back (its queues are empty: every path of its jobs
// has ended).
void regen_drop_lane(xt_dev *d, RegenHost &R, int l) {
    regen_free(d, R.lanes[l].blk, R.lanes[l].bytes);
    R.lanes[l] = RegenLaneH();
    R.dirty = true;
}

// The lane of (scene, integrator, accel), created on first use (its queues,
// hit and re-test buffers sized for every path slot). -1: all lanes busy.
int regen_lane(xt_dev *d, ServeState &S, const SJob &J, int &lane) {
    RegenHost &R = S.R;
    lane = -1;
    int idle = -1;
    for (int l = 0; l < XT_REGEN_MAX_LANES; ++l) {
        const RegenLaneH &L = R.lanes[l];
        if (L.used && L.scene == J.job.scene_id && L.integrator == J.jd.integrator && L.accel == J.job.accel) {
            lane = l;
            return XT_OK;
        }
        if (!L.used && idle < 0) idle = l;
    }
    if (idle < 0) {
        // every lane is set up: take one whose jobs have all finished (lanes
        // stay set up between jobs, so a scene's next job starts at once)
        for (int l = 0; l < XT_REGEN_MAX_LANES && idle < 0; ++l)
            if (!regen_lane_busy(S, R.lanes[l])) idle = l;
        if (idle < 0) return XT_OK;
        regen_drop_lane(d, R, idle);
    }
    RegenLaneH &L = R.lanes[idle];

    const DevScene &ds = d->scenes.at(J.job.scene_id);
    const bool chan = xt_needs_chan(J.jd.integrator, ds.view);
    const uint64_t P = R.P;
    const size_t qsz = align256(16 * P), psz = align256(4 * P), isz = qsz;
    // per parity: r0, r1, path, s0, s1, info; then hit, retest, recheck, channels
    const size_t total = 2 * (4 * qsz + psz + isz) + qsz + psz + qsz + (chan ? 3 * qsz : 0);
    int rc = regen_malloc(d, &L.blk, total);
    if (rc) return rc;
    L.bytes = total;
    char *b = (char *)L.blk;
    for (int qi = 0; qi < 2; ++qi) {
        Queue &q = R.lq_host[qi * XT_REGEN_MAX_LANES + idle];
        RegenState &st = R.ls_host[qi * XT_REGEN_MAX_LANES + idle];
        q.r0 = (float4 *)b;
        q.r1 = (float4 *)(b + qsz);
        st.s0 = (float4 *)(b + 2 * qsz);
        st.s1 = (float4 *)(b + 3 * qsz);
        q.path = (uint32_t *)(b + 4 * qsz);
        st.info = (uint4 *)(b + 4 * qsz + psz);
        b += 4 * qsz + psz + isz;
    }
    L.hit = (float4 *)b;
    b += qsz;
    L.retest = (uint32_t *)b;
    b += psz;
    L.recheck = (uint4 *)b;
    b += qsz;
    L.chan = chan ? (float4 *)b : nullptr;
    L.used = true;
    L.scene = J.job.scene_id;
    L.integrator = J.jd.integrator;
    L.accel = J.job.accel;
    hipLaunchKernelGGL(k_reset_lanes, dim3(1), dim3(64), 0, d->stream, R.V.cnt + idle, 0x800000ffu);
    XT_LAUNCHED();
    XT_TRY(hipMemcpyAsync(R.lq, R.lq_host, sizeof R.lq_host, hipMemcpyHostToDevice, d->stream));
    XT_TRY(hipMemcpyAsync(R.ls, R.ls_host, sizeof R.ls_host, hipMemcpyHostToDevice, d->stream));
    R.queries_seen[idle] = R.retests_seen[idle] = 0;
    R.dirty = true;
    lane = idle;
    return XT_OK;
}

// One query round over queue parity qi: plan, new camera rays, trace and
// shade per lane, retire.
int regen_round(xt_dev *d, ServeState &S, uint32_t qi) {
    RegenHost &R = S.R;
    RegenView &V = R.V;
    const hipStream_t st = d->stream;
    const unsigned gP = grid_for(R.P, XT_BLOCK, 8192);
    if (S.cfg.loop == 2)
        hipLaunchKernelGGL(k_regen_intake, dim3(1), dim3(256), 0, st, V, (const RegenRing *)R.ring_dev);
    hipLaunchKernelGGL(k_regen_plan, dim3(1), dim3(64), 0, st, V, qi);
    hipLaunchKernelGGL(k_regen_gen, dim3(gP), dim3(XT_BLOCK), 0, st, V, qi);
    hipLaunchKernelGGL(k_regen_round_start, dim3(1), dim3(1), 0, st, V.ctl);
    XT_LAUNCHED();
    // lanes share nothing but the retire list, the job table and the done
    // shards, which they update with atomics whose results do not depend on
    // their order: with several lanes they run on their own streams, forked
    // from and joined back into the round's stream
    int used[XT_REGEN_MAX_LANES], nu = 0;
    for (int l = 0; l < XT_REGEN_MAX_LANES; ++l)
        if (R.lanes[l].used) used[nu++] = l;
    const bool par = S.cfg.lane_streams == 2 && nu > 1;
    const int ns = par ? std::min(nu, XT_LANE_STREAMS) : 0;
    if (par) {
        XT_TRY(hipEventRecord(R.fork_ev[qi], st));
        for (int k = 0; k < ns; ++k) XT_TRY(hipStreamWaitEvent(d->lane_stream[k], R.fork_ev[qi], 0));
    }
    for (int u = 0; u < nu; ++u) {
        const int l = used[u];
        RegenLaneH &L = R.lanes[l];
        const hipStream_t ls = par ? d->lane_stream[u % ns] : st;
        const DevScene &ds = d->scenes.at(L.scene);
        Counters *cnt = V.cnt + l;
        hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, ls, cnt,
                           (1u << (qi ^ 1)) | (1u << CNT_RETEST) | (1u << CNT_CURSOR) | (1u << CNT_RECHECK));
        const Queue in = R.lq_host[qi * XT_REGEN_MAX_LANES + l], out = R.lq_host[(qi ^ 1) * XT_REGEN_MAX_LANES + l];
        TraceBufs B;
        B.q = in;
        B.count = &cnt->n[qi];
        B.hit = L.hit;
        B.retest = L.retest;
        B.recheck = L.recheck;
        B.cnt = cnt;
        B.cap = R.P;
        int rc = xt_launch_trace_bufs(d, ds, (int)L.accel, B, nullptr, nullptr, false, ls);
        if (rc) return rc;
        const unsigned g = grid_for(R.P, XT_BLOCK, 8192);
        const RegenState ist = R.ls_host[qi * XT_REGEN_MAX_LANES + l], ost = R.ls_host[(qi ^ 1) * XT_REGEN_MAX_LANES + l];
        if (L.integrator == XT_INTEGRATOR_AO) {
            hipLaunchKernelGGL(k_regen_shade_ao, dim3(g), dim3(XT_BLOCK), 0, ls, ds.view, V, in, ist, &cnt->n[qi], L.hit,
                               out, ost, &cnt->n[qi ^ 1]);
        } else if (L.chan) {
            hipLaunchKernelGGL(k_channels, dim3(g), dim3(XT_BLOCK), 0, ls, ds.view, in, &cnt->n[qi], L.hit, L.chan);
            XT_LAUNCHED();
            hipLaunchKernelGGL(k_regen_shade_pt<true>, dim3(g), dim3(XT_BLOCK), 0, ls, ds.view, V, in, ist, &cnt->n[qi],
                               L.hit, out, ost, &cnt->n[qi ^ 1], (const float4 *)L.chan);
        } else {
            hipLaunchKernelGGL(k_regen_shade_pt<false>, dim3(g), dim3(XT_BLOCK), 0, ls, ds.view, V, in, ist,
                               &cnt->n[qi], L.hit, out, ost, &cnt->n[qi ^ 1], (const float4 *)nullptr);
        }
        XT_LAUNCHED();
    }
    if (par)
        for (int k = 0; k < ns; ++k) {
            XT_TRY(hipEventRecord(R.join_ev[qi][k], d->lane_stream[k]));
            XT_TRY(hipStreamWaitEvent(st, R.join_ev[qi][k], 0));
        }
    hipLaunchKernelGGL(k_regen_retire, dim3(gP), dim3(XT_BLOCK), 0, st, V);
    XT_LAUNCHED();
    return XT_OK;
}

// Device-fed serving, after a replay's two rounds: finish, resolve and post
// the jobs whose paths have all ended.
int regen_tail(xt_dev *d, ServeState &S) {
    RegenHost &R = S.R;
    hipLaunchKernelGGL(k_regen_done_sum, dim3(XT_REGEN_MAX_JOBS / 256), dim3(256), 0, d->stream, R.V);
    hipLaunchKernelGGL(k_regen_finish, dim3(1), dim3(256), 0, d->stream, R.V);
    hipLaunchKernelGGL(k_regen_resolve_fin, dim3(1024), dim3(XT_BLOCK), 0, d->stream, R.V);
    hipLaunchKernelGGL(k_regen_post, dim3(1), dim3(64), 0, d->stream, R.V, R.ring_dev);
    XT_LAUNCHED();
    return XT_OK;
}

// Captures rounds over parity 0 then 1 (queue parity is back at 0 after a
// replay) as one hipGraph.
int regen_capture(xt_dev *d, ServeState &S) {
    RegenHost &R = S.R;
    regen_drop_graph(R);
    XT_TRY(hipStreamBeginCapture(d->stream, hipStreamCaptureModeThreadLocal));
    int rc = regen_round(d, S, 0);
    if (!rc) rc = regen_round(d, S, 1);
    if (!rc && S.cfg.loop == 2) rc = regen_tail(d, S);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(d->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) {
        d->err = std::string("hipStreamEndCapture: ") + hipGetErrorString(e);
        return XT_E_DEVICE;
    }
    R.graph = g;
    XT_TRY(hipGraphInstantiate(&R.exec, R.graph, nullptr, nullptr, 0));
    R.captured = d->tuning;
    R.dirty = false;
    return XT_OK;
}

// Priority order of the started, unfinished jobs (table entries) by policy.
void regen_priorities(ServeState &S, std::vector<uint32_t> &prio, uint32_t &cap) {
    std::vector<std::pair<uint64_t, const SJob *>> js;  // (admission rank, job)
    uint64_t rank = 0;
    for (uint64_t t : S.active) {
        const SJob &J = S.jobs[t];
        ++rank;
        if (J.regen && J.entry >= 0) js.push_back({rank, &J});
    }
    auto left = [](const SJob *J) { return (uint64_t)J->jd.total * J->jd.NT - J->paths_done; };
    cap = 0;
    const int pol = S.policy;
    if (pol == XT_SERVE_SRPT) {
        std::stable_sort(js.begin(), js.end(), [&](const auto &a, const auto &b) { return left(a.second) < left(b.second); });
    } else if (pol == XT_SERVE_HYBRID) {
        const uint64_t shortw = S.R.P / 2;
        std::stable_sort(js.begin(), js.end(), [&](const auto &a, const auto &b) {
            const bool sa = left(a.second) <= shortw, sb = left(b.second) <= shortw;
            if (sa != sb) return sa;
            if (sa) return left(a.second) < left(b.second);
            return a.first < b.first;
        });
    } else if (pol == XT_SERVE_FAIR && !js.empty()) {
        cap = std::max<uint32_t>(64, S.R.P / (uint32_t)js.size());
        std::rotate(js.begin(), js.begin() + (S.rr % js.size()), js.end());
        S.rr++;
    }
    prio.clear();
    for (const auto &e : js) prio.push_back((uint32_t)e.second->entry);
}

// One step of the engine: start admitted jobs, set priorities, run the
// configured rounds, then finish the jobs whose paths have all ended.
int regen_step(xt_dev *d, ServeState &S, std::vector<uint64_t> &finished) {
    int rc = regen_init(d, S);
    if (rc) return rc;
    RegenHost &R = S.R;
    RegenView &V = R.V;
    // start admitted jobs that have no table entry yet
    for (uint64_t t : S.active) {
        SJob &J = S.jobs[t];
        if (!J.regen || J.entry >= 0 || R.free_entries.empty()) continue;
        int lane = -1;
        rc = regen_lane(d, S, J, lane);
        if (rc) return rc;
        if (lane < 0) continue;  // every lane serves another key: wait
        J.entry = R.free_entries.back();
        R.free_entries.pop_back();
        RegenJob &h = R.h_job[J.entry];
        std::memset(&h, 0, sizeof h);
        h.jd = J.jd;
        h.total = (unsigned long long)J.jd.total * J.jd.NT;
        h.next = 0;
        h.acc = J.acc;
        h.sbuf = J.sbuf;
        h.lane = (uint32_t)lane;
        R.h_done[J.entry] = 0;
        XT_TRY(hipMemcpyAsync(V.jobs + J.entry, &h, sizeof h, hipMemcpyHostToDevice, d->stream));
        hipLaunchKernelGGL(k_regen_job_reset, dim3(1), dim3(64), 0, d->stream, V, (uint32_t)J.entry);
        XT_LAUNCHED();
        J.paths_done = 0;
    }
    std::vector<uint32_t> prio;
    uint32_t cap = 0;
    regen_priorities(S, prio, cap);
    if (prio.empty()) return XT_OK;
    R.h_prio[0] = (uint32_t)prio.size();
    R.h_prio[1] = cap;
    std::memcpy(R.h_prio + 2, prio.data(), prio.size() * 4);
    XT_TRY(hipMemcpyAsync(&V.ctl->n_prio, R.h_prio, 8, hipMemcpyHostToDevice, d->stream));
    XT_TRY(hipMemcpyAsync(V.ctl->prio, R.h_prio + 2, prio.size() * 4, hipMemcpyHostToDevice, d->stream));
    XT_TRY(hipEventRecord(S.p0, d->stream));
    const uint32_t pairs = std::max<uint32_t>(1, S.cfg.rounds_per_step / 2);
    if (S.cfg.graph == 2) {
        if (R.dirty || std::memcmp(&R.captured, &d->tuning, sizeof(xt_tuning)) != 0) {
            rc = regen_capture(d, S);
            if (rc) return rc;
        }
        for (uint32_t k = 0; k < pairs; ++k) XT_TRY(hipGraphLaunch(R.exec, d->stream));
    } else {
        for (uint32_t k = 0; k < pairs; ++k) {
            rc = regen_round(d, S, 0);
            if (!rc) rc = regen_round(d, S, 1);
            if (rc) return rc;
        }
    }
    hipLaunchKernelGGL(k_regen_done_sum, dim3(XT_REGEN_MAX_JOBS / 256), dim3(256), 0, d->stream, V);
    XT_LAUNCHED();
    XT_TRY(hipEventRecord(S.p1, d->stream));
    XT_TRY(hipMemcpyAsync(R.h_done, V.done, 8ull * XT_REGEN_MAX_JOBS, hipMemcpyDeviceToHost, d->stream));
    XT_TRY(hipMemcpyAsync(R.h_cnt, V.cnt, sizeof(Counters) * XT_REGEN_MAX_LANES, hipMemcpyDeviceToHost, d->stream));
    XT_TRY(hipStreamSynchronize(d->stream));
    float ms = 0.f;
    hipEventElapsedTime(&ms, S.p0, S.p1);
    S.info.ms_device += ms;
    for (int l = 0; l < XT_REGEN_MAX_LANES; ++l) {
        if (!R.lanes[l].used) continue;
        S.info.queries += R.h_cnt[l].queries - R.queries_seen[l];
        S.info.retests += R.h_cnt[l].retests - R.retests_seen[l];
        R.queries_seen[l] = R.h_cnt[l].queries;
        R.retests_seen[l] = R.h_cnt[l].retests;
    }
    bool any_done = false;
    for (uint64_t t : S.active) {
        SJob &J = S.jobs[t];
        if (!J.regen || J.entry < 0) continue;
        J.passes += 1;  // steps the job has taken part in
        const uint64_t total = (uint64_t)J.jd.total * J.jd.NT;
        const uint64_t done = R.h_done[J.entry];
        S.info.samples += (done - J.paths_done) * J.npix / J.jd.NT;
        J.paths_done = done;
        if (done < total) continue;
        if (J.sbuf)
            hipLaunchKernelGGL(k_regen_resolve_ordered, dim3(grid_for(J.npix, XT_BLOCK, 8192)), dim3(XT_BLOCK), 0,
                               d->stream, J.jd, (const float4 *)J.sbuf, J.out);
        else
            hipLaunchKernelGGL(k_regen_resolve, dim3(grid_for(J.npix, XT_BLOCK, 8192)), dim3(XT_BLOCK), 0, d->stream,
                               J.jd, (const long long *)J.acc, J.out);
        XT_LAUNCHED();
        rc = take_host_block(d, S, J.npix * sizeof(float4), J.host);
        if (rc) return rc;
        XT_TRY(hipMemcpyAsync(J.host.p, J.out, J.npix * sizeof(float4), hipMemcpyDeviceToHost, d->stream));
        XT_TRY(hipEventRecord(J.done_ev, d->stream));
        J.issued = J.jd.total;
        R.free_entries.push_back(J.entry);
        J.entry = -1;
        finished.push_back(t);
        any_done = true;
    }
    // completion events must have happened before they are timed
    if (any_done) XT_TRY(hipStreamSynchronize(d->stream));
    S.info.passes += 1;
    return XT_OK;
}


// The lane of J's (scene, integrator, accel) if it is set up, else -1.
int regen_find_lane(const RegenHost &R, const SJob &J) {
    for (int l = 0; l < XT_REGEN_MAX_LANES; ++l) {
        const RegenLaneH &L = R.lanes[l];
        if (L.used && L.scene == J.job.scene_id && L.integrator == J.jd.integrator && L.accel == J.job.accel) return l;
    }
    return -1;
}

// Device-fed serving: takes the lane counters a finished replay left in its
// snapshot slot.
void regen_take_snapshot(ServeState &S, int slot) {
    RegenHost &R = S.R;
    const Counters *c = R.h_snap + (size_t)slot * XT_REGEN_MAX_LANES;
    for (int l = 0; l < XT_REGEN_MAX_LANES; ++l) {
        if (!R.lanes[l].used || c[l].queries < R.queries_seen[l]) continue;
        S.info.queries += c[l].queries - R.queries_seen[l];
        S.info.retests += c[l].retests - R.retests_seen[l];
        R.queries_seen[l] = c[l].queries;
        R.retests_seen[l] = c[l].retests;
    }
}

// Device-fed serving: retires the oldest queued replay (waiting for it when
// `wait`); false when it has not ended.
int regen_retire_replay(xt_dev *d, ServeState &S, bool wait, bool &done) {
    RegenHost &R = S.R;
    done = false;
    if (R.inflight.empty()) return XT_OK;
    const auto f = R.inflight.front();
    if (wait) {
        XT_TRY(hipEventSynchronize(f.first));
    } else {
        const hipError_t e = hipEventQuery(f.first);
        if (e == hipErrorNotReady) return XT_OK;
        if (e != hipSuccess) XT_TRY(e);
    }
    regen_take_snapshot(S, f.second);
    for (uint64_t t : S.active) {
        SJob &J = S.jobs[t];
        if (J.regen && J.entry >= 0) J.passes += 1;  // replays the job has taken part in
    }
    R.free_ev.push_back(f.first);
    R.inflight.pop_front();
    S.info.passes += 1;
    done = true;
    return XT_OK;
}

int regen_drain_replays(xt_dev *d, ServeState &S) {
    bool done = true;
    while (!S.R.inflight.empty()) {
        const int rc = regen_retire_replay(d, S, true, done);
        if (rc) return rc;
    }
    return XT_OK;
}

// One step of device-fed serving. The host writes newly admitted jobs to the
// request ring, copies out the images the device posted to the completion
// ring, and keeps XT_REGEN_INFLIGHT replays queued; it waits only for the
// oldest replay when nothing else moved (so a drain loop does not spin), or
// for an image copy when nothing else runs.
int regen_step_device(xt_dev *d, ServeState &S, std::vector<uint64_t> &finished) {
    int rc = regen_init(d, S);
    if (rc) return rc;
    RegenHost &R = S.R;
    RegenView &V = R.V;
    if (R.policy_sent != (uint32_t)S.policy) {
        *R.h_policy = (uint32_t)S.policy;
        XT_TRY(hipMemcpyAsync(&V.ctl->policy, R.h_policy, 4, hipMemcpyHostToDevice, d->stream));
        R.policy_sent = (uint32_t)S.policy;
    }
    // requests: admitted jobs without a table entry
    bool wrote = false;
    for (uint64_t t : S.active) {
        SJob &J = S.jobs[t];
        if (!J.regen || J.entry >= 0 || J.copying || R.free_entries.empty()) continue;
        if (J.zeroing) {
            const hipError_t e = hipEventQuery(J.done_ev);
            if (e == hipErrorNotReady) continue;
            if (e != hipSuccess) XT_TRY(e);
            J.zeroing = false;
        }
        int lane = regen_find_lane(R, J);
        if (lane < 0) {
            // a new lane changes the captured rounds: let the queued replays end first
            rc = regen_drain_replays(d, S);
            if (!rc) rc = regen_lane(d, S, J, lane);
            if (rc) return rc;
            if (lane < 0) continue;  // every lane serves another key: wait
        }
        J.entry = R.free_entries.back();
        R.free_entries.pop_back();
        R.entry_ticket[J.entry] = t;
        RegenReq &q = R.ring->req[R.req_head % XT_REGEN_RING];
        std::memset(&q, 0, sizeof q);
        q.job.jd = J.jd;
        q.job.total = (unsigned long long)J.jd.total * J.jd.NT;
        q.job.next = 0;
        q.job.sbuf = J.sbuf;
        q.job.acc = J.acc;
        q.job.lane = (uint32_t)lane;
        q.job.state = 1u;
        q.job.out = J.out;
        q.job.seq = R.seq++;
        q.entry = (uint32_t)J.entry;
        ++R.req_head;
        J.paths_done = 0;
        wrote = true;
    }
    if (wrote) __atomic_store_n(&R.ring->req_head, R.req_head, __ATOMIC_RELEASE);
    // completions: the image goes to host memory behind the queued replays
    const uint32_t fh = __atomic_load_n(&R.ring->fin_head, __ATOMIC_ACQUIRE);
    if (fh - R.fin_tail > XT_REGEN_RING) {
        d->err = "device-fed serving: completion ring overrun";
        return XT_E_STATE;
    }
    for (; R.fin_tail != fh; ++R.fin_tail) {
        const uint32_t e = R.ring->fin[R.fin_tail % XT_REGEN_RING];
        if (e >= XT_REGEN_MAX_JOBS) continue;
        const uint64_t t = R.entry_ticket[e];
        SJob &J = S.jobs.at(t);
        rc = take_host_block(d, S, J.npix * sizeof(float4), J.host);
        if (rc) return rc;
        XT_TRY(hipMemcpyAsync(J.host.p, J.out, J.npix * sizeof(float4), hipMemcpyDeviceToHost, d->stream));
        XT_TRY(hipEventRecord(J.done_ev, d->stream));
        J.copying = true;
        J.entry = -1;
        J.issued = J.jd.total;
        J.paths_done = (uint64_t)J.jd.total * J.jd.NT;
        S.info.samples += (uint64_t)J.jd.total * J.npix;
        R.free_entries.push_back((int)e);
        R.resolving.push_back(t);
    }
    // keep the device fed while any job is on it
    bool on_device = false;
    for (uint64_t t : S.active) on_device = on_device || (S.jobs[t].regen && S.jobs[t].entry >= 0);
    while (on_device && R.inflight.size() < XT_REGEN_INFLIGHT) {
        if (S.cfg.graph == 2 && (R.dirty || std::memcmp(&R.captured, &d->tuning, sizeof(xt_tuning)) != 0)) {
            rc = regen_drain_replays(d, S);
            if (!rc) rc = regen_capture(d, S);
            if (rc) return rc;
        }
        if (S.cfg.graph == 2) {
            XT_TRY(hipGraphLaunch(R.exec, d->stream));
        } else {
            rc = regen_round(d, S, 0);
            if (!rc) rc = regen_round(d, S, 1);
            if (!rc) rc = regen_tail(d, S);
            if (rc) return rc;
        }
        const int slot = (int)(R.launches++ % XT_REGEN_INFLIGHT);
        XT_TRY(hipMemcpyAsync(R.h_snap + (size_t)slot * XT_REGEN_MAX_LANES, V.cnt,
                              sizeof(Counters) * XT_REGEN_MAX_LANES, hipMemcpyDeviceToHost, d->stream));
        hipEvent_t ev = nullptr;
        if (!R.free_ev.empty()) {
            ev = R.free_ev.back();
            R.free_ev.pop_back();
        } else {
            XT_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        }
        XT_TRY(hipEventRecord(ev, d->stream));
        R.inflight.push_back({ev, slot});
    }
    bool done = true;
    while (done && !R.inflight.empty()) {
        rc = regen_retire_replay(d, S, false, done);
        if (rc) return rc;
    }
    auto collect = [&]() -> int {
        for (size_t k = 0; k < R.resolving.size();) {
            const uint64_t t = R.resolving[k];
            SJob &J = S.jobs.at(t);
            const hipError_t e = hipEventQuery(J.done_ev);
            if (e == hipErrorNotReady) {
                ++k;
                continue;
            }
            if (e != hipSuccess) XT_TRY(e);
            J.copying = false;
            finished.push_back(t);
            R.resolving[k] = R.resolving.back();
            R.resolving.pop_back();
        }
        return XT_OK;
    };
    rc = collect();
    if (rc) return rc;
    if (finished.empty()) {
        if (R.inflight.size() >= XT_REGEN_INFLIGHT) {
            rc = regen_retire_replay(d, S, true, done);  // the device has the next replay queued
        } else if (!R.resolving.empty()) {
            XT_TRY(hipEventSynchronize(S.jobs.at(R.resolving.front()).done_ev));
            rc = collect();
        }
        if (rc) return rc;
    }
    return XT_OK;
}

void release_block(ServeState &S, SJob &J) {
    if (J.blk.p) S.free_dev.push_back(J.blk);
    J.blk = DevBlock();
    J.accf = J.out = nullptr;
    J.acc64 = nullptr;
    J.acc = nullptr;
    J.sbuf = nullptr;
}

}  // namespace

void xt_serve_forget_scene(xt_dev *d, uint32_t scene_id) {
    if (!d->serve) return;
    RegenHost &R = d->serve->R;
    for (int l = 0; l < XT_REGEN_MAX_LANES; ++l)
        if (R.lanes[l].used && R.lanes[l].scene == scene_id) regen_drop_lane(d, R, l);
}

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
    regen_shutdown(d, *S);
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
    if (fixed + (regen_eligible(S, jd) ? regen_job_bytes(S.cfg, npix, jd) : job_bytes(npix)) > d->budget) {
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
        stats->samples = J.regen && J.state != FINISHED ? J.paths_done * J.npix / J.jd.NT : (uint64_t)J.issued * J.npix;
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
    std::vector<uint64_t> finished;
    bool any_regen = false;
    for (uint64_t t : S.active) any_regen = any_regen || S.jobs[t].regen;
    if (any_regen) {
        rc = S.cfg.loop == 2 ? regen_step_device(d, S, finished) : regen_step(d, S, finished);
        if (rc) return rc;
    }
    std::vector<Segment> segs;
    plan_pass(d, S, segs);
    if (!segs.empty()) {
        rc = run_pass(d, S, segs, finished);
        if (rc) return rc;
    }
    if (!finished.empty()) {
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
    if (!d || policy < XT_SERVE_FAIR || policy > XT_SERVE_HYBRID) return XT_E_INVALID;
    int rc = serve_init(d);
    if (rc) return rc;
    d->serve->policy = policy;
    return XT_OK;
}

extern "C" int xt_serve_configure(xt_dev *d, const xt_serve_config *c) {
    if (!d || !c) return XT_E_INVALID;
    if (c->regen > 2 || c->graph > 2 || (c->rounds_per_step & 1u) ||
        c->rounds_per_step > 1024 || (c->paths && (c->paths < 4096 || c->paths > (1u << 28))) ||
        c->sample_buffer_mb > (1u << 20) || c->lane_streams > 2 || c->loop > 2) {
        d->err = "bad serve configuration";
        return XT_E_INVALID;
    }
    int rc = serve_init(d);
    if (rc) return rc;
    ServeState &S = *d->serve;
    bool regen_busy = false;
    for (uint64_t t : S.active) regen_busy = regen_busy || S.jobs[t].regen;
    if ((c->paths && c->paths != S.R.P && S.R.ready) || (c->regen && c->regen != S.cfg.regen) ||
        (c->loop && c->loop != S.cfg.loop)) {
        if (regen_busy || !S.waiting.empty()) {
            d->err = "the regeneration engine holds jobs";
            return XT_E_STATE;
        }
        regen_shutdown(d, S);
    }
    if (c->regen) S.cfg.regen = c->regen;
    if (c->graph) S.cfg.graph = c->graph;
    if (c->rounds_per_step) S.cfg.rounds_per_step = c->rounds_per_step;
    if (c->paths) S.cfg.paths = c->paths;
    if (c->sample_buffer_mb) S.cfg.sample_buffer_mb = c->sample_buffer_mb;
    if (c->loop) S.cfg.loop = c->loop;
    if (c->lane_streams && c->lane_streams != S.cfg.lane_streams) {
        S.cfg.lane_streams = c->lane_streams;
        S.R.dirty = true;  // the captured rounds follow the lane streams
    }
    return XT_OK;
}

extern "C" int xt_serve_get_config(xt_dev *d, xt_serve_config *c) {
    if (!d || !c) return XT_E_INVALID;
    int rc = serve_init(d);
    if (rc) return rc;
    *c = d->serve->cfg;
    if (!c->paths) c->paths = d->serve->R.ready ? d->serve->R.P : XT_REGEN_PATHS;
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
