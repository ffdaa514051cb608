// Kernels of the server's path-regeneration engine (ao and pathtracer jobs).
//
// A pool of P path slots is shared by every active job: the queues of all
// lanes together hold at most P live paths. Each query round is one fixed
// launch sequence, captured once as a hipGraph:
//
//   k_regen_plan   one wave: hands the free room (P minus the live paths) to
//                  jobs in priority order (a segment per job: its next slot
//                  indices, its lane and where its rays go in that lane's
//                  queue)
//   k_regen_gen    camera rays of the new segments, appended to each lane's
//                  queue behind the paths that are still running
//   per lane       k_trace / k_recheck / k_retest as for one job, then the
//                  shading step; a path that continues is appended to the
//                  lane's next queue, one that ends to the retire list
//   k_regen_retire each ended path's radiance stored in its job's sample
//                  buffer at (pixel, sample), or, for a job too large for
//                  one, added to its pixel in 64-bit fixed point (exact, so
//                  the order in which paths end does not matter); the job's
//                  done count raised
//
// A finished job's sample buffer is summed per pixel in sample order with the
// compensated FP64 sum of a one-job render (k_regen_resolve_ordered), so its
// image is bit for bit the xt_render image.
//
// A path's state (throughput, last factor, random key, query count, job and
// pixel) travels with its ray in the queue, so every kernel reads and writes
// it at consecutive queue positions. A lane is one (scene, integrator, accel)
// with its own queues and counters, as in a one-job render. Jobs are
// admitted, prioritised and finished by the host between graph launches; the
// device never waits on the host inside a round.
#pragma once
#include "xt_kernels.h"

#define XT_REGEN_MAX_JOBS 1024
#define XT_REGEN_PATHS (16u << 20)  // default path slots of the engine
#define XT_REGEN_MAX_LANES 8
#define XT_REGEN_FIX 4294967296.0  // 2^32: fixed-point scale of the radiance sums
#define XT_REGEN_SHARDS 64         // done counters per job, summed at the end of a step

// One job of the device table (written by the host when the job is admitted).
struct RegenJob {
    JobDev jd;
    unsigned long long total;  // path slots of the job: NT tile positions x samples per pixel
    unsigned long long next;   // first slot not yet handed out (k_regen_plan)
    float4 *sbuf;              // radiance per (pixel, sample): pixel * samples + sample; or null
    long long *acc;            // without sbuf: 3 fixed-point sums per pixel
    uint32_t lane;
    uint32_t state;            // device-fed serving: 0 free, 1 active, 2 finished (being resolved)
    float4 *out;               // device-fed serving: the job's resolved image
    unsigned long long seq;    // device-fed serving: admission order
};

// A job's share of one round: `count` slots from slot index `slot0`, new rays
// [first, first + count) of the round, written at lane queue position qpos.
// Slot index k of a job is sample k % total of tile position k / total (the
// samples of one pixel follow each other); slot0 is sample s0 of position q0.
struct RegenSeg {
    unsigned long long slot0;
    uint32_t job, lane, first, count, qpos, q0, s0, pad;
};

struct RegenCtl {
    uint32_t P;
    uint32_t n_prio;       // jobs in prio[], highest priority first
    uint32_t cap_per_job;  // 0: none; else the most slots a job takes per round
    uint32_t n_seg, n_new;
    uint32_t retire_n;
    uint32_t prio[XT_REGEN_MAX_JOBS];
    // device-fed serving
    uint32_t policy;    // XT_SERVE_*: how k_regen_intake orders the jobs
    uint32_t req_tail;  // requests taken from the ring
    uint32_t fin_head;  // completions posted to the ring
    uint32_t n_fin;     // jobs finished in this replay
    uint32_t fin[XT_REGEN_MAX_JOBS];
};

// Device-fed serving: the host admits a job (buffers, table entry, lane) and
// writes it to the request ring; the device takes new requests at the start
// of every round, orders the jobs itself, resolves finished jobs in the
// replay that finishes them and posts them to the completion ring. Both
// rings sit in fine-grained host memory; the heads are published with
// system-scope release stores and read with acquire loads. The host never
// waits on the device to start, order or finish a job.
#define XT_REGEN_RING XT_REGEN_MAX_JOBS
struct RegenReq {
    RegenJob job;
    uint32_t entry, pad[3];
};
struct RegenRing {
    uint32_t req_head;  // requests written (host)
    uint32_t pad0[31];
    uint32_t fin_head;  // completions written (device)
    uint32_t pad1[31];
    RegenReq req[XT_REGEN_RING];
    uint32_t fin[XT_REGEN_RING];
};

// Path state next to a lane queue's rays: s0 = (T rgb, random key),
// s1 = (F rgb, query word), info = (job-table entry, pixel or XT_NONE for a
// slot past the image edge, sample, 0).
struct RegenState {
    float4 *s0, *s1;
    uint4 *info;
};

// Everything one round reads and writes besides the lanes' queues.
struct RegenView {
    RegenJob *jobs;
    unsigned long long *done;  // paths finished, per job-table entry (k_regen_done_sum)
    unsigned long long *done_shard;  // the same in XT_REGEN_SHARDS parts: [shard * XT_REGEN_MAX_JOBS + entry]
    RegenCtl *ctl;
    RegenSeg *segs;
    float4 *ret;           // ended paths: radiance rgb, pixel bits
    uint2 *ret_js;         // ended paths: job-table entry, sample
    Counters *cnt;         // one counter block per lane
    const Queue *lq;       // lane queues: lq[qi * XT_REGEN_MAX_LANES + lane]
    const RegenState *ls;  // their path state, same indexing
};

__device__ __forceinline__ uint32_t wave_excl_sum(uint32_t v, uint32_t &total) {
    const uint32_t lane = lane_id();
    uint32_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o);
        if (lane >= (uint32_t)o) x += y;
    }
    total = __shfl(x, 63);
    return x - v;
}

// One wave. The room left by the live paths goes to the jobs of ctl->prio in
// order, each taking what it still needs (at most cap_per_job); each job with
// a share becomes a segment, and each segment's rays are placed behind the
// lane's current queue (queue qi).
static __global__ __launch_bounds__(64) void k_regen_plan(RegenView V, uint32_t qi) {
    RegenCtl *C = V.ctl;
    const uint32_t lane = lane_id();
    uint32_t lane_count = lane < XT_REGEN_MAX_LANES ? V.cnt[lane].n[qi] : 0u;
    uint32_t live;
    wave_excl_sum(lane_count, live);
    uint32_t room = live < C->P ? C->P - live : 0u;
    const uint32_t cap = C->cap_per_job ? C->cap_per_job : 0xffffffffu;
    uint32_t given = 0, nseg = 0;
    for (uint32_t c = 0; c < C->n_prio && room > 0; c += 64) {
        const uint32_t k = c + lane;
        uint32_t job = XT_NONE, want = 0;
        unsigned long long nx = 0;
        if (k < C->n_prio) {
            job = C->prio[k];
            const RegenJob &J = V.jobs[job];
            nx = J.next;
            const unsigned long long left = J.total - nx;
            want = (uint32_t)(left < (unsigned long long)cap ? left : cap);
        }
        uint32_t tot;
        const uint32_t pre = wave_excl_sum(want, tot);
        const uint32_t give = pre >= room ? 0u : (want < room - pre ? want : room - pre);
        uint32_t gtot;
        const uint32_t gpre = wave_excl_sum(give, gtot);
        const unsigned long long m = __ballot(give > 0);
        const uint32_t below = (uint32_t)__popcll(lane ? (m & ((~0ull) >> (64 - lane))) : 0ull);
        const uint32_t jl = give ? V.jobs[job].lane : XT_NONE;
        // queue position: the lane's count so far plus the earlier segments
        // of the same lane in this chunk
        uint32_t qpos = 0;
        for (uint32_t l = 0; l < XT_REGEN_MAX_LANES; ++l) {
            uint32_t ltot;
            const uint32_t lpre = wave_excl_sum(jl == l ? give : 0u, ltot);
            const uint32_t base = __shfl(lane_count, (int)l);
            if (jl == l) qpos = base + lpre;
            if (lane == l) lane_count += ltot;
        }
        if (give) {
            RegenSeg &g = V.segs[nseg + below];
            const uint32_t T = V.jobs[job].jd.total;
            g.slot0 = nx;
            g.job = job;
            g.lane = jl;
            g.first = given + gpre;
            g.count = give;
            g.qpos = qpos;
            g.q0 = (uint32_t)(nx / T);
            g.s0 = (uint32_t)(nx - (unsigned long long)g.q0 * T);
            V.jobs[job].next = nx + give;
        }
        nseg += (uint32_t)__popcll(m);
        given += gtot;
        room -= gtot;
    }
    if (lane < XT_REGEN_MAX_LANES) V.cnt[lane].n[qi] = lane_count;
    if (lane == 0) {
        C->n_seg = nseg;
        C->n_new = given;
    }
}

// New camera rays: ray k of the round belongs to the segment whose range holds
// k (the segments' first entries are searched in LDS).
static __global__ __launch_bounds__(XT_BLOCK) void k_regen_gen(RegenView V, uint32_t qi) {
    __shared__ uint32_t sfirst[XT_REGEN_MAX_JOBS];
    const RegenCtl *C = V.ctl;
    const uint32_t n = C->n_new, nseg = C->n_seg;
    if ((uint32_t)blockIdx.x * blockDim.x >= n) return;
    for (uint32_t k = threadIdx.x; k < nseg; k += blockDim.x) sfirst[k] = V.segs[k].first;
    __syncthreads();
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
        uint32_t lo = 0, hi = nseg;  // last segment with first <= k
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (sfirst[mid] <= k) lo = mid;
            else hi = mid;
        }
        const RegenSeg g = V.segs[lo];
        const JobDev &j = V.jobs[g.job].jd;
        const uint32_t u = g.s0 + (k - g.first);
        const uint32_t q = g.q0 + u / j.total, s = u % j.total;
        uint32_t x = 0, y = 0;
        const bool live = slot_pixel(j, q, x, y);
        const Queue out = V.lq[qi * XT_REGEN_MAX_LANES + g.lane];
        const RegenState st = V.ls[qi * XT_REGEN_MAX_LANES + g.lane];
        const uint32_t at = g.qpos + (k - g.first);
        st.info[at] = make_uint4(g.job, live ? y * j.W + x : XT_NONE, s, 0u);
        st.s1[at] = make_float4(1.f, 1.f, 1.f, __uint_as_float(1u | (j.rdepth << 16)));
        if (!live) {
            out.r0[at] = make_float4(0.f, 0.f, 0.f, __uint_as_float(XT_DEAD));
            out.r1[at] = make_float4(0.f, 0.f, 1.f, 0.f);
            st.s0[at] = make_float4(1.f, 1.f, 1.f, 0.f);
            continue;
        }
        const uint32_t key = xt_path_key(j.seed, j.job_key, y * j.W + x, s);
        float cx, cy;
        aa_coords(j, x, y, s / j.spp, cx, cy);
        double o[3], d[3];
        camera_ray(j, cx, cy, key, o, d);
        st.s0[at] = make_float4(1.f, 1.f, 1.f, __uint_as_float(key));
        out.r0[at] = make_float4((float)o[0], (float)o[1], (float)o[2], __uint_as_float(XT_NONE));
        out.r1[at] = make_float4((float)d[0], (float)d[1], (float)d[2], 0.f);
    }
}

// Clears the round's retire count (a kernel rather than a memset: the round
// is replayed from a graph).
static __global__ void k_regen_round_start(RegenCtl *C) { C->retire_n = 0; }

// Appends the block's continuing paths (ray and state) to `out` and its ended
// paths to the retire list; every thread of the block calls it.
__device__ __forceinline__ void regen_emit(const RegenView &V, uint32_t *lds, bool have, bool cont, Queue &out,
                                           const RegenState &ost, uint32_t *out_count, f3 no, f3 nd, uint32_t nprim,
                                           float nself, float4 s0, float4 s1, uint4 info, f3 L) {
    const uint32_t slot = block_append<XT_BLOCK>(cont, out_count, lds);
    if (cont) {
        out.r0[slot] = make_float4(no.x, no.y, no.z, __uint_as_float(nprim));
        out.r1[slot] = make_float4(nd.x, nd.y, nd.z, nself);
        ost.s0[slot] = s0;
        ost.s1[slot] = s1;
        ost.info[slot] = info;
    }
    const bool ended = have && !cont;
    const uint32_t rs = block_append<XT_BLOCK>(ended, &V.ctl->retire_n, lds);
    if (ended) {
        V.ret[rs] = make_float4(L.x, L.y, L.z, __uint_as_float(info.y));
        V.ret_js[rs] = make_uint2(info.x, info.z);
    }
}

// Pathtracer shading of a lane (pt_step with the queue-resident state).
template <bool SAMPLED>
static __global__ __launch_bounds__(XT_BLOCK) void k_regen_shade_pt(const SceneView S, const RegenView V,
                                                             const Queue in, const RegenState ist,
                                                             const uint32_t *in_count, const float4 *hit, Queue out,
                                                             const RegenState ost, uint32_t *out_count,
                                                             const float4 *chan) {
    __shared__ uint32_t lds[XT_BLOCK / 64 + 1];
    const uint32_t n = *in_count;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {
        const uint32_t i = base + threadIdx.x;
        const bool have = i < n;
        bool cont = false;
        f3 no = mk(0, 0, 0), nd = mk(0, 0, 1), L = mk(0, 0, 0);
        uint32_t nprim = XT_NONE;
        float nself = 0.f;
        float4 s0 = make_float4(0.f, 0.f, 0.f, 0.f), s1 = s0;
        uint4 info = make_uint4(XT_NONE, XT_NONE, 0u, 0u);
        if (have) {
            s0 = ist.s0[i];
            s1 = ist.s1[i];
            info = ist.info[i];
            pt_step<SAMPLED>(S, hit[i], in.r0[i], in.r1[i], s0, s1, SAMPLED ? chan + 3 * (size_t)i : nullptr, L, cont,
                             no, nd, nprim, nself);
        }
        regen_emit(V, lds, have, cont, out, ost, out_count, no, nd, nprim, nself, s0, s1, info, L);
    }
}

// AO for lanes whose queue mixes first and second queries (the query index is
// in the path state): src/xtcore/integrator/ao/integrator.cc:12-53 as in
// k_shade_ao1 / k_shade_ao2.
static __global__ __launch_bounds__(XT_BLOCK) void k_regen_shade_ao(const SceneView S, const RegenView V,
                                                             const Queue in, const RegenState ist,
                                                             const uint32_t *in_count, const float4 *hit, Queue out,
                                                             const RegenState ost, uint32_t *out_count) {
    __shared__ uint32_t lds[XT_BLOCK / 64 + 1];
    const uint32_t n = *in_count;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {
        const uint32_t i = base + threadIdx.x;
        const bool have = i < n;
        bool cont = false;
        f3 no = mk(0, 0, 0), nd = mk(0, 0, 1), L = mk(0, 0, 0);
        uint32_t nprim = XT_NONE;
        float nself = 0.f;
        float4 s0 = make_float4(0.f, 0.f, 0.f, 0.f), s1 = s0;
        uint4 info = make_uint4(XT_NONE, XT_NONE, 0u, 0u);
        if (have) {
            s0 = ist.s0[i];
            s1 = ist.s1[i];
            info = ist.info[i];
            const float4 h = hit[i];
            const uint32_t prim = __float_as_uint(h.y);
            const uint32_t qword = __float_as_uint(s1.w);
            const float4 b = in.r1[i];
            if ((qword & 0xffffu) >= 2u) {  // the AO ray
                float w = 1.f;
                if (prim != XT_NONE) {
                    const float dist = h.x * sqrtf(b.x * b.x + b.y * b.y + b.z * b.z);
                    w = fminf(dist, 100.0f) / 100.0f;
                }
                L = mk(w, w, w);
            } else if (prim == XT_NONE) {
                L = mk(1.f, 1.f, 1.f);
            } else {
                const float4 a = in.r0[i];
                const f3 o = mk(a.x, a.y, a.z), d = mk(b.x, b.y, b.z);
                f3 p, N, g;
                uint32_t obj;
                surface(S, prim, o, d, h.x, h.z, h.w, p, N, g, obj);
                Draws r{xt_stream(__float_as_uint(s0.w), 1), 0};
                const f3 dir = smp_hemisphere(N, N, r);
                s1.w = __uint_as_float(qword + 1);
                cont = true;
                no = p;
                nd = dir;
                nprim = prim;
                nself = self_hit_t(g, N, dir);
            }
        }
        regen_emit(V, lds, have, cont, out, ost, out_count, no, nd, nprim, nself, s0, s1, info, L);
    }
}

__device__ __forceinline__ long long to_fixed(float v) {
    const double x = fmin(fmax((double)v, 0.0), 1e6);
    return (long long)__double2ll_rn(x * XT_REGEN_FIX);
}

// Ended paths: radiance into the job's sample buffer (or its pixel's
// fixed-point sums) and the job's done count. For fixed-point jobs,
// consecutive entries of a wave for the same (job, pixel) (the samples of a
// pixel sit next to each other in the queues) are summed first, so each run
// costs one set of atomics. Lanes of a wave that share the first lane's job
// count with one atomic.
static __global__ __launch_bounds__(XT_BLOCK) void k_regen_retire(RegenView V) {
    const uint32_t n = V.ctl->retire_n;
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t lane = lane_id();
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {
        const uint32_t i = base + threadIdx.x;
        const bool have = i < n;
        uint32_t job = XT_NONE, pixel = XT_NONE;
        bool fixed = false;
        long long f[3] = {0, 0, 0};
        if (have) {
            const float4 r = V.ret[i];
            const uint2 js = V.ret_js[i];
            job = js.x;
            pixel = __float_as_uint(r.w);
            if (pixel != XT_NONE) {
                const RegenJob &J = V.jobs[job];
                float4 *sb = J.sbuf;
                if (sb) {
                    sb[(size_t)pixel * J.jd.total + js.y] = make_float4(r.x, r.y, r.z, 0.f);
                } else {
                    fixed = true;
                    f[0] = to_fixed(r.x);
                    f[1] = to_fixed(r.y);
                    f[2] = to_fixed(r.z);
                }
            }
        }
        // segmented inclusive sum over runs of equal (job, pixel)
        const unsigned long long key = fixed ? ((unsigned long long)job << 32 | pixel) : ~0ull - lane;
        const unsigned long long prev = __shfl_up(key, 1), next = __shfl_down(key, 1);
        bool head = lane == 0 || prev != key;
        const bool tail = lane == 63 || next != key;
        if (__ballot(fixed)) {
            for (int o = 1; o < 64; o <<= 1) {
                long long y[3];
                for (int c = 0; c < 3; ++c) y[c] = __shfl_up(f[c], o);
                const bool hy = __shfl_up(head, o);
                if (lane >= (uint32_t)o && !head) {
                    for (int c = 0; c < 3; ++c) f[c] += y[c];
                    head = hy;
                }
            }
        }
        if (fixed && tail) {
            long long *a = V.jobs[job].acc + 3 * (size_t)pixel;
            for (int c = 0; c < 3; ++c)
                if (f[c]) atomicAdd((unsigned long long *)(a + c), (unsigned long long)f[c]);
        }
        // done counts: one atomic per run of equal jobs in the wave, on one of
        // XT_REGEN_SHARDS counters of the job so that the waves of the GPU do
        // not queue on one address
        const uint32_t jkey = have ? job : XT_NONE;
        const uint32_t jprev = __shfl_up(jkey, 1), jnext = __shfl_down(jkey, 1);
        const bool jhead = lane == 0 || jprev != jkey, jtail = lane == 63 || jnext != jkey;
        const unsigned long long heads = __ballot(jhead);
        // run length = lane - (last head at or below this lane) + 1
        const unsigned long long upto = heads & (lane == 63 ? ~0ull : ((2ull << lane) - 1));
        const uint32_t start = 63 - (uint32_t)__clzll((long long)upto);
        if (have && jtail) {
            const uint32_t shard = (blockIdx.x * (XT_BLOCK / 64) + (threadIdx.x >> 6)) % XT_REGEN_SHARDS;
            atomicAdd(&V.done_shard[shard * XT_REGEN_MAX_JOBS + job], (unsigned long long)(lane - start + 1));
        }
    }
}

// done[e] = the sum of entry e's shards.
static __global__ void k_regen_done_sum(RegenView V) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= XT_REGEN_MAX_JOBS) return;
    unsigned long long t = 0;
    for (uint32_t sh = 0; sh < XT_REGEN_SHARDS; ++sh) t += V.done_shard[sh * XT_REGEN_MAX_JOBS + e];
    V.done[e] = t;
}

// Zeroes the done shards of job-table entry e (a new job).
static __global__ void k_regen_job_reset(RegenView V, uint32_t e) {
    const uint32_t sh = threadIdx.x;
    if (sh < XT_REGEN_SHARDS) V.done_shard[sh * XT_REGEN_MAX_JOBS + e] = 0;
    if (sh == 0) V.done[e] = 0;
}

// Resolve of a finished job from its sample buffer: per pixel, the samples
// in order with k_accumulate's compensated FP64 sum and k_resolve's rounding,
// so the image is the one-job render's bit for bit.
__device__ __forceinline__ float4 regen_resolve_ordered_px(const JobDev &j, const float4 *sbuf, uint32_t p) {
    double2 acc[3] = {make_double2(0.0, 0.0), make_double2(0.0, 0.0), make_double2(0.0, 0.0)};
    const float4 *sp = sbuf + (size_t)p * j.total;
    for (uint32_t s = 0; s < j.total; ++s) {
        const float4 v = sp[s];
        neumaier_add(acc[0], weighted(v.x, j.weight));
        neumaier_add(acc[1], weighted(v.y, j.weight));
        neumaier_add(acc[2], weighted(v.z, j.weight));
    }
    return make_float4(fmaxf((float)(acc[0].x + acc[0].y), 0.f), fmaxf((float)(acc[1].x + acc[1].y), 0.f),
                       fmaxf((float)(acc[2].x + acc[2].y), 0.f), 1.f);
}

__device__ __forceinline__ float4 regen_resolve_fixed_px(const JobDev &j, const long long *acc, uint32_t p) {
    float c[3];
    for (int k = 0; k < 3; ++k) c[k] = fmaxf((float)((double)acc[3 * (size_t)p + k] / XT_REGEN_FIX * j.weight), 0.f);
    return make_float4(c[0], c[1], c[2], 1.f);
}

static __global__ __launch_bounds__(XT_BLOCK) void k_regen_resolve_ordered(const JobDev j, const float4 *sbuf,
                                                                    float4 *out) {
    const uint32_t n = j.W * j.H;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += stride) out[p] = regen_resolve_ordered_px(j, sbuf, p);
}

// Resolve of a finished job: fixed-point sums times the per-sample weight.
static __global__ __launch_bounds__(XT_BLOCK) void k_regen_resolve(const JobDev j, const long long *acc, float4 *out) {
    const uint32_t n = j.W * j.H;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += stride) out[p] = regen_resolve_fixed_px(j, acc, p);
}

// Device-fed serving, start of a round (one block of 256): new requests from
// the ring into the job table (their done counts zeroed), then every active
// job with slots left ordered by the policy into ctl->prio for k_regen_plan.
// SRPT: fewest slots left; FCFS and FAIR: admission order (FAIR also caps a
// job's slots per round at P / active jobs); HYBRID: jobs with at most P / 2
// slots left by that count, then the rest in admission order.
static __global__ __launch_bounds__(256) void k_regen_intake(RegenView V, const RegenRing *ring) {
    __shared__ unsigned long long key[XT_REGEN_MAX_JOBS];
    __shared__ uint32_t s_head, s_valid, s_active;
    RegenCtl *C = V.ctl;
    const uint32_t t = threadIdx.x;
    if (t == 0) {
        s_head = __hip_atomic_load(&ring->req_head, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
        s_valid = 0;
        s_active = 0;
    }
    __syncthreads();
    const uint32_t head = s_head, tail = C->req_tail;
    // at most a ring of requests can be pending (each holds a table entry)
    const uint32_t pending = head - tail <= XT_REGEN_RING ? head - tail : 0u;
    constexpr uint32_t WORDS = sizeof(RegenJob) / 4;
    for (uint32_t k = 0; k < pending; ++k) {
        const RegenReq *q = &ring->req[(tail + k) % XT_REGEN_RING];
        const uint32_t e = __hip_atomic_load(&q->entry, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        if (e >= XT_REGEN_MAX_JOBS) continue;  // not a table entry: skipped
        const uint32_t *src = (const uint32_t *)&q->job;
        uint32_t *dst = (uint32_t *)&V.jobs[e];
        for (uint32_t w = t; w < WORDS; w += 256) dst[w] = __hip_atomic_load(src + w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        if (t < XT_REGEN_SHARDS) V.done_shard[t * XT_REGEN_MAX_JOBS + e] = 0;
        if (t == 0) V.done[e] = 0;
    }
    __syncthreads();
    if (t == 0) C->req_tail = tail + pending;
    const uint32_t policy = C->policy;
    for (uint32_t e = t; e < XT_REGEN_MAX_JOBS; e += 256) {
        const RegenJob &J = V.jobs[e];
        unsigned long long k = ~0ull;
        if (J.state == 1u) atomicAdd(&s_active, 1u);
        if (J.state == 1u && J.next < J.total) {
            const unsigned long long left = J.total - J.next, seq = J.seq & 0x3fffull;
            unsigned long long primary = J.seq & 0xffffffffffull;  // FCFS, FAIR
            if (policy == XT_SERVE_SRPT) primary = left;
            else if (policy == XT_SERVE_HYBRID) primary = left <= C->P / 2 ? left : (1ull << 39) + (J.seq & 0x7fffffffffull);
            k = (primary << 24) | (seq << 10) | e;
            atomicAdd(&s_valid, 1u);
        }
        key[e] = k;
    }
    __syncthreads();
    // bitonic sort of the 1024 keys, ascending
    for (uint32_t size = 2; size <= XT_REGEN_MAX_JOBS; size <<= 1) {
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            for (uint32_t i = t; i < XT_REGEN_MAX_JOBS; i += 256) {
                const uint32_t j = i ^ stride;
                if (j > i) {
                    const unsigned long long a = key[i], b = key[j];
                    const bool up = (i & size) == 0;
                    if ((a > b) == up) {
                        key[i] = b;
                        key[j] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    const uint32_t nv = s_valid;
    for (uint32_t k = t; k < nv; k += 256) C->prio[k] = (uint32_t)(key[k] & 0x3ffull);
    if (t == 0) {
        C->n_prio = nv;
        const uint32_t na = s_active ? s_active : 1u;
        C->cap_per_job = policy == XT_SERVE_FAIR ? (C->P / na > 64u ? C->P / na : 64u) : 0u;
    }
}

// Device-fed serving, after the replay's rounds and k_regen_done_sum: the
// active jobs whose paths have all ended, listed for resolving.
static __global__ __launch_bounds__(256) void k_regen_finish(RegenView V) {
    RegenCtl *C = V.ctl;
    for (uint32_t e = threadIdx.x; e < XT_REGEN_MAX_JOBS; e += 256) {
        RegenJob &J = V.jobs[e];
        if (J.state == 1u && V.done[e] >= J.total) {
            J.state = 2u;
            C->fin[atomicAdd(&C->n_fin, 1u)] = e;
        }
    }
}

// Resolves every job that k_regen_finish listed into its image.
static __global__ __launch_bounds__(XT_BLOCK) void k_regen_resolve_fin(RegenView V) {
    const RegenCtl *C = V.ctl;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t f = 0; f < C->n_fin; ++f) {
        const RegenJob &J = V.jobs[C->fin[f]];
        const uint32_t n = J.jd.W * J.jd.H;
        for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += stride)
            J.out[p] = J.sbuf ? regen_resolve_ordered_px(J.jd, J.sbuf, p) : regen_resolve_fixed_px(J.jd, J.acc, p);
    }
}

// Posts the resolved jobs to the completion ring and frees their entries.
static __global__ void k_regen_post(RegenView V, RegenRing *ring) {
    RegenCtl *C = V.ctl;
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const uint32_t n = C->n_fin, h = C->fin_head;
    for (uint32_t f = 0; f < n; ++f) {
        const uint32_t e = C->fin[f];
        __hip_atomic_store(&ring->fin[(h + f) % XT_REGEN_RING], e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        V.jobs[e].state = 0u;
    }
    __hip_atomic_store(&ring->fin_head, h + n, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    C->fin_head = h + n;
    C->n_fin = 0;
}

static __global__ void k_regen_init(RegenCtl *C, uint32_t P) {
    C->P = P;
    C->retire_n = 0;
    C->n_seg = C->n_new = 0;
}
