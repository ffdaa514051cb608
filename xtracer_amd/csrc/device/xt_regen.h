// Kernels of the server's path-regeneration engine (ao and pathtracer jobs).
//
// A fixed pool of P path slots is shared by every active job. Each query
// round is one fixed launch sequence, captured once as a hipGraph:
//
//   k_regen_plan   one wave: hands the free slots to jobs in priority order
//                  (a segment per job: its next slot indices, its lane and
//                  where its rays go in that lane's queue)
//   k_regen_gen    camera rays of the new segments, appended to each lane's
//                  queue behind the paths that are still running
//   per lane       k_trace / k_recheck / k_retest / shading, as for one job;
//                  shading appends every path that ends to the retire list
//   k_regen_retire each ended path's radiance added to its job's pixel in
//                  64-bit fixed point (exact, so the order in which paths end
//                  does not matter), its job's done count raised and its slot
//                  returned to the free ring
//
// A lane is one (scene, integrator, accel) with its own queues and counters,
// as in a one-job render. Jobs are admitted, prioritised and finished by the
// host between graph launches; the device never waits on the host inside a
// round.
#pragma once
#include "xt_kernels.h"

#define XT_REGEN_MAX_JOBS 1024
#define XT_REGEN_MAX_LANES 8
#define XT_REGEN_FIX 4294967296.0  // 2^32: fixed-point scale of the radiance sums

// One job of the device table (written by the host when the job is admitted).
struct RegenJob {
    JobDev jd;
    unsigned long long total;  // path slots of the job: NT tile positions x samples per pixel
    unsigned long long next;   // first slot not yet handed out (k_regen_plan)
    long long *acc;            // 3 fixed-point sums per pixel
    uint32_t lane;
    uint32_t pad;
};

// A job's share of one round: `count` slots from slot index `slot0`, new rays
// [first, first + count) of the round, written at lane queue position qpos.
struct RegenSeg {
    unsigned long long slot0;
    uint32_t job, lane, first, count, qpos, pad;
};

struct RegenCtl {
    unsigned long long fl_head, fl_tail;  // free ring: slots [fl_head, fl_tail) mod P are free
    unsigned long long gen_head;          // fl_head when this round's plan ran
    uint32_t P;
    uint32_t n_prio;       // jobs in prio[], highest priority first
    uint32_t cap_per_job;  // 0: none; else the most slots a job takes per round
    uint32_t n_seg, n_new;
    uint32_t retire_n;
    uint32_t lane_in_use[XT_REGEN_MAX_LANES];
    uint32_t prio[XT_REGEN_MAX_JOBS];
};

// Everything one round reads and writes besides the lanes' queues.
struct RegenView {
    RegenJob *jobs;
    unsigned long long *done;  // paths finished, per job-table entry
    RegenCtl *ctl;
    RegenSeg *segs;
    float4 *ps0, *ps1, *rad;
    uint2 *pinfo;     // per slot: job-table entry, pixel (XT_NONE: a slot past the image edge)
    uint32_t *ring;   // free ring, P entries
    uint32_t *retire; // ended paths of the round
    Counters *cnt;    // one counter block per lane
    const Queue *lq;  // lane queues: lq[qi * XT_REGEN_MAX_LANES + lane]
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

// One wave. Free slots go to the jobs of ctl->prio in order, each taking what
// it still needs (at most cap_per_job); each job with a share becomes a
// segment, and each segment's rays are placed behind the lane's current queue
// (queue qi).
static __global__ __launch_bounds__(64) void k_regen_plan(RegenView V, uint32_t qi) {
    RegenCtl *C = V.ctl;
    const uint32_t lane = lane_id();
    const unsigned long long head = C->fl_head, tail = C->fl_tail;
    uint32_t room = (uint32_t)(tail - head);
    const uint32_t cap = C->cap_per_job ? C->cap_per_job : 0xffffffffu;
    uint32_t given = 0, nseg = 0;
    uint32_t lane_count = lane < XT_REGEN_MAX_LANES ? V.cnt[lane].n[qi] : 0u;
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
            g.slot0 = nx;
            g.job = job;
            g.lane = jl;
            g.first = given + gpre;
            g.count = give;
            g.qpos = qpos;
            V.jobs[job].next = nx + give;
        }
        nseg += (uint32_t)__popcll(m);
        given += gtot;
        room -= gtot;
    }
    if (lane < XT_REGEN_MAX_LANES) V.cnt[lane].n[qi] = lane_count;
    if (lane == 0) {
        C->gen_head = head;
        C->fl_head = head + given;
        C->n_seg = nseg;
        C->n_new = given;
    }
}

// New camera rays: ray k of the round belongs to the segment whose range holds
// k; slot index s of a job is sample s / NT at tile position s % NT, like a
// one-job pass, and takes path slot ring[(gen_head + k) mod P].
static __global__ __launch_bounds__(XT_BLOCK) void k_regen_gen(RegenView V, uint32_t qi) {
    const RegenCtl *C = V.ctl;
    const uint32_t n = C->n_new, nseg = C->n_seg, P = C->P;
    const unsigned long long head = C->gen_head;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
        uint32_t lo = 0, hi = nseg;  // last segment with first <= k
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (V.segs[mid].first <= k) lo = mid;
            else hi = mid;
        }
        const RegenSeg g = V.segs[lo];
        const uint32_t off = k - g.first;
        const JobDev &j = V.jobs[g.job].jd;
        const unsigned long long sidx = g.slot0 + off;
        const uint32_t s = (uint32_t)(sidx / j.NT), q = (uint32_t)(sidx - (unsigned long long)s * j.NT);
        const uint32_t path = V.ring[(uint32_t)((head + k) % P)];
        uint32_t x = 0, y = 0;
        const bool live = slot_pixel(j, q, x, y) && s < j.total;
        const Queue out = V.lq[qi * XT_REGEN_MAX_LANES + g.lane];
        const uint32_t at = g.qpos + off;
        V.rad[path] = make_float4(0.f, 0.f, 0.f, 0.f);
        V.pinfo[path] = make_uint2(g.job, live ? y * j.W + x : XT_NONE);
        out.path[at] = path;
        if (!live) {
            out.r0[at] = make_float4(0.f, 0.f, 0.f, __uint_as_float(XT_DEAD));
            out.r1[at] = make_float4(0.f, 0.f, 1.f, 0.f);
            continue;
        }
        const uint32_t key = xt_path_key(j.seed, j.job_key, y * j.W + x, s);
        float cx, cy;
        aa_coords(j, x, y, s / j.spp, cx, cy);
        double o[3], d[3];
        camera_ray(j, cx, cy, key, o, d);
        V.ps0[path] = make_float4(1.f, 1.f, 1.f, __uint_as_float(key));
        V.ps1[path] = make_float4(1.f, 1.f, 1.f, __uint_as_float(1u | (j.rdepth << 16)));
        out.r0[at] = make_float4((float)o[0], (float)o[1], (float)o[2], __uint_as_float(XT_NONE));
        out.r1[at] = make_float4((float)d[0], (float)d[1], (float)d[2], 0.f);
    }
}

// Clears the round's retire count (a kernel rather than a memset: the round
// is replayed from a graph).
static __global__ void k_regen_round_start(RegenCtl *C) { C->retire_n = 0; }

// AO for lanes whose queue mixes first and second queries (the query index is
// in the path state): src/xtcore/integrator/ao/integrator.cc:12-53 as in
// k_shade_ao1 / k_shade_ao2. Ended paths go to the retire list.
static __global__ __launch_bounds__(XT_BLOCK) void k_regen_shade_ao(const SceneView S, const Queue in,
                                                             const uint32_t *in_count, const float4 *hit, Queue out,
                                                             uint32_t *out_count, const float4 *ps0, float4 *ps1,
                                                             float4 *rad, uint32_t *retire, uint32_t *retire_n) {
    __shared__ uint32_t lds[XT_BLOCK / 64 + 1];
    const uint32_t n = *in_count;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {
        const uint32_t i = base + threadIdx.x;
        bool cont = false, ended = false;
        f3 no = mk(0, 0, 0), nd = mk(0, 0, 1);
        uint32_t nprim = XT_NONE, path = 0;
        float nself = 0.f;
        if (i < n) {
            path = in.path[i];
            const float4 h = hit[i];
            const uint32_t prim = __float_as_uint(h.y);
            const float4 s1 = ps1[path];
            const uint32_t qword = __float_as_uint(s1.w);
            const float4 b = in.r1[i];
            if ((qword & 0xffffu) >= 2u) {  // the AO ray
                float w = 1.f;
                if (prim != XT_NONE) {
                    const float dist = h.x * sqrtf(b.x * b.x + b.y * b.y + b.z * b.z);
                    w = fminf(dist, 100.0f) / 100.0f;
                }
                rad[path] = make_float4(w, w, w, 0.f);
                ended = true;
            } else if (prim == XT_NONE) {
                rad[path] = make_float4(1.f, 1.f, 1.f, 0.f);
                ended = true;
            } else {
                const float4 a = in.r0[i];
                const f3 o = mk(a.x, a.y, a.z), d = mk(b.x, b.y, b.z);
                f3 p, N, g;
                uint32_t obj;
                surface(S, prim, o, d, h.x, h.z, h.w, p, N, g, obj);
                Draws r{xt_stream(__float_as_uint(ps0[path].w), 1), 0};
                const f3 dir = smp_hemisphere(N, N, r);
                ps1[path] = make_float4(s1.x, s1.y, s1.z, __uint_as_float(qword + 1));
                cont = true;
                no = p;
                nd = dir;
                nprim = prim;
                nself = self_hit_t(g, N, dir);
            }
        }
        const uint32_t slot = block_append<XT_BLOCK>(cont, out_count, lds);
        if (cont) push_ray(out, slot, no, nd, nprim, nself, path);
        const uint32_t rs = block_append<XT_BLOCK>(ended, retire_n, lds);
        if (ended) retire[rs] = path;
    }
}

// 64-bit wave-aggregated append on a monotonic counter.
__device__ __forceinline__ unsigned long long wave_append64(bool want, unsigned long long *counter) {
    const unsigned long long m = __ballot(want);
    const uint32_t n = (uint32_t)__popcll(m);
    const uint32_t lane = lane_id();
    const uint32_t leader = m ? (uint32_t)__ffsll((long long)m) - 1 : 0;
    unsigned long long base = 0;
    if (lane == leader && n) base = atomicAdd(counter, (unsigned long long)n);
    base = __shfl(base, (int)leader);
    const unsigned long long below = lane ? (m & ((~0ull) >> (64 - lane))) : 0ull;
    return base + (unsigned long long)__popcll(below);
}

__device__ __forceinline__ long long to_fixed(float v) {
    const double x = fmin(fmax((double)v, 0.0), 1e6);
    return (long long)__double2ll_rn(x * XT_REGEN_FIX);
}

// Ended paths: radiance into the job's pixel, done count, slot back to the
// free ring. Lanes of a wave that share the first lane's job count with one
// atomic.
static __global__ __launch_bounds__(XT_BLOCK) void k_regen_retire(RegenView V) {
    RegenCtl *C = V.ctl;
    const uint32_t n = C->retire_n, P = C->P;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {
        const uint32_t i = base + threadIdx.x;
        const bool have = i < n;
        uint32_t path = 0, job = XT_NONE;
        if (have) {
            path = V.retire[i];
            const uint2 pi = V.pinfo[path];
            job = pi.x;
            if (pi.y != XT_NONE) {
                const float4 L = V.rad[path];
                long long *a = V.jobs[job].acc + 3 * (size_t)pi.y;
                atomicAdd((unsigned long long *)a, (unsigned long long)to_fixed(L.x));
                atomicAdd((unsigned long long *)(a + 1), (unsigned long long)to_fixed(L.y));
                atomicAdd((unsigned long long *)(a + 2), (unsigned long long)to_fixed(L.z));
            }
        }
        const unsigned long long act = __ballot(have);
        const uint32_t first = act ? (uint32_t)__ffsll((long long)act) - 1 : 0;
        const uint32_t j0 = __shfl(job, (int)first);
        const unsigned long long same = __ballot(have && job == j0);
        if (lane_id() == first && same) atomicAdd(&V.done[j0], (unsigned long long)__popcll(same));
        if (have && job != j0) atomicAdd(&V.done[job], 1ull);
        const unsigned long long at = wave_append64(have, &C->fl_tail);
        if (have) V.ring[(uint32_t)(at % P)] = path;
    }
}

// Resolve of a finished job: fixed-point sums times the per-sample weight.
static __global__ __launch_bounds__(XT_BLOCK) void k_regen_resolve(const JobDev j, const long long *acc, float4 *out) {
    const uint32_t n = j.W * j.H;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += stride) {
        float c[3];
        for (int k = 0; k < 3; ++k)
            c[k] = fmaxf((float)((double)acc[3 * (size_t)p + k] / XT_REGEN_FIX * j.weight), 0.f);
        out[p] = make_float4(c[0], c[1], c[2], 1.f);
    }
}

// Free ring with every slot free: ring[k] = k, head 0, tail P.
static __global__ void k_regen_init(RegenCtl *C, uint32_t *ring, uint32_t P) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < P; k += stride) ring[k] = k;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        C->fl_head = 0;
        C->fl_tail = P;
        C->P = P;
        C->retire_n = 0;
        C->n_seg = C->n_new = 0;
    }
}
