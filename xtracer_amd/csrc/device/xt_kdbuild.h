// Kernels of the device kd-tree build: the host build's binned SAH with
// triangle clipping (host/kdtree.cc), run breadth-first over the levels of
// the tree instead of depth-first over nodes.
//
// One level at a time, with every node of the level split in the same
// launches:
//
//   k_kdb_chunks / k_kdb_bin_big  nodes above XT_KDB_BIG references: their
//                                 start / end bin counts, a block per chunk
//                                 of references, summed in LDS, then added
//                                 to the node's bins
//   k_kdb_sah                     one wave per node: bin counts (in LDS for
//                                 the smaller nodes), the 3 x 63 candidate
//                                 planes' SAH costs, the first cheapest one,
//                                 or a leaf
//   k_kdb_classify                per reference: left, right or both, and
//                                 its bounds in each child (a straddling
//                                 triangle clipped to the child's box in FP64)
//   exclusive scans               of the (left, right) flags over the level's
//                                 references and of (inner, leaf references)
//                                 over its nodes
//   k_kdb_scatter / k_kdb_nodes   each child's references in their parent's
//                                 order, leaf references into a pool, the
//                                 level's nodes (children numbered level by
//                                 level) and the next level's nodes
//
// The finished tree is renumbered depth-first (k_kdb_sizes bottom-up,
// k_kdb_place top-down), which is the order in which the host build
// allocates child pairs and leaf references on one thread. Every decision
// uses the host build's float and double operations in the same order
// (contraction off), so the two trees are the same node for node.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../xt_layout.h"

#define XT_KDB_BINS 64       // SAH bins per axis: one per lane of a wave
#define XT_KDB_BIG 4096      // nodes with more references are binned by blocks
#define XT_KDB_CHUNK 2048    // references per binning block of such a node
#define XT_KDB_TILE 4096     // elements per scan block (256 threads x 16)

struct KdbParams {
    int max_depth, leaf_prims;
    float cost_traverse, cost_intersect, empty_bonus;
};

// A node of the level being split: its box and its references
// [first, first + count) in the level's reference arrays.
struct KdbNode {
    float lo[3];
    uint32_t first;
    float hi[3];
    uint32_t count;
};

struct KdbChunk {
    uint32_t node, big, begin, end;
};

// Per-level counters; error bits: 1 chunk list, 2 reference or pool
// capacity, 4 node capacity, 8 offset beyond the 30-bit node word, 16 an
// index out of its array (node, reference range, primitive, child).
struct KdbCtl {
    uint32_t n_big, n_chunks, error, pad;
};

#pragma clang fp contract(off)

// std::min / std::max as the host build calls them: the first argument
// unless the second is strictly smaller / larger.
__device__ __forceinline__ float kdb_min(float a, float b) { return b < a ? b : a; }
__device__ __forceinline__ float kdb_max(float a, float b) { return a < b ? b : a; }
__device__ __forceinline__ double kdb_mind(double a, double b) { return b < a ? b : a; }
__device__ __forceinline__ double kdb_maxd(double a, double b) { return a < b ? b : a; }

// (int) of a float as the host's cvttss2si gives it: INT_MIN outside the
// int range or for NaN.
__device__ __forceinline__ int kdb_to_int(float x) {
    return (x > -2147483648.f && x < 2147483648.f) ? (int)x : (int)0x80000000u;
}

__device__ __forceinline__ int kdb_bin(float v, float lo, float scale) {
    int s = kdb_to_int((v - lo) * scale);
    s = s < 0 ? 0 : s;
    return s > XT_KDB_BINS - 1 ? XT_KDB_BINS - 1 : s;
}

__device__ __forceinline__ float kdb_area(const float lo[3], const float hi[3]) {
    const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    return 2.f * (dx * dy + dy * dz + dz * dx);
}

__device__ __forceinline__ float kdb_round_down(double v) {
    float f = (float)v;
    if ((double)f > v) f = nextafterf(f, -INFINITY);
    return f;
}

__device__ __forceinline__ float kdb_round_up(double v) {
    float f = (float)v;
    if ((double)f < v) f = nextafterf(f, INFINITY);
    return f;
}

// Bounds of triangle t clipped to the box [blo, bhi] (Sutherland-Hodgman over
// the six planes in FP64, a vertex on a plane kept), widened outward by 1e-6
// of the box extent plus 1e-12 of the coordinates, rounded outward to float
// and cut to the box. False when nothing of the triangle is in the box.
__device__ bool kdb_clip(const XtPrim64 &t, const float blo[3], const float bhi[3], float olo[3], float ohi[3]) {
    double poly[2][12][3];
    for (int c = 0; c < 3; ++c) {
        poly[0][0][c] = t.a[c];
        poly[0][1][c] = t.b[c];
        poly[0][2][c] = t.c[c];
    }
    int n = 3, cur = 0;
    for (int a = 0; a < 3 && n > 0; ++a) {
        for (int side = 0; side < 2 && n > 0; ++side) {
            const double plane = side == 0 ? (double)blo[a] : (double)bhi[a];
            const double sgn = side == 0 ? 1.0 : -1.0;
            const int nxt = cur ^ 1;
            int m = 0;
            for (int k = 0; k < n && m < 11; ++k) {
                const int k1 = k + 1 == n ? 0 : k + 1;
                const double pa = poly[cur][k][a], qa = poly[cur][k1][a];
                const double dp = sgn * (pa - plane), dq = sgn * (qa - plane);
                if (dp >= 0) {
                    for (int c = 0; c < 3; ++c) poly[nxt][m][c] = poly[cur][k][c];
                    ++m;
                }
                if ((dp >= 0) != (dq >= 0) && m < 11) {
                    const double tt = dp / (dp - dq);
                    for (int c = 0; c < 3; ++c) {
                        const double p = poly[cur][k][c], q = poly[cur][k1][c];
                        poly[nxt][m][c] = p + tt * (q - p);
                    }
                    poly[nxt][m][a] = plane;
                    ++m;
                }
            }
            n = m;
            cur = nxt;
        }
    }
    if (n == 0) return false;
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int k = 0; k < n; ++k)
        for (int a = 0; a < 3; ++a) {
            lo[a] = kdb_mind(lo[a], poly[cur][k][a]);
            hi[a] = kdb_maxd(hi[a], poly[cur][k][a]);
        }
    for (int a = 0; a < 3; ++a) {
        const double ext = (double)bhi[a] - (double)blo[a];
        const double pad = 1e-6 * ext + 1e-12 * (fabs(lo[a]) + fabs(hi[a]));
        const float l = nextafterf((float)(lo[a] - pad), -INFINITY);
        const float h = nextafterf((float)(hi[a] + pad), INFINITY);
        olo[a] = kdb_max(l, blo[a]);
        ohi[a] = kdb_min(h, bhi[a]);
        if (olo[a] > ohi[a]) return false;
    }
    return true;
}

// Level-0 references: each soup primitive's box (its FP64 box rounded
// outward to float), prim id in r0.w, node 0 in r1.w.
static __global__ void k_kdb_boxes(const float4 *prims, const XtPrim64 *p64, uint32_t n, float4 *r0, float4 *r1) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const XtPrim64 q = p64[i];
        const uint32_t tf = __float_as_uint(prims[3 * (size_t)i].w);
        double lo[3], hi[3];
        if (XT_PRIM_TYPE(tf) == XT_PRIM_SPHERE) {
            const double r = fabs(q.b[0]);
            for (int a = 0; a < 3; ++a) {
                lo[a] = q.a[a] - r;
                hi[a] = q.a[a] + r;
            }
        } else {
            for (int a = 0; a < 3; ++a) {
                lo[a] = INFINITY;
                hi[a] = -INFINITY;
            }
            const double *v[3] = {q.a, q.b, q.c};
            for (int k = 0; k < 3; ++k)
                for (int a = 0; a < 3; ++a) {
                    if (v[k][a] < lo[a]) lo[a] = v[k][a];
                    if (v[k][a] > hi[a]) hi[a] = v[k][a];
                }
        }
        r0[i] = make_float4(kdb_round_down(lo[0]), kdb_round_down(lo[1]), kdb_round_down(lo[2]), __uint_as_float(i));
        r1[i] = make_float4(kdb_round_up(hi[0]), kdb_round_up(hi[1]), kdb_round_up(hi[2]), __uint_as_float(0u));
    }
}

// Root box, part 1: each thread folds a contiguous range of boxes in order,
// thread 0 folds its block's 256 results in order (min keeps the earlier of
// equal values, as the host's sequential fold does). part: 6 floats a block.
static __global__ __launch_bounds__(256) void k_kdb_root_partial(const float4 *r0, const float4 *r1, uint32_t n,
                                                                 float *part) {
    __shared__ float s[256][6];
    const uint32_t threads = gridDim.x * 256u, per = (n + threads - 1) / threads;
    const uint32_t g = blockIdx.x * 256u + threadIdx.x;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    const uint64_t b0 = (uint64_t)g * per, b1 = b0 + per < n ? b0 + per : n;
    for (uint64_t i = b0; i < b1; ++i) {
        const float4 a = r0[i], b = r1[i];
        lo[0] = kdb_min(lo[0], a.x);
        lo[1] = kdb_min(lo[1], a.y);
        lo[2] = kdb_min(lo[2], a.z);
        hi[0] = kdb_max(hi[0], b.x);
        hi[1] = kdb_max(hi[1], b.y);
        hi[2] = kdb_max(hi[2], b.z);
    }
    for (int a = 0; a < 3; ++a) {
        s[threadIdx.x][a] = lo[a];
        s[threadIdx.x][3 + a] = hi[a];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (uint32_t t = 1; t < 256; ++t)
            for (int a = 0; a < 3; ++a) {
                lo[a] = kdb_min(lo[a], s[t][a]);
                hi[a] = kdb_max(hi[a], s[t][3 + a]);
            }
        for (int a = 0; a < 3; ++a) {
            part[6 * blockIdx.x + a] = lo[a];
            part[6 * blockIdx.x + 3 + a] = hi[a];
        }
    }
}

// Root box, part 2 (one thread): the blocks' results in order; the root node
// holds all n references.
static __global__ void k_kdb_root_final(const float *part, uint32_t nb, uint32_t n, KdbNode *root, float *bounds) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t b = 0; b < nb; ++b)
        for (int a = 0; a < 3; ++a) {
            lo[a] = kdb_min(lo[a], part[6 * b + a]);
            hi[a] = kdb_max(hi[a], part[6 * b + 3 + a]);
        }
    KdbNode r;
    for (int a = 0; a < 3; ++a) {
        r.lo[a] = lo[a];
        r.hi[a] = hi[a];
        bounds[a] = lo[a];
        bounds[3 + a] = hi[a];
    }
    r.first = 0;
    r.count = n;
    *root = r;
}

__device__ __forceinline__ bool kdb_may_split(uint32_t count, int depth, const KdbParams &P) {
    return count > (uint32_t)P.leaf_prims && depth < P.max_depth;
}

// Chunks of the references of the level's big nodes (one binning block
// each); big_of[j] = the node's slot in the big-node bins, or -1.
static __global__ void k_kdb_chunks(const KdbNode *nodes, uint32_t n_nodes, uint32_t n_refs, int depth, KdbParams P,
                                    int32_t *big_of, KdbChunk *chunks, uint32_t chunk_cap, uint32_t big_cap,
                                    KdbCtl *ctl) {
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n_nodes; j += gridDim.x * blockDim.x) {
        const KdbNode nd = nodes[j];
        if ((uint64_t)nd.first + nd.count > n_refs) {
            big_of[j] = -1;
            atomicOr(&ctl->error, 16u);
            continue;
        }
        if (nd.count <= XT_KDB_BIG || !kdb_may_split(nd.count, depth, P)) {
            big_of[j] = -1;
            continue;
        }
        const uint32_t b = atomicAdd(&ctl->n_big, 1u);
        if (b >= big_cap) {
            big_of[j] = -1;
            atomicOr(&ctl->error, 1u);
            continue;
        }
        big_of[j] = (int32_t)b;
        const uint32_t nch = (nd.count + XT_KDB_CHUNK - 1) / XT_KDB_CHUNK;
        const uint32_t c0 = atomicAdd(&ctl->n_chunks, nch);
        if (c0 + nch > chunk_cap) {
            atomicOr(&ctl->error, 1u);
            continue;
        }
        for (uint32_t k = 0; k < nch; ++k) {
            KdbChunk c;
            c.node = j;
            c.big = b;
            c.begin = nd.first + k * XT_KDB_CHUNK;
            c.end = k + 1 == nch ? nd.first + nd.count : c.begin + XT_KDB_CHUNK;
            chunks[c0 + k] = c;
        }
    }
}

// Adds one chunk's bin counts to its node's bins (gbins: 6 x 64 counters per
// big node: axis-major, start counts then end counts).
static __global__ __launch_bounds__(256) void k_kdb_bin_big(const KdbNode *nodes, const KdbChunk *chunks,
                                                            const KdbCtl *ctl, const float4 *r0, const float4 *r1,
                                                            uint32_t *gbins) {
    __shared__ uint32_t bins[6 * XT_KDB_BINS];
    if (blockIdx.x >= ctl->n_chunks) return;
    const KdbChunk ch = chunks[blockIdx.x];
    const KdbNode nd = nodes[ch.node];
    for (uint32_t t = threadIdx.x; t < 6 * XT_KDB_BINS; t += 256) bins[t] = 0;
    __syncthreads();
    float scale[3];
    bool ok[3];
    for (int a = 0; a < 3; ++a) {
        const float ext = nd.hi[a] - nd.lo[a];
        ok[a] = ext > 0.f;
        scale[a] = ok[a] ? (float)XT_KDB_BINS / ext : 0.f;
    }
    for (uint32_t i = ch.begin + threadIdx.x; i < ch.end; i += 256) {
        const float4 a0 = r0[i], a1 = r1[i];
        const float rl[3] = {a0.x, a0.y, a0.z}, rh[3] = {a1.x, a1.y, a1.z};
        for (int a = 0; a < 3; ++a) {
            if (!ok[a]) continue;
            const float l = kdb_max(rl[a], nd.lo[a]), h = kdb_min(rh[a], nd.hi[a]);
            atomicAdd(&bins[(2 * a) * XT_KDB_BINS + kdb_bin(l, nd.lo[a], scale[a])], 1u);
            atomicAdd(&bins[(2 * a + 1) * XT_KDB_BINS + kdb_bin(h, nd.lo[a], scale[a])], 1u);
        }
    }
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < 6 * XT_KDB_BINS; t += 256)
        if (bins[t]) atomicAdd(&gbins[(size_t)ch.big * 6 * XT_KDB_BINS + t], bins[t]);
}

__device__ __forceinline__ uint32_t kdb_wave_excl(uint32_t v, uint32_t lane) {
    uint32_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= (uint32_t)o) x += y;
    }
    return x - v;
}

// One wave per node of the level: the split plane (axis, position), or a
// leaf (split[j].x = -1). The candidates are the 63 inner bin boundaries of
// each axis with an extent; the first (in axis, then boundary order) of the
// cheapest is taken when it beats making a leaf.
static __global__ __launch_bounds__(256) void k_kdb_sah(const KdbNode *nodes, uint32_t n_nodes, uint32_t n_refs,
                                                        int depth, KdbParams P,
                                                        const int32_t *big_of, const uint32_t *gbins, const float4 *r0,
                                                        const float4 *r1, int2 *split) {
    __shared__ uint32_t lb[4][6 * XT_KDB_BINS];
    const uint32_t w = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint32_t j = blockIdx.x * 4u + w;
    const bool active = j < n_nodes;
    KdbNode nd;
    bool leaf = true;
    float sa = 0.f;
    int32_t big = -1;
    if (active) {
        nd = nodes[j];
        sa = kdb_area(nd.lo, nd.hi);
        leaf = !kdb_may_split(nd.count, depth, P) || !(sa > 0.f);
        big = big_of[j];
    }
    float scale[3];
    bool ok[3];
    for (int a = 0; a < 3; ++a) {
        const float ext = active ? nd.hi[a] - nd.lo[a] : 0.f;
        ok[a] = ext > 0.f;
        scale[a] = ok[a] ? (float)XT_KDB_BINS / ext : 0.f;
    }
    const bool local = active && !leaf && big < 0 && nd.count <= XT_KDB_BIG && (uint64_t)nd.first + nd.count <= n_refs;
    for (uint32_t t = lane; t < 6 * XT_KDB_BINS; t += 64) lb[w][t] = 0;
    __syncthreads();
    if (local) {
        for (uint32_t i = nd.first + lane; i < nd.first + nd.count; i += 64) {
            const float4 a0 = r0[i], a1 = r1[i];
            const float rl[3] = {a0.x, a0.y, a0.z}, rh[3] = {a1.x, a1.y, a1.z};
            for (int a = 0; a < 3; ++a) {
                if (!ok[a]) continue;
                const float l = kdb_max(rl[a], nd.lo[a]), h = kdb_min(rh[a], nd.hi[a]);
                atomicAdd(&lb[w][(2 * a) * XT_KDB_BINS + kdb_bin(l, nd.lo[a], scale[a])], 1u);
                atomicAdd(&lb[w][(2 * a + 1) * XT_KDB_BINS + kdb_bin(h, nd.lo[a], scale[a])], 1u);
            }
        }
    }
    __syncthreads();
    if (!active) return;
    float best = P.cost_intersect * (float)nd.count;
    uint32_t best_idx = 0xffffffffu;
    if (!leaf) {
        const float ebf = 1.f - P.empty_bonus;
        for (int a = 0; a < 3; ++a) {
            const uint32_t st = big >= 0 ? gbins[(size_t)big * 6 * XT_KDB_BINS + (2 * a) * XT_KDB_BINS + lane]
                                         : lb[w][(2 * a) * XT_KDB_BINS + lane];
            const uint32_t en = big >= 0 ? gbins[(size_t)big * 6 * XT_KDB_BINS + (2 * a + 1) * XT_KDB_BINS + lane]
                                         : lb[w][(2 * a + 1) * XT_KDB_BINS + lane];
            const uint32_t nl = kdb_wave_excl(st, lane);
            const uint32_t nr = nd.count - kdb_wave_excl(en, lane);
            if (!ok[a] || lane == 0) continue;
            const float lo = nd.lo[a], ext = nd.hi[a] - lo;
            const float pos = lo + ext * (float)lane / (float)XT_KDB_BINS;
            if (!(pos > lo && pos < nd.hi[a])) continue;
            float llo[3], lhi[3], rlo[3], rhi[3];
            for (int c = 0; c < 3; ++c) {
                llo[c] = rlo[c] = nd.lo[c];
                lhi[c] = rhi[c] = nd.hi[c];
            }
            lhi[a] = pos;
            rlo[a] = pos;
            float c = P.cost_traverse +
                      P.cost_intersect * (kdb_area(llo, lhi) * (float)nl + kdb_area(rlo, rhi) * (float)nr) / sa;
            if (nl == 0 || nr == 0) c *= ebf;
            if (c < best) {  // axes in order: an equal cost on a later axis loses
                best = c;
                best_idx = (uint32_t)a * XT_KDB_BINS + lane;
            }
        }
    }
    // the cheapest over the wave; equal costs go to the lower (axis, boundary)
    for (int o = 32; o >= 1; o >>= 1) {
        const float ob = __shfl_xor(best, o, 64);
        const uint32_t oi = __shfl_xor(best_idx, o, 64);
        if (oi != 0xffffffffu && (best_idx == 0xffffffffu || ob < best || (ob == best && oi < best_idx))) {
            best = ob;
            best_idx = oi;
        }
    }
    if (lane == 0) {
        if (best_idx == 0xffffffffu) {
            split[j] = make_int2(-1, 0);
        } else {
            const int a = (int)(best_idx / XT_KDB_BINS);
            const uint32_t k = best_idx % XT_KDB_BINS;
            const float lo = nd.lo[a], ext = nd.hi[a] - lo;
            const float pos = lo + ext * (float)k / (float)XT_KDB_BINS;
            split[j] = make_int2(a, __float_as_int(pos));
        }
    }
}

// Per reference of an inner node: flags (left in the low word, right in the
// high word) and its bounds in each child (l0/l1 left, q0/q1 right).
static __global__ void k_kdb_classify(const float4 *r0, const float4 *r1, uint32_t n_refs, const KdbNode *nodes,
                                      uint32_t n_nodes, const int2 *split, const float4 *prims, const XtPrim64 *p64,
                                      uint32_t n_prims, uint64_t *flags, float4 *l0, float4 *l1, float4 *q0,
                                      float4 *q1, KdbCtl *ctl) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_refs; i += gridDim.x * blockDim.x) {
        const float4 a0 = r0[i], a1 = r1[i];
        const uint32_t j = __float_as_uint(a1.w), id = __float_as_uint(a0.w);
        if (j >= n_nodes || id >= n_prims) {
            flags[i] = 0;
            atomicOr(&ctl->error, 16u);
            continue;
        }
        const int2 s = split[j];
        if (s.x < 0) {
            flags[i] = 0;
            continue;
        }
        const int ax = s.x;
        const float sp = __int_as_float(s.y);
        float blo[3] = {a0.x, a0.y, a0.z}, bhi[3] = {a1.x, a1.y, a1.z};
        const float l = blo[ax], h = bhi[ax];
        const bool to_l = l < sp || (l <= sp && h <= sp);
        const bool to_r = h > sp || (l >= sp && h >= sp);
        float Llo[3], Lhi[3], Rlo[3], Rhi[3];
        bool fl, fr;
        const uint32_t tf = __float_as_uint(prims[3 * (size_t)id].w);
        if (to_l && to_r && l < sp && h > sp && XT_PRIM_TYPE(tf) == XT_PRIM_TRIANGLE) {
            const KdbNode nd = nodes[j];
            float clo[3], chi[3];
            for (int c = 0; c < 3; ++c) {
                clo[c] = nd.lo[c];
                chi[c] = nd.hi[c];
            }
            const XtPrim64 t = p64[id];
            chi[ax] = sp;
            fl = kdb_clip(t, clo, chi, Llo, Lhi);
            chi[ax] = nd.hi[ax];
            clo[ax] = sp;
            fr = kdb_clip(t, clo, chi, Rlo, Rhi);
        } else {
            fl = to_l;
            fr = to_r;
            for (int c = 0; c < 3; ++c) {
                Llo[c] = Rlo[c] = blo[c];
                Lhi[c] = Rhi[c] = bhi[c];
            }
            Lhi[ax] = kdb_min(bhi[ax], sp);
            Rlo[ax] = kdb_max(blo[ax], sp);
        }
        flags[i] = (uint64_t)(fl ? 1u : 0u) | ((uint64_t)(fr ? 1u : 0u) << 32);
        const float idf = __uint_as_float(id);
        if (fl) {
            l0[i] = make_float4(Llo[0], Llo[1], Llo[2], idf);
            l1[i] = make_float4(Lhi[0], Lhi[1], Lhi[2], 0.f);
        }
        if (fr) {
            q0[i] = make_float4(Rlo[0], Rlo[1], Rlo[2], idf);
            q1[i] = make_float4(Rhi[0], Rhi[1], Rhi[2], 0.f);
        }
    }
}

// Per node: 1 for an inner node (its ordinal among the level's inner nodes
// after the scan), its reference count in the high word for a leaf (its
// offset in the level's part of the leaf pool).
static __global__ void k_kdb_node_flags(const KdbNode *nodes, uint32_t n_nodes, const int2 *split, uint64_t *nf) {
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n_nodes; j += gridDim.x * blockDim.x)
        nf[j] = split[j].x >= 0 ? 1ull : ((uint64_t)nodes[j].count << 32);
}

__device__ __forceinline__ uint32_t kdb_lo32(uint64_t v) { return (uint32_t)v; }
__device__ __forceinline__ uint32_t kdb_hi32(uint64_t v) { return (uint32_t)(v >> 32); }

// References to the next level (left child's in their parent's order, then
// the right child's) or, for a leaf, into the leaf pool.
static __global__ void k_kdb_scatter(const float4 *r0, const float4 *r1, uint32_t n_refs, const KdbNode *nodes,
                                     uint32_t n_nodes, const int2 *split, const uint64_t *flags, const uint64_t *scan,
                                     const uint64_t *nscan, const float4 *l0, const float4 *l1, const float4 *q0,
                                     const float4 *q1, uint32_t pool_base, uint32_t *pool, uint32_t pool_cap, float4 *n0,
                                     float4 *n1, uint32_t next_cap, KdbCtl *ctl) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_refs; i += gridDim.x * blockDim.x) {
        const uint32_t j = __float_as_uint(r1[i].w);
        if (j >= n_nodes) {
            atomicOr(&ctl->error, 16u);
            continue;
        }
        const KdbNode nd = nodes[j];
        if ((uint64_t)nd.first + nd.count > n_refs || i < nd.first) {
            atomicOr(&ctl->error, 16u);
            continue;
        }
        if (split[j].x < 0) {
            const uint32_t pos = pool_base + kdb_hi32(nscan[j]) + (i - nd.first);
            if (pos < pool_cap) pool[pos] = __float_as_uint(r0[i].w);
            else atomicOr(&ctl->error, 2u);
            continue;
        }
        const uint64_t S = scan[i], Sf = scan[nd.first], Se = scan[nd.first + nd.count];
        const uint32_t base = kdb_lo32(Sf) + kdb_hi32(Sf), nl = kdb_lo32(Se) - kdb_lo32(Sf);
        const uint32_t ord = kdb_lo32(nscan[j]);
        const uint64_t f = flags[i];
        if (kdb_lo32(f)) {
            const uint32_t pos = base + (kdb_lo32(S) - kdb_lo32(Sf));
            if (pos < next_cap) {
                n0[pos] = l0[i];
                const float4 h = l1[i];
                n1[pos] = make_float4(h.x, h.y, h.z, __uint_as_float(2u * ord));
            } else {
                atomicOr(&ctl->error, 2u);
            }
        }
        if (kdb_hi32(f)) {
            const uint32_t pos = base + nl + (kdb_hi32(S) - kdb_hi32(Sf));
            if (pos < next_cap) {
                n0[pos] = q0[i];
                const float4 h = q1[i];
                n1[pos] = make_float4(h.x, h.y, h.z, __uint_as_float(2u * ord + 1u));
            } else {
                atomicOr(&ctl->error, 2u);
            }
        }
    }
}

// The level's nodes (global ids level_base + j; an inner node's children
// are the pair next_base + 2 * ordinal) and the next level's nodes.
static __global__ void k_kdb_nodes(const KdbNode *nodes, uint32_t n_nodes, uint32_t n_refs, const int2 *split,
                                   const uint64_t *scan,
                                   const uint64_t *nscan, uint32_t level_base, uint32_t next_base, uint32_t pool_base,
                                   XtKdNode *gnodes, uint32_t gcap, KdbNode *next, uint32_t next_cap, KdbCtl *ctl) {
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n_nodes; j += gridDim.x * blockDim.x) {
        const uint32_t g = level_base + j;
        if (g >= gcap) {
            atomicOr(&ctl->error, 4u);
            continue;
        }
        const KdbNode nd = nodes[j];
        if ((uint64_t)nd.first + nd.count > n_refs) {
            atomicOr(&ctl->error, 16u);
            continue;
        }
        const int2 s = split[j];
        XtKdNode out;
        if (s.x < 0) {
            const uint32_t off = pool_base + kdb_hi32(nscan[j]);
            if (off >= (1u << 30)) atomicOr(&ctl->error, 8u);
            out.word = (off << 2) | 3u;
            out.count = nd.count;
            gnodes[g] = out;
            continue;
        }
        const uint32_t ord = kdb_lo32(nscan[j]), pair = next_base + 2u * ord;
        if (pair + 1 >= (1u << 30)) atomicOr(&ctl->error, 8u);
        out.word = (pair << 2) | (uint32_t)s.x;
        out.split = __int_as_float(s.y);
        gnodes[g] = out;
        if (2u * ord + 1u >= next_cap) {
            atomicOr(&ctl->error, 4u);
            continue;
        }
        const uint64_t Sf = scan[nd.first], Se = scan[nd.first + nd.count];
        const uint32_t base = kdb_lo32(Sf) + kdb_hi32(Sf);
        const uint32_t nl = kdb_lo32(Se) - kdb_lo32(Sf), nr = kdb_hi32(Se) - kdb_hi32(Sf);
        KdbNode L = nd, R = nd;
        L.hi[s.x] = __int_as_float(s.y);
        R.lo[s.x] = __int_as_float(s.y);
        L.first = base;
        L.count = nl;
        R.first = base + nl;
        R.count = nr;
        next[2u * ord] = L;
        next[2u * ord + 1u] = R;
    }
}

// Exclusive scan of n 64-bit values in tiles of XT_KDB_TILE: in-tile
// prefixes and tile sums, the tile sums' prefixes (one block; the total goes
// to out[n]), then each tile's offset added.
static __global__ __launch_bounds__(256) void k_kdb_scan_tiles(const uint64_t *in, uint64_t *out, uint32_t n,
                                                               uint64_t *tile_sum) {
    __shared__ uint64_t s[256];
    const uint64_t t0 = (uint64_t)blockIdx.x * XT_KDB_TILE + (uint64_t)threadIdx.x * 16u;
    uint64_t v[16], sum = 0;
    for (int k = 0; k < 16; ++k) {
        v[k] = t0 + k < n ? in[t0 + k] : 0ull;
        sum += v[k];
    }
    s[threadIdx.x] = sum;
    __syncthreads();
    for (uint32_t o = 1; o < 256; o <<= 1) {
        const uint64_t y = threadIdx.x >= o ? s[threadIdx.x - o] : 0ull;
        __syncthreads();
        s[threadIdx.x] += y;
        __syncthreads();
    }
    uint64_t run = s[threadIdx.x] - sum;
    for (int k = 0; k < 16; ++k) {
        if (t0 + k < n) out[t0 + k] = run;
        run += v[k];
    }
    if (threadIdx.x == 255) tile_sum[blockIdx.x] = s[255];
}

static __global__ __launch_bounds__(256) void k_kdb_scan_sums(uint64_t *tile_sum, uint32_t nt, uint64_t *total) {
    __shared__ uint64_t s[256];
    const uint32_t per = (nt + 255) / 256, b0 = threadIdx.x * per;
    uint64_t sum = 0;
    for (uint32_t k = 0; k < per; ++k)
        if (b0 + k < nt) sum += tile_sum[b0 + k];
    s[threadIdx.x] = sum;
    __syncthreads();
    for (uint32_t o = 1; o < 256; o <<= 1) {
        const uint64_t y = threadIdx.x >= o ? s[threadIdx.x - o] : 0ull;
        __syncthreads();
        s[threadIdx.x] += y;
        __syncthreads();
    }
    uint64_t run = s[threadIdx.x] - sum;
    for (uint32_t k = 0; k < per; ++k)
        if (b0 + k < nt) {
            const uint64_t x = tile_sum[b0 + k];
            tile_sum[b0 + k] = run;
            run += x;
        }
    if (threadIdx.x == 255) *total = s[255];
}

static __global__ __launch_bounds__(256) void k_kdb_scan_add(uint64_t *out, uint32_t n, const uint64_t *tile_sum) {
    const uint64_t t0 = (uint64_t)blockIdx.x * XT_KDB_TILE;
    const uint64_t add = tile_sum[blockIdx.x];
    for (uint32_t k = threadIdx.x; k < XT_KDB_TILE; k += 256)
        if (t0 + k < n) out[t0 + k] += add;
}

// Depth-first renumbering, bottom-up: for the nodes [base, base + n) of one
// level, desc = nodes below the node, refc = leaf references below it.
static __global__ void k_kdb_sizes(const XtKdNode *gnodes, uint32_t base, uint32_t n, uint32_t total, uint32_t *desc,
                                   uint32_t *refc, KdbCtl *ctl) {
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const uint32_t g = base + j;
        const XtKdNode nd = gnodes[g];
        if ((nd.word & 3u) == 3u) {
            desc[g] = 0;
            refc[g] = nd.count;
        } else {
            const uint32_t c = nd.word >> 2;
            if (c <= g || c + 1 >= total) {
                atomicOr(&ctl->error, 16u);
                desc[g] = refc[g] = 0;
                continue;
            }
            desc[g] = 2u + desc[c] + desc[c + 1];
            refc[g] = refc[c] + refc[c + 1];
        }
    }
}

// Top-down: a node at pos with its child pair at pairpos and its leaf
// references from refpos writes itself into the depth-first arrays and
// places its children: the left child's pair right after its own, the
// right child's after the left child's subtree.
static __global__ void k_kdb_place(const XtKdNode *gnodes, uint32_t base, uint32_t n, const uint32_t *desc,
                                   const uint32_t *refc, uint32_t *pos, uint32_t *pairpos, uint32_t *refpos,
                                   const uint32_t *pool, uint32_t pool_n, uint32_t total, XtKdNode *out,
                                   uint32_t out_cap, uint32_t *out_refs, uint32_t refs_cap, KdbCtl *ctl) {
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const uint32_t g = base + j;
        const XtKdNode nd = gnodes[g];
        const uint32_t p = pos[g];
        if (p >= out_cap) {
            atomicOr(&ctl->error, 4u);
            continue;
        }
        XtKdNode o = nd;
        if ((nd.word & 3u) == 3u) {
            const uint32_t R = refpos[g], src = nd.word >> 2;
            o.word = (R << 2) | 3u;
            if ((uint64_t)R + nd.count > refs_cap || (uint64_t)src + nd.count > pool_n) {
                atomicOr(&ctl->error, 2u);
            } else {
                for (uint32_t k = 0; k < nd.count; ++k) out_refs[R + k] = pool[src + k];
            }
        } else {
            const uint32_t P = pairpos[g], c = nd.word >> 2;
            if (c <= g || c + 1 >= total) {
                atomicOr(&ctl->error, 16u);
                continue;
            }
            o.word = (P << 2) | (nd.word & 3u);
            pos[c] = P;
            pos[c + 1] = P + 1;
            pairpos[c] = P + 2;
            pairpos[c + 1] = P + 2 + desc[c];
            refpos[c] = refpos[g];
            refpos[c + 1] = refpos[g] + refc[c];
        }
        out[p] = o;
    }
}

// Each leaf reference's primitive record with the soup index as object word
// (the FP32 walk's leaf copies).
static __global__ void k_kdb_leaf_copies(const float4 *prims, uint32_t n_prims, const uint32_t *refs, uint32_t n,
                                         float4 *leaf, KdbCtl *ctl) {
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
        const size_t s = refs[r];
        if (s >= n_prims) {
            atomicOr(&ctl->error, 16u);
            continue;
        }
        leaf[3 * (size_t)r] = prims[3 * s];
        const float4 b = prims[3 * s + 1];
        leaf[3 * (size_t)r + 1] = make_float4(b.x, b.y, b.z, __uint_as_float((uint32_t)s));
        leaf[3 * (size_t)r + 2] = prims[3 * s + 2];
    }
}

#pragma clang fp contract(on)
