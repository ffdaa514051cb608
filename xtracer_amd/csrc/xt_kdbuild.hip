// Device kd-tree build for resident scenes: drives the level-by-level
// kernels of device/xt_kdbuild.h and installs the tree in the scene's
// SceneView. The tree is the host build's (host/kdtree.cc, one thread) node
// for node: same splits, same leaf references in the same order, same
// depth-first numbering.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "capi_internal.h"
#include "device/xt_kdbuild.h"
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

// A growable device array of one build, charged to the device's budget.
struct Buf {
    void *p = nullptr;
    size_t bytes = 0;
};

// Device memory of one build; everything still held is freed (after the
// stream has drained) when the build ends.
struct Work {
    xt_dev *d;
    std::vector<Buf *> held;
    explicit Work(xt_dev *dev) : d(dev) {}
    ~Work() {
        (void)hipStreamSynchronize(d->stream);
        for (Buf *b : held)
            if (b->p) {
                (void)hipFree(b->p);
                d->used -= std::min<uint64_t>(d->used, b->bytes);
                b->p = nullptr;
                b->bytes = 0;
            }
    }
    // At least `need` bytes in b; `keep` copies the old contents over.
    int reserve(Buf &b, size_t need, bool keep) {
        if (std::find(held.begin(), held.end(), &b) == held.end()) held.push_back(&b);
        if (need == 0) need = 256;
        if (need <= b.bytes) return XT_OK;
        size_t nb = std::max(need, b.bytes + b.bytes / 2);
        nb = (nb + 255) & ~(size_t)255;
        if (d->budget && d->used + nb > d->budget) {
            d->err = "device kd-tree build exceeds the HBM budget";
            return XT_E_BUDGET;
        }
        void *p = nullptr;
        XT_TRY(hipMalloc(&p, nb));
        d->used += nb;
        if (b.p) {
            if (keep) XT_TRY(hipMemcpyAsync(p, b.p, b.bytes, hipMemcpyDeviceToDevice, d->stream));
            XT_TRY(hipStreamSynchronize(d->stream));
            (void)hipFree(b.p);
            d->used -= std::min<uint64_t>(d->used, b.bytes);
        }
        b.p = p;
        b.bytes = nb;
        return XT_OK;
    }
};

template <class T>
T *as(const Buf &b) {
    return (T *)b.p;
}

int scan64(xt_dev *d, const uint64_t *in, uint64_t *out, uint32_t n, uint64_t *tiles) {
    const uint32_t nt = (uint32_t)(((uint64_t)n + XT_KDB_TILE - 1) / XT_KDB_TILE);
    if (nt) hipLaunchKernelGGL(k_kdb_scan_tiles, dim3(nt), dim3(256), 0, d->stream, in, out, n, tiles);
    hipLaunchKernelGGL(k_kdb_scan_sums, dim3(1), dim3(256), 0, d->stream, tiles, nt, out + n);
    if (nt) hipLaunchKernelGGL(k_kdb_scan_add, dim3(nt), dim3(256), 0, d->stream, out, n, (const uint64_t *)tiles);
    XT_LAUNCHED();
    return XT_OK;
}

struct Readback {
    uint64_t ref_total, node_total;
    KdbCtl ctl;
};

const char *kdb_error(uint32_t e) {
    if (e & 1u) return "device kd-tree build: chunk list overflow";
    if (e & 2u) return "device kd-tree build: reference capacity overflow";
    if (e & 4u) return "device kd-tree build: node capacity overflow";
    return "device kd-tree build: offsets beyond the node word";
}

// The result of one build: device arrays owned by the caller afterwards.
struct KdOut {
    XtKdNode *nodes = nullptr;
    uint32_t *refs = nullptr;
    float4 *leaf = nullptr;
    uint64_t nodes_bytes = 0, refs_bytes = 0, leaf_bytes = 0;
    float bounds[6] = {0, 0, 0, 0, 0, 0};
    xt_kd_build_stats stats{};
};

int out_alloc(xt_dev *d, void **p, uint64_t bytes) {
    if (bytes == 0) bytes = 256;
    if (d->budget && d->used + bytes > d->budget) {
        d->err = "device kd-tree exceeds the HBM budget";
        return XT_E_BUDGET;
    }
    XT_TRY(hipMalloc(p, bytes));
    d->used += bytes;
    return XT_OK;
}

void out_free(xt_dev *d, KdOut &o) {
    for (auto pr : {std::make_pair((void *)o.nodes, o.nodes_bytes), std::make_pair((void *)o.refs, o.refs_bytes),
                    std::make_pair((void *)o.leaf, o.leaf_bytes)})
        if (pr.first) {
            (void)hipFree(pr.first);
            d->used -= std::min<uint64_t>(d->used, pr.second);
        }
    o = KdOut();
}

int build(xt_dev *d, const SceneView &v, const xt_build_params *bp, KdOut &o) {
    const uint32_t n = v.n_prims;
    KdbParams P;
    xt::KdParams host;  // the host build's defaults
    P.leaf_prims = bp && bp->kd_leaf_prims > 0 ? bp->kd_leaf_prims : host.leaf_prims;
    P.cost_intersect = bp && bp->kd_cost_intersect > 0.f ? bp->kd_cost_intersect : host.cost_intersect;
    P.cost_traverse = host.cost_traverse;
    P.empty_bonus = host.empty_bonus;
    P.max_depth = bp && bp->kd_max_depth > 0 ? bp->kd_max_depth
                                            : (int)std::lround(8.0 + 1.3 * std::log2((double)std::max<uint32_t>(n, 1)));
    if (host.bins != XT_KDB_BINS) {
        d->err = "device kd-tree build: bin count differs from the host build's";
        return XT_E_STATE;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    XT_TRY(hipEventCreate(&e0));
    XT_TRY(hipEventCreate(&e1));
    struct Events {
        hipEvent_t a, b;
        ~Events() {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    } evs{e0, e1};
    XT_TRY(hipEventRecord(e0, d->stream));
    Readback *hr = nullptr;
    XT_TRY(hipHostMalloc((void **)&hr, sizeof(Readback), hipHostMallocDefault));
    struct Pinned {
        Readback *p;
        ~Pinned() { (void)hipHostFree(p); }
    } pinned{hr};

    if (n == 0) {
        // the host build's empty tree: a root leaf without references
        int rc = out_alloc(d, (void **)&o.nodes, 256);
        if (!rc) {
            o.nodes_bytes = 256;
            rc = out_alloc(d, (void **)&o.refs, 256);
        }
        if (!rc) {
            o.refs_bytes = 256;
            rc = out_alloc(d, (void **)&o.leaf, 256);
        }
        if (!rc) o.leaf_bytes = 256;
        if (rc) {
            out_free(d, o);
            return rc;
        }
        static const XtKdNode leaves[2] = {{3u, {0.f}}, {3u, {0.f}}};  // the root and the padding leaf
        XT_TRY(hipMemcpyAsync(o.nodes, leaves, sizeof leaves, hipMemcpyHostToDevice, d->stream));
        XT_TRY(hipEventRecord(e1, d->stream));
        XT_TRY(hipStreamSynchronize(d->stream));
        o.stats.n_nodes = 2;
        o.stats.leaves = 1;
        o.stats.levels = 1;
        float ms = 0.f;
        hipEventElapsedTime(&ms, e0, e1);
        o.stats.ms = ms;
        return XT_OK;
    }

    Work W(d);
    Buf r0[2], r1[2], nodes[2], split, big_of, nflags, nscan, flags, scan, l0, l1, q0, q1, tiles, gbins, chunks, ctl,
        part, bounds, gnodes, pool;
    int rc = XT_OK;
    auto need = [&](Buf &b, size_t bytes, bool keep = false) {
        if (!rc) rc = W.reserve(b, bytes, keep);
        return rc == XT_OK;
    };
    const unsigned nb_root = grid_for(n, 256u * 64u, 1024);
    need(r0[0], 16ull * n);
    need(r1[0], 16ull * n);
    need(nodes[0], sizeof(KdbNode));
    need(part, 24ull * nb_root);
    need(bounds, 24);
    need(ctl, sizeof(KdbCtl));
    if (rc) return rc;
    hipLaunchKernelGGL(k_kdb_boxes, dim3(grid_for(n, 256, 8192)), dim3(256), 0, d->stream, v.prims, v.prims64, n,
                       as<float4>(r0[0]), as<float4>(r1[0]));
    hipLaunchKernelGGL(k_kdb_root_partial, dim3(nb_root), dim3(256), 0, d->stream, (const float4 *)as<float4>(r0[0]),
                       (const float4 *)as<float4>(r1[0]), n, as<float>(part));
    hipLaunchKernelGGL(k_kdb_root_final, dim3(1), dim3(1), 0, d->stream, (const float *)as<float>(part), nb_root, n,
                       as<KdbNode>(nodes[0]), as<float>(bounds));
    XT_LAUNCHED();

    std::vector<std::pair<uint32_t, uint32_t>> levels;  // (first global id, nodes)
    uint32_t level_nodes = 1, level_refs = n, level_base = 0, pool_total = 0, depth = 0, leaves = 0;
    int cur = 0;
    for (int level = 0; level_nodes > 0; ++level) {
        if (level > P.max_depth + 1) {
            d->err = "device kd-tree build: deeper than its depth limit";
            return XT_E_STATE;
        }
        const uint32_t NL = level_nodes, NR = level_refs;
        const uint32_t big_cap = NR / XT_KDB_BIG + 1, chunk_cap = NR / XT_KDB_CHUNK + big_cap + 1;
        const uint32_t tiles_n = (uint32_t)((std::max(NR, NL) + (uint64_t)XT_KDB_TILE) / XT_KDB_TILE) + 1;
        need(gnodes, sizeof(XtKdNode) * ((size_t)level_base + NL), true);
        need(split, 8ull * NL);
        need(big_of, 4ull * NL);
        need(nflags, 8ull * NL);
        need(nscan, 8ull * (NL + 1));
        need(flags, 8ull * NR);
        need(scan, 8ull * (NR + 1));
        need(l0, 16ull * NR);
        need(l1, 16ull * NR);
        need(q0, 16ull * NR);
        need(q1, 16ull * NR);
        need(tiles, 8ull * tiles_n);
        need(gbins, 4ull * 6 * XT_KDB_BINS * big_cap);
        need(chunks, sizeof(KdbChunk) * (size_t)chunk_cap);
        if (rc) return rc;
        const uint32_t gcap = (uint32_t)std::min<size_t>(gnodes.bytes / sizeof(XtKdNode), 0xffffffffu);
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, d->stream, as<uint32_t>(ctl), 0u,
                           sizeof(KdbCtl) / 4);
        hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(6ull * XT_KDB_BINS * big_cap, 256, 4096)), dim3(256), 0,
                           d->stream, as<uint32_t>(gbins), 0u, 6ull * XT_KDB_BINS * big_cap);
        const KdbNode *nd = as<KdbNode>(nodes[cur]);
        const float4 *a0 = as<float4>(r0[cur]), *a1 = as<float4>(r1[cur]);
        hipLaunchKernelGGL(k_kdb_chunks, dim3(grid_for(NL, 256, 4096)), dim3(256), 0, d->stream, nd, NL, NR, level, P,
                           as<int32_t>(big_of), as<KdbChunk>(chunks), chunk_cap, big_cap, as<KdbCtl>(ctl));
        hipLaunchKernelGGL(k_kdb_bin_big, dim3(chunk_cap), dim3(256), 0, d->stream, nd,
                           (const KdbChunk *)as<KdbChunk>(chunks), (const KdbCtl *)as<KdbCtl>(ctl), a0, a1,
                           as<uint32_t>(gbins));
        hipLaunchKernelGGL(k_kdb_sah, dim3((NL + 3) / 4), dim3(256), 0, d->stream, nd, NL, NR, level, P,
                           (const int32_t *)as<int32_t>(big_of), (const uint32_t *)as<uint32_t>(gbins), a0, a1,
                           as<int2>(split));
        hipLaunchKernelGGL(k_kdb_classify, dim3(grid_for(NR, 256, 16384)), dim3(256), 0, d->stream, a0, a1, NR, nd, NL,
                           (const int2 *)as<int2>(split), v.prims, v.prims64, n, as<uint64_t>(flags), as<float4>(l0),
                           as<float4>(l1), as<float4>(q0), as<float4>(q1), as<KdbCtl>(ctl));
        XT_LAUNCHED();
        rc = scan64(d, as<uint64_t>(flags), as<uint64_t>(scan), NR, as<uint64_t>(tiles));
        if (rc) return rc;
        hipLaunchKernelGGL(k_kdb_node_flags, dim3(grid_for(NL, 256, 4096)), dim3(256), 0, d->stream, nd, NL,
                           (const int2 *)as<int2>(split), as<uint64_t>(nflags));
        XT_LAUNCHED();
        rc = scan64(d, as<uint64_t>(nflags), as<uint64_t>(nscan), NL, as<uint64_t>(tiles));
        if (rc) return rc;
        XT_TRY(hipMemcpyAsync(&hr->ref_total, as<uint64_t>(scan) + NR, 8, hipMemcpyDeviceToHost, d->stream));
        XT_TRY(hipMemcpyAsync(&hr->node_total, as<uint64_t>(nscan) + NL, 8, hipMemcpyDeviceToHost, d->stream));
        XT_TRY(hipMemcpyAsync(&hr->ctl, ctl.p, sizeof(KdbCtl), hipMemcpyDeviceToHost, d->stream));
        XT_TRY(hipStreamSynchronize(d->stream));
        if (hr->ctl.error) {
            d->err = kdb_error(hr->ctl.error);
            return XT_E_STATE;
        }
        const uint32_t n_inner = (uint32_t)hr->node_total, leaf_refs = (uint32_t)(hr->node_total >> 32);
        const uint64_t next_refs = (hr->ref_total & 0xffffffffull) + (hr->ref_total >> 32);
        if (next_refs >= 0xffffffffull || (uint64_t)pool_total + leaf_refs >= (1ull << 30)) {
            d->err = "device kd-tree build: more references than the node word holds";
            return XT_E_STATE;
        }
        const int nxt = cur ^ 1;
        need(pool, 4ull * (pool_total + leaf_refs), true);
        need(r0[nxt], 16ull * next_refs);
        need(r1[nxt], 16ull * next_refs);
        need(nodes[nxt], sizeof(KdbNode) * 2ull * n_inner);
        if (rc) return rc;
        const uint32_t pool_cap = (uint32_t)std::min<size_t>(pool.bytes / 4, 0xffffffffu);
        const uint32_t next_cap = (uint32_t)std::min<size_t>(r0[nxt].bytes / 16, 0xffffffffu);
        const uint32_t next_nodes_cap = (uint32_t)std::min<size_t>(nodes[nxt].bytes / sizeof(KdbNode), 0xffffffffu);
        hipLaunchKernelGGL(k_kdb_scatter, dim3(grid_for(NR, 256, 16384)), dim3(256), 0, d->stream, a0, a1, NR, nd, NL,
                           (const int2 *)as<int2>(split), (const uint64_t *)as<uint64_t>(flags),
                           (const uint64_t *)as<uint64_t>(scan), (const uint64_t *)as<uint64_t>(nscan),
                           (const float4 *)as<float4>(l0), (const float4 *)as<float4>(l1),
                           (const float4 *)as<float4>(q0), (const float4 *)as<float4>(q1), pool_total,
                           as<uint32_t>(pool), pool_cap, as<float4>(r0[nxt]), as<float4>(r1[nxt]), next_cap,
                           as<KdbCtl>(ctl));
        hipLaunchKernelGGL(k_kdb_nodes, dim3(grid_for(NL, 256, 4096)), dim3(256), 0, d->stream, nd, NL, NR,
                           (const int2 *)as<int2>(split), (const uint64_t *)as<uint64_t>(scan),
                           (const uint64_t *)as<uint64_t>(nscan), level_base, level_base + NL, pool_total,
                           as<XtKdNode>(gnodes), gcap, as<KdbNode>(nodes[nxt]), next_nodes_cap, as<KdbCtl>(ctl));
        XT_LAUNCHED();
        levels.push_back({level_base, NL});
        if (NL > n_inner) {
            depth = (uint32_t)level;
            leaves += NL - n_inner;
        }
        pool_total += leaf_refs;
        level_base += NL;
        level_nodes = 2 * n_inner;
        level_refs = (uint32_t)next_refs;
        cur = nxt;
    }
    // errors of the last level's scatter
    XT_TRY(hipMemcpyAsync(&hr->ctl, ctl.p, sizeof(KdbCtl), hipMemcpyDeviceToHost, d->stream));
    XT_TRY(hipStreamSynchronize(d->stream));
    if (hr->ctl.error) {
        d->err = kdb_error(hr->ctl.error);
        return XT_E_STATE;
    }

    // depth-first numbering
    const uint32_t total = level_base;
    Buf desc, refc, pos, pairpos, refpos;
    need(desc, 4ull * total);
    need(refc, 4ull * total);
    need(pos, 4ull * total);
    need(pairpos, 4ull * total);
    need(refpos, 4ull * total);
    if (rc) return rc;
    for (auto it = levels.rbegin(); it != levels.rend(); ++it)
        hipLaunchKernelGGL(k_kdb_sizes, dim3(grid_for(it->second, 256, 8192)), dim3(256), 0, d->stream,
                           (const XtKdNode *)as<XtKdNode>(gnodes), it->first, it->second, total, as<uint32_t>(desc),
                           as<uint32_t>(refc), as<KdbCtl>(ctl));
    XT_LAUNCHED();
    // the root at 0, an unreferenced empty leaf at 1 (child pairs on even
    // nodes, as the host build lays them out), the root's pair from 2
    static const uint32_t root_init[3] = {0u, 2u, 0u};  // pos, pair position, reference position of the root
    static const XtKdNode pad = {3u, {0.f}};
    XT_TRY(hipMemcpyAsync(pos.p, &root_init[0], 4, hipMemcpyHostToDevice, d->stream));
    XT_TRY(hipMemcpyAsync(pairpos.p, &root_init[1], 4, hipMemcpyHostToDevice, d->stream));
    XT_TRY(hipMemcpyAsync(refpos.p, &root_init[2], 4, hipMemcpyHostToDevice, d->stream));
    o.nodes_bytes = (sizeof(XtKdNode) * ((uint64_t)total + 1) + 255) & ~255ull;
    o.refs_bytes = (4ull * pool_total + 255) & ~255ull;
    o.leaf_bytes = (48ull * pool_total + 255) & ~255ull;
    rc = out_alloc(d, (void **)&o.nodes, o.nodes_bytes);
    if (!rc) rc = out_alloc(d, (void **)&o.refs, o.refs_bytes);
    if (!rc) rc = out_alloc(d, (void **)&o.leaf, o.leaf_bytes);
    if (rc) {
        (void)hipStreamSynchronize(d->stream);
        out_free(d, o);
        return rc;
    }
    XT_TRY(hipMemcpyAsync(o.nodes + 1, &pad, sizeof pad, hipMemcpyHostToDevice, d->stream));
    for (const auto &lv : levels)
        hipLaunchKernelGGL(k_kdb_place, dim3(grid_for(lv.second, 256, 8192)), dim3(256), 0, d->stream,
                           (const XtKdNode *)as<XtKdNode>(gnodes), lv.first, lv.second,
                           (const uint32_t *)as<uint32_t>(desc), (const uint32_t *)as<uint32_t>(refc),
                           as<uint32_t>(pos), as<uint32_t>(pairpos), as<uint32_t>(refpos),
                           (const uint32_t *)as<uint32_t>(pool), pool_total, total, o.nodes, total + 1, o.refs,
                           pool_total, as<KdbCtl>(ctl));
    if (pool_total)
        hipLaunchKernelGGL(k_kdb_leaf_copies, dim3(grid_for(pool_total, 256, 16384)), dim3(256), 0, d->stream,
                           v.prims, n, (const uint32_t *)o.refs, pool_total, o.leaf, as<KdbCtl>(ctl));
    XT_LAUNCHED();
    XT_TRY(hipEventRecord(e1, d->stream));
    XT_TRY(hipMemcpyAsync(&hr->ctl, ctl.p, sizeof(KdbCtl), hipMemcpyDeviceToHost, d->stream));
    XT_TRY(hipMemcpyAsync(o.bounds, bounds.p, 24, hipMemcpyDeviceToHost, d->stream));
    XT_TRY(hipStreamSynchronize(d->stream));
    if (hr->ctl.error) {
        d->err = kdb_error(hr->ctl.error);
        out_free(d, o);
        return XT_E_STATE;
    }
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    o.stats.ms = ms;
    o.stats.n_nodes = total + 1;
    o.stats.n_refs = pool_total;
    o.stats.depth = depth;
    o.stats.leaves = leaves;
    o.stats.levels = (uint32_t)levels.size();
    return XT_OK;
}

// Frees the scene's current kd arrays (tracked in its allocation list).
void drop_kd(xt_dev *d, DevScene &ds) {
    for (const void *p : {(const void *)ds.view.kd, (const void *)ds.view.kd_refs, (const void *)ds.view.kd_leaf}) {
        auto it = std::find(ds.allocs.begin(), ds.allocs.end(), p);
        if (p && it != ds.allocs.end()) {
            (void)hipFree(*it);
            ds.allocs.erase(it);
        }
    }
    ds.bytes -= std::min<uint64_t>(ds.bytes, ds.kd_bytes);
    d->used -= std::min<uint64_t>(d->used, ds.kd_bytes);
    ds.kd_bytes = 0;
    ds.view.kd = nullptr;
    ds.view.kd_refs = nullptr;
    ds.view.kd_leaf = nullptr;
}

}  // namespace

int xt_kd_build_into(xt_dev *d, DevScene &ds, const xt_build_params *p) {
    KdOut o;
    int rc = build(d, ds.view, p, o);
    if (rc) return rc;
    XT_TRY(hipStreamSynchronize(d->stream));
    drop_kd(d, ds);
    ds.allocs.push_back(o.nodes);
    ds.allocs.push_back(o.refs);
    ds.allocs.push_back(o.leaf);
    ds.kd_bytes = o.nodes_bytes + o.refs_bytes + o.leaf_bytes;
    ds.bytes += ds.kd_bytes;
    ds.view.kd = o.nodes;
    ds.view.kd_refs = o.refs;
    ds.view.kd_leaf = o.leaf;
    for (int a = 0; a < 3; ++a) {
        ds.view.kb_lo[a] = o.bounds[a];
        ds.view.kb_hi[a] = o.bounds[3 + a];
    }
    ds.n_kd_nodes = o.stats.n_nodes;
    ds.n_kd_refs = o.stats.n_refs;
    ds.kd_stats = o.stats;
    return XT_OK;
}

extern "C" {

int xt_scene_build_kd_device(xt_dev *d, uint32_t scene_id, const xt_build_params *p, xt_kd_build_stats *stats) {
    if (!d) return XT_E_INVALID;
    auto it = d->scenes.find(scene_id);
    if (it == d->scenes.end()) return XT_E_NOTFOUND;
    if (p && (p->kd_max_depth < 0 || p->kd_leaf_prims < 0 || p->kd_cost_intersect < 0.f || p->kd_max_depth > 64)) {
        d->err = "bad kd-tree build parameters";
        return XT_E_INVALID;
    }
    if (xt_serve_uses_scene(d, scene_id)) {
        d->err = "scene is in use by served jobs";
        return XT_E_STATE;
    }
    XT_TRY(hipSetDevice(d->device));
    XT_TRY(hipStreamSynchronize(d->stream));
    xt_serve_forget_scene(d, scene_id);  // idle serving lanes hold the scene's old arrays
    int rc = xt_kd_build_into(d, it->second, p);
    if (rc) return rc;
    if (stats) *stats = it->second.kd_stats;
    return XT_OK;
}

int xt_scene_get_kd_device(xt_dev *d, uint32_t scene_id, uint32_t *nodes, uint32_t *refs, float *bounds,
                           xt_kd_build_stats *stats) {
    if (!d) return XT_E_INVALID;
    auto it = d->scenes.find(scene_id);
    if (it == d->scenes.end()) return XT_E_NOTFOUND;
    const DevScene &ds = it->second;
    XT_TRY(hipSetDevice(d->device));
    if (nodes && ds.n_kd_nodes)
        XT_TRY(hipMemcpyAsync(nodes, ds.view.kd, 8ull * ds.n_kd_nodes, hipMemcpyDeviceToHost, d->stream));
    if (refs && ds.n_kd_refs)
        XT_TRY(hipMemcpyAsync(refs, ds.view.kd_refs, 4ull * ds.n_kd_refs, hipMemcpyDeviceToHost, d->stream));
    XT_TRY(hipStreamSynchronize(d->stream));
    if (bounds)
        for (int a = 0; a < 3; ++a) {
            bounds[a] = ds.view.kb_lo[a];
            bounds[3 + a] = ds.view.kb_hi[a];
        }
    if (stats) {
        *stats = ds.kd_stats;
        stats->n_nodes = ds.n_kd_nodes;
        stats->n_refs = ds.n_kd_refs;
    }
    return XT_OK;
}

}  // extern "C"
