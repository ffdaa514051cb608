#include "kdtree.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <thread>

namespace xt {

namespace {

struct Local {
    std::vector<XtKdNode> nodes;
    std::vector<uint32_t> refs;
    int depth = 0;
    size_t leaves = 0;
};

struct Task {
    uint32_t slot;
    Box32 box;
    std::vector<uint32_t> refs;
    int depth;
};

inline float area(const Box32 &b) {
    float dx = b.hi[0] - b.lo[0], dy = b.hi[1] - b.lo[1], dz = b.hi[2] - b.lo[2];
    return 2.f * (dx * dy + dy * dz + dz * dx);
}

struct Builder {
    const std::vector<Box32> &pb;
    KdParams p;
    int spawn_depth;  // -1: never spawn
    std::vector<Task> *spawned;

    void leaf(Local &L, uint32_t slot, const std::vector<uint32_t> &refs, int depth) {
        XtKdNode n;
        n.word = ((uint32_t)L.refs.size() << 2) | 3u;
        n.count = (uint32_t)refs.size();
        L.nodes[slot] = n;
        L.refs.insert(L.refs.end(), refs.begin(), refs.end());
        L.depth = std::max(L.depth, depth);
        L.leaves++;
    }

    bool find_split(const Box32 &box, const std::vector<uint32_t> &refs, int &axis, float &split) {
        const size_t n = refs.size();
        const float sa = area(box);
        if (!(sa > 0.f)) return false;
        float best = p.cost_intersect * (float)n;
        bool found = false;
        const int K = p.bins;
        std::vector<uint32_t> st(K), en(K);
        for (int a = 0; a < 3; ++a) {
            const float lo = box.lo[a], ext = box.hi[a] - lo;
            if (!(ext > 0.f)) continue;
            std::fill(st.begin(), st.end(), 0u);
            std::fill(en.begin(), en.end(), 0u);
            const float scale = (float)K / ext;
            for (uint32_t r : refs) {
                float l = std::max(pb[r].lo[a], lo), h = std::min(pb[r].hi[a], box.hi[a]);
                int s = (int)((l - lo) * scale), e = (int)((h - lo) * scale);
                s = std::min(std::max(s, 0), K - 1);
                e = std::min(std::max(e, 0), K - 1);
                st[s]++;
                en[e]++;
            }
            size_t nl = 0, nr = n;
            for (int k = 1; k < K; ++k) {
                nl += st[k - 1];
                nr -= en[k - 1];
                float pos = lo + ext * (float)k / (float)K;
                if (!(pos > lo && pos < box.hi[a])) continue;
                Box32 lb = box, rb = box;
                lb.hi[a] = pos;
                rb.lo[a] = pos;
                float c = p.cost_traverse + p.cost_intersect * (area(lb) * (float)nl + area(rb) * (float)nr) / sa;
                if (nl == 0 || nr == 0) c *= (1.f - p.empty_bonus);
                if (c < best) {
                    best = c;
                    axis = a;
                    split = pos;
                    found = true;
                }
            }
        }
        return found;
    }

    void build(Local &L, uint32_t slot, const Box32 &box, std::vector<uint32_t> &refs, int depth) {
        if (spawn_depth >= 0 && depth == spawn_depth && refs.size() > (size_t)p.leaf_prims) {
            spawned->push_back(Task{slot, box, std::move(refs), depth});
            return;
        }
        int axis = 0;
        float split = 0.f;
        if ((int)refs.size() <= p.leaf_prims || depth >= p.max_depth || !find_split(box, refs, axis, split)) {
            leaf(L, slot, refs, depth);
            return;
        }
        std::vector<uint32_t> left, right;
        left.reserve(refs.size());
        right.reserve(refs.size());
        for (uint32_t r : refs) {
            float l = pb[r].lo[axis], h = pb[r].hi[axis];
            bool to_l = l < split || (l <= split && h <= split);
            bool to_r = h > split || (l >= split && h >= split);
            if (to_l) left.push_back(r);
            if (to_r) right.push_back(r);
        }
        std::vector<uint32_t>().swap(refs);
        uint32_t c = (uint32_t)L.nodes.size();
        L.nodes.resize(c + 2);
        XtKdNode in;
        in.word = (c << 2) | (uint32_t)axis;
        in.split = split;
        L.nodes[slot] = in;
        Box32 lb = box, rb = box;
        lb.hi[axis] = split;
        rb.lo[axis] = split;
        build(L, c, lb, left, depth + 1);
        build(L, c + 1, rb, right, depth + 1);
    }
};

}  // namespace

void kdtree_build(const std::vector<Box32> &pb, const KdParams &params, KdTree &out) {
    KdParams p = params;
    const size_t n = pb.size();
    if (p.max_depth <= 0) p.max_depth = (int)std::lround(8.0 + 1.3 * std::log2((double)std::max<size_t>(n, 1)));
    if (p.bins < 4) p.bins = 4;
    int threads = p.threads > 0 ? p.threads : (int)std::max(1u, std::thread::hardware_concurrency());

    Box32 root;
    for (int a = 0; a < 3; ++a) {
        root.lo[a] = INFINITY;
        root.hi[a] = -INFINITY;
    }
    for (const Box32 &b : pb)
        for (int a = 0; a < 3; ++a) {
            root.lo[a] = std::min(root.lo[a], b.lo[a]);
            root.hi[a] = std::max(root.hi[a], b.hi[a]);
        }
    if (n == 0)
        for (int a = 0; a < 3; ++a) root.lo[a] = root.hi[a] = 0.f;

    std::vector<uint32_t> refs(n);
    for (size_t i = 0; i < n; ++i) refs[i] = (uint32_t)i;

    int spawn = 0;
    while ((1 << spawn) < threads * 4 && spawn < 10) ++spawn;
    if (n < 50000 || threads == 1) spawn = -1;

    std::vector<Task> tasks;
    Builder top{pb, p, spawn, &tasks};
    Local G;
    G.nodes.resize(1);
    top.build(G, 0, root, refs, 0);

    std::vector<Local> locals(tasks.size());
    std::vector<std::thread> pool;
    {
        std::vector<size_t> order(tasks.size());
        for (size_t i = 0; i < order.size(); ++i) order[i] = i;
        std::sort(order.begin(), order.end(),
                  [&](size_t a, size_t b) { return tasks[a].refs.size() > tasks[b].refs.size(); });
        std::atomic_size_t cursor{0};
        auto worker = [&]() {
            for (;;) {
                size_t k = cursor.fetch_add(1);
                if (k >= order.size()) return;
                Task &t = tasks[order[k]];
                Builder b{pb, p, -1, nullptr};
                Local &L = locals[order[k]];
                L.nodes.resize(1);
                b.build(L, 0, t.box, t.refs, t.depth);
            }
        };
        int nt = std::min<int>(threads, (int)tasks.size());
        for (int i = 0; i < nt; ++i) pool.emplace_back(worker);
        for (auto &th : pool) th.join();
    }

    // stitch spawned subtrees into the global arrays
    for (size_t k = 0; k < tasks.size(); ++k) {
        Local &L = locals[k];
        const uint32_t base = (uint32_t)G.nodes.size();
        const uint32_t rbase = (uint32_t)G.refs.size();
        auto reloc = [&](XtKdNode nd) {
            if ((nd.word & 3u) == 3u) nd.word = (((nd.word >> 2) + rbase) << 2) | 3u;
            else nd.word = (((nd.word >> 2) + base - 1) << 2) | (nd.word & 3u);
            return nd;
        };
        G.nodes[tasks[k].slot] = reloc(L.nodes[0]);
        for (size_t i = 1; i < L.nodes.size(); ++i) G.nodes.push_back(reloc(L.nodes[i]));
        G.refs.insert(G.refs.end(), L.refs.begin(), L.refs.end());
        G.depth = std::max(G.depth, L.depth);
        G.leaves += L.leaves;
    }
    out.nodes.swap(G.nodes);
    out.refs.swap(G.refs);
    out.bounds = root;
    out.depth = G.depth;
    out.leaves = G.leaves;
}

}  // namespace xt
