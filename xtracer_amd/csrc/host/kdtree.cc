#include "kdtree.h"

#include <algorithm>
#include <atomic>
#include <deque>
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


// A reference: primitive id and its bounds inside the current node.
struct Ref {
    uint32_t id;
    Box32 b;
};

struct Task {
    uint32_t slot;
    Box32 box;
    std::vector<Ref> refs;
    int depth;
};

// Bounds of triangle v (9 doubles) clipped to box (Sutherland-Hodgman in
// FP64, inclusive at the planes), widened outward by a relative 1e-6 of the
// box extent and intersected with the box. False when the triangle does not
// reach the box.
bool clip_triangle(const double *v, const Box32 &box, Box32 &out) {
    double poly[2][12][3];
    int n = 3, cur = 0;
    for (int k = 0; k < 3; ++k)
        for (int a = 0; a < 3; ++a) poly[0][k][a] = v[3 * k + a];
    for (int a = 0; a < 3 && n > 0; ++a) {
        for (int side = 0; side < 2 && n > 0; ++side) {
            const double plane = side == 0 ? (double)box.lo[a] : (double)box.hi[a];
            const double sgn = side == 0 ? 1.0 : -1.0;  // inside: sgn * (x - plane) >= 0
            int m = 0;
            double(*src)[3] = poly[cur];
            double(*dst)[3] = poly[cur ^ 1];
            for (int k = 0; k < n && m < 11; ++k) {
                const double *p = src[k], *q = src[(k + 1) % n];
                const double dp = sgn * (p[a] - plane), dq = sgn * (q[a] - plane);
                if (dp >= 0) {
                    for (int c = 0; c < 3; ++c) dst[m][c] = p[c];
                    ++m;
                }
                if ((dp >= 0) != (dq >= 0) && m < 11) {
                    const double t = dp / (dp - dq);
                    for (int c = 0; c < 3; ++c) dst[m][c] = p[c] + t * (q[c] - p[c]);
                    dst[m][a] = plane;
                    ++m;
                }
            }
            n = m;
            cur ^= 1;
        }
    }
    if (n == 0) return false;
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int k = 0; k < n; ++k)
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], poly[cur][k][a]);
            hi[a] = std::max(hi[a], poly[cur][k][a]);
        }
    for (int a = 0; a < 3; ++a) {
        const double ext = (double)box.hi[a] - (double)box.lo[a];
        const double pad = 1e-6 * ext + 1e-12 * (std::fabs(lo[a]) + std::fabs(hi[a]));
        const float l = std::nextafter((float)(lo[a] - pad), -INFINITY);
        const float h = std::nextafter((float)(hi[a] + pad), INFINITY);
        out.lo[a] = std::max(l, box.lo[a]);
        out.hi[a] = std::min(h, box.hi[a]);
        if (out.lo[a] > out.hi[a]) return false;
    }
    return true;
}

inline float area(const Box32 &b) {
    float dx = b.hi[0] - b.lo[0], dy = b.hi[1] - b.lo[1], dz = b.hi[2] - b.lo[2];
    return 2.f * (dx * dy + dy * dz + dz * dx);
}

struct Builder {
    const double *verts;  // 9 per prim, or null
    const uint8_t *is_tri;
    KdParams p;
    int spawn_depth;  // -1: never spawn
    std::vector<Task> *spawned;

    void leaf(Local &L, uint32_t slot, const std::vector<Ref> &refs, int depth) {
        XtKdNode n;
        n.word = ((uint32_t)L.refs.size() << 2) | 3u;
        n.count = (uint32_t)refs.size();
        L.nodes[slot] = n;
        for (const Ref &r : refs) L.refs.push_back(r.id);
        L.depth = std::max(L.depth, depth);
        L.leaves++;
    }

    bool find_split(const Box32 &box, const std::vector<Ref> &refs, int &axis, float &split) {
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
            for (const Ref &r : refs) {
                float l = std::max(r.b.lo[a], lo), h = std::min(r.b.hi[a], box.hi[a]);
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

    void build(Local &L, uint32_t slot, const Box32 &box, std::vector<Ref> &refs, int depth) {
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
        std::vector<Ref> left, right;
        left.reserve(refs.size());
        right.reserve(refs.size());
        Box32 lb = box, rb = box;
        lb.hi[axis] = split;
        rb.lo[axis] = split;
        for (const Ref &r : refs) {
            const float l = r.b.lo[axis], h = r.b.hi[axis];
            const bool to_l = l < split || (l <= split && h <= split);
            const bool to_r = h > split || (l >= split && h >= split);
            if (to_l && to_r && l < split && h > split && verts && is_tri[r.id]) {
                // straddler: keep only the sides its surface reaches
                Box32 cb;
                if (clip_triangle(verts + 9 * (size_t)r.id, lb, cb)) left.push_back(Ref{r.id, cb});
                if (clip_triangle(verts + 9 * (size_t)r.id, rb, cb)) right.push_back(Ref{r.id, cb});
                continue;
            }
            if (to_l) {
                Ref c = r;
                c.b.hi[axis] = std::min(c.b.hi[axis], split);
                left.push_back(c);
            }
            if (to_r) {
                Ref c = r;
                c.b.lo[axis] = std::max(c.b.lo[axis], split);
                right.push_back(c);
            }
        }
        std::vector<Ref>().swap(refs);
        uint32_t c = (uint32_t)L.nodes.size();
        L.nodes.resize(c + 2);
        XtKdNode in;
        in.word = (c << 2) | (uint32_t)axis;
        in.split = split;
        L.nodes[slot] = in;
        build(L, c, lb, left, depth + 1);
        build(L, c + 1, rb, right, depth + 1);
    }
};

}  // namespace

void kdtree_build(const std::vector<Box32> &pb, const double *tri_verts, const uint8_t *is_tri,
                  const KdParams &params, KdTree &out) {
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

    std::vector<Ref> refs(n);
    for (size_t i = 0; i < n; ++i) refs[i] = Ref{(uint32_t)i, pb[i]};

    int spawn = 0;
    while ((1 << spawn) < threads * 4 && spawn < 10) ++spawn;
    if (n < 50000 || threads == 1) spawn = -1;

    std::vector<Task> tasks;
    Builder top{tri_verts, is_tri, p, spawn, &tasks};
    Local G;
    // node 1 is an empty leaf that nothing references, so that child pairs
    // start on even nodes: a pair is one aligned 16-byte record, which the
    // traversal loads whole (spawned subtrees keep the parity: each adds an
    // even number of nodes)
    G.nodes.resize(2);
    G.nodes[1].word = 3u;
    G.nodes[1].count = 0;
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
                Builder b{tri_verts, is_tri, p, -1, nullptr};
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

std::vector<XtKdNode> kd_line_layout(const std::vector<XtKdNode> &nodes) {
    if (nodes.size() < 3 || (nodes[0].word & 3u) == 3u) return nodes;
    struct Pair {
        uint32_t old_first;  // first node of the pair in `nodes`
        uint32_t parent;     // the parent's index in the output
    };
    std::vector<XtKdNode> out;
    out.reserve(nodes.size() + 1);
    out.push_back(nodes[0]);
    out.push_back(XtKdNode{});
    out[1].word = 3u;  // padding: an empty leaf
    out[1].count = 0;
    std::deque<Pair> roots;
    roots.push_back(Pair{nodes[0].word >> 2, 0});
    while (!roots.empty()) {
        std::deque<Pair> line;
        line.push_back(roots.front());
        roots.pop_front();
        do {
            const Pair p = line.front();
            line.pop_front();
            const uint32_t at = (uint32_t)out.size();
            out[p.parent].word = (at << 2) | (out[p.parent].word & 3u);
            for (uint32_t k = 0; k < 2; ++k) {
                const XtKdNode nd = nodes[p.old_first + k];
                out.push_back(nd);
                if ((nd.word & 3u) != 3u) line.push_back(Pair{nd.word >> 2, at + k});
            }
        } while (!line.empty() && out.size() % 16 != 0);
        // the line's frontier: each pair roots a line of its own, next
        for (auto it = line.rbegin(); it != line.rend(); ++it) roots.push_front(*it);
    }
    return out;
}

}  // namespace xt
