#include "octree.h"

#include <cmath>
#include <deque>

namespace xt {

namespace {

inline bool contains(const Box64 &b, const dvec3 &p) {
    return p.x >= b.lo.x && p.y >= b.lo.y && p.z >= b.lo.z && p.x <= b.hi.x && p.y <= b.hi.y && p.z <= b.hi.z;
}

// AABB3(a, b): per-axis min/max of two corners (src/xtcore/math/aabb.cc:20-24).
inline Box64 make_box(const dvec3 &a, const dvec3 &b) {
    Box64 r;
    r.lo = dvec3(a.x <= b.x ? a.x : b.x, a.y <= b.y ? a.y : b.y, a.z <= b.z ? a.z : b.z);
    r.hi = dvec3(a.x >= b.x ? a.x : b.x, a.y >= b.y ? a.y : b.y, a.z >= b.z ? a.z : b.z);
    return r;
}

struct Builder {
    const std::vector<Box64> &items;
    size_t max_items, max_depth;
    double root_max_z;
    OctreeBuild &out;

    void subdivide(int32_t ni, size_t level) {
        if (out.nodes[ni].items.size() <= max_items || level >= max_depth) return;
        const Box64 nb = out.nodes[ni].box;
        const dvec3 c = (nb.lo + nb.hi) / 2.0;
        Box64 cb[8];
        cb[0] = make_box(nb.lo, c);
        cb[1] = make_box(dvec3(c.x, nb.lo.y, nb.lo.z), dvec3(nb.hi.x, c.y, c.z));
        cb[2] = make_box(dvec3(c.x, c.y, nb.lo.z), dvec3(nb.hi.x, nb.hi.y, c.z));
        cb[3] = make_box(dvec3(nb.lo.x, c.y, nb.lo.z), dvec3(c.x, nb.hi.y, c.z));
        cb[4] = make_box(dvec3(nb.lo.x, nb.lo.y, c.z), dvec3(c.x, c.y, nb.hi.z));
        cb[5] = make_box(dvec3(c.x, nb.lo.y, c.z), dvec3(nb.hi.x, c.y, nb.hi.z));
        cb[6] = make_box(c, nb.hi);
        cb[7] = make_box(dvec3(nb.lo.x, c.y, c.z), dvec3(c.x, nb.hi.y, root_max_z));
        for (int i = 0; i < 8; ++i) {
            OctreeBuild::Node child;
            child.box = cb[i];
            child.level = (uint32_t)level + 1;
            std::vector<uint32_t> keep;
            keep.reserve(out.nodes[ni].items.size());
            for (uint32_t it : out.nodes[ni].items) {
                const Box64 &ib = items[it];
                if (contains(cb[i], ib.lo) && contains(cb[i], ib.hi)) child.items.push_back(it);
                else keep.push_back(it);
            }
            out.nodes[ni].items.swap(keep);
            int32_t ci = (int32_t)out.nodes.size();
            out.nodes.push_back(std::move(child));
            out.nodes[ni].child[i] = ci;
        }
        for (int i = 0; i < 8; ++i) subdivide(out.nodes[ni].child[i], level + 1);
    }
};

}  // namespace

void octree_build(const std::vector<Box64> &items, size_t max_items, size_t max_depth, OctreeBuild &out) {
    out.nodes.clear();
    Box64 root;  // starts as the (0,0,0) box and is augmented by every item box
    for (const Box64 &b : items) {
        for (int a = 0; a < 3; ++a) {
            if (b.hi[a] > root.hi[a]) root.hi[a] = b.hi[a];
            if (b.lo[a] < root.lo[a]) root.lo[a] = b.lo[a];
        }
    }
    OctreeBuild::Node r;
    r.box = root;
    r.items.resize(items.size());
    for (size_t i = 0; i < items.size(); ++i) r.items[i] = (uint32_t)i;
    out.nodes.push_back(std::move(r));
    Builder b{items, max_items > 0 ? max_items : 1, max_depth > 0 ? max_depth : 1, root.hi.z, out};
    if (out.nodes[0].items.size() > b.max_items) b.subdivide(0, 0);
}

void octree_preorder(const OctreeBuild &t, std::vector<uint32_t> &order) {
    order.clear();
    if (t.nodes.empty()) return;
    std::vector<int32_t> stack{0};
    while (!stack.empty()) {
        int32_t n = stack.back();
        stack.pop_back();
        const auto &node = t.nodes[n];
        order.insert(order.end(), node.items.begin(), node.items.end());
        for (int i = 7; i >= 0; --i)
            if (node.child[i] >= 0) stack.push_back(node.child[i]);
    }
}

void octree_flatten(const OctreeBuild &t, const std::vector<uint32_t> &item_ref, std::vector<XtOctNode> &nodes,
                    std::vector<uint32_t> &refs) {
    nodes.clear();
    refs.clear();
    if (t.nodes.empty()) return;
    // subtree item counts (children are created after their parent)
    std::vector<uint32_t> sub(t.nodes.size(), 0);
    for (size_t i = t.nodes.size(); i-- > 0;) {
        uint32_t s = (uint32_t)t.nodes[i].items.size();
        for (int c = 0; c < 8; ++c)
            if (t.nodes[i].child[c] >= 0) s += sub[t.nodes[i].child[c]];
        sub[i] = s;
    }
    std::deque<int32_t> q{0};
    std::vector<int32_t> order;
    std::vector<uint32_t> pos(t.nodes.size(), 0);
    // breadth-first: assign positions so that the 8 children are consecutive
    order.push_back(0);
    for (size_t k = 0; k < order.size(); ++k) {
        const auto &n = t.nodes[order[k]];
        if (n.child[0] >= 0)
            for (int c = 0; c < 8; ++c) order.push_back(n.child[c]);
    }
    for (size_t k = 0; k < order.size(); ++k) pos[order[k]] = (uint32_t)k;
    nodes.resize(order.size());
    for (size_t k = 0; k < order.size(); ++k) {
        const auto &n = t.nodes[order[k]];
        XtOctNode &o = nodes[k];
        for (int a = 0; a < 3; ++a) {
            float lo = (float)n.box.lo[a], hi = (float)n.box.hi[a];
            if ((double)lo > n.box.lo[a]) lo = std::nextafter(lo, -INFINITY);
            if ((double)hi < n.box.hi[a]) hi = std::nextafter(hi, INFINITY);
            o.lo[a] = lo;
            o.hi[a] = hi;
        }
        o.first_item = (uint32_t)refs.size();
        o.n_items = (uint32_t)n.items.size();
        for (uint32_t it : n.items) refs.push_back(item_ref[it]);
        o.first_child = n.child[0] >= 0 ? pos[n.child[0]] : 0;
        o.subtree_items = sub[order[k]];
        o.pad[0] = o.pad[1] = 0;
    }
}

}  // namespace xt
