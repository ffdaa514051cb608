// SAH kd-tree over the primitive soup. The reference has no kd-tree (its
// README lists one, src/xtcore/bvh.cc is an empty stub), so the structure is
// this project's own; only its closest-hit results are checked against the
// reference. Binned SAH with an empty-space bonus. A triangle that straddles
// a split is clipped to each child's box (in FP64, bounds widened outward)
// and is referenced only by the children its surface reaches, with the
// clipped bounds driving the deeper splits.
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

#include "../xt_layout.h"

namespace xt {

struct Box32 {
    float lo[3], hi[3];
};

struct KdParams {
    int max_depth = 0;         // 0: 8 + 1.3 * log2(N)
    int leaf_prims = 4;        // stop splitting at or below this many refs
    float cost_traverse = 1.f;
    float cost_intersect = 12.f;
    float empty_bonus = 0.2f;  // cost reduction when one side is empty
    int bins = 64;
    int threads = 0;           // 0: hardware concurrency
    bool on_device = false;    // leave the tree to the device build at upload
};

struct KdTree {
    std::vector<XtKdNode> nodes;  // nodes[0] is the root, nodes[1] an unreferenced empty leaf; child pairs
                                  // start on even nodes (16-byte aligned)
    std::vector<uint32_t> refs;   // leaf primitive references (soup indices)
    Box32 bounds;                 // root box
    int depth = 0;                // deepest leaf level
    size_t leaves = 0;
};

// tri_verts: 9 doubles per primitive (v0 v1 v2) for the triangles marked in
// is_tri; null disables clipping (every straddler goes to both children).
void kdtree_build(const std::vector<Box32> &prim_boxes, const double *tri_verts, const uint8_t *is_tri,
                  const KdParams &p, KdTree &out);

// The same tree with its child pairs regrouped so that a descent stays in one
// 128-byte line (16 nodes, 8 child pairs) for about three levels at a time:
// each line is filled breadth-first from one pair down, and the pairs left at
// its frontier start the next lines, depth-first. The root stays node 0 and
// node 1 is padding, so pairs start on even nodes and lines on multiples of
// 16. Leaves keep their reference ranges.
std::vector<XtKdNode> kd_line_layout(const std::vector<XtKdNode> &nodes);

}  // namespace xt
