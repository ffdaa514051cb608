// Octree build with the reference's rules (src/xtcore/octree.tml:97-198,
// src/xtcore/mesh.cc:30-113, limits src/xtcore/mesh.h:11-12):
//   * the root box is the union of the item boxes and the origin
//     (Octree::clear() starts from a (0,0,0) box, octree.tml:89-95);
//   * a node with more than max_items items below max_depth gets all 8
//     children; child boxes follow octree.tml:150-166, including child 7
//     whose z-maximum is the ROOT's z-maximum;
//   * each item moves to the FIRST child (0..7) whose closed box contains the
//     item's box (for a triangle: all three vertices, octree.tml:175-187);
//     items that fit no child stay in the node, in insertion order.
// The pre-order of items (node items, then children 0..7) is the order in
// which Octree::r_intersection (octree.tml:206-249) meets candidates, i.e.
// the reference's tie order between hits at equal t.
#pragma once
#include <cstdint>
#include <vector>

#include "../xt_layout.h"
#include "xt_vec.h"

namespace xt {

struct Box64 {
    dvec3 lo, hi;
};

struct OctreeBuild {
    struct Node {
        Box64 box;
        std::vector<uint32_t> items;  // item indices, insertion order
        int32_t child[8] = {-1, -1, -1, -1, -1, -1, -1, -1};
        uint32_t level = 0;
    };
    std::vector<Node> nodes;  // nodes[0] is the root
};

void octree_build(const std::vector<Box64> &items, size_t max_items, size_t max_depth, OctreeBuild &out);
void octree_preorder(const OctreeBuild &t, std::vector<uint32_t> &order);

// Breadth-first device layout; `item_ref` maps an item index to the value
// stored in the reference list (e.g. its soup index).
void octree_flatten(const OctreeBuild &t, const std::vector<uint32_t> &item_ref, std::vector<XtOctNode> &nodes,
                    std::vector<uint32_t> &refs);

}  // namespace xt
