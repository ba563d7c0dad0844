// Binary BVH builders used by the scene layer.
//
//  * BuildSahBvh: per-mesh BLAS, binned object-split SAH. Same decisions as the
//    reference's SplitBvh(2.0, 64, 0, 0.001, 0) (src/core/Mesh.h:37), whose
//    spatial splits never run because max_split_depth is 0
//    (thirdparty/RadeonRays/split_bvh.cpp:60-218, FindObjectSahSplit :220-337).
//  * BuildMedianBvh: TLAS over instance boxes, the reference's Bvh with SAH off
//    (thirdparty/RadeonRays/bvh.cpp:68-243, built at src/core/Scene.cpp:148-187).
//
// Both produce an index-based tree plus the packed primitive order that the
// flattening step turns into the GPU-visible arrays. The order in which
// subtrees are visited fixes the packed order, so it is part of the contract:
// SAH builds the right child first, the median builder the left child first.
#pragma once

#include <vector>

#include "pt_math.h"

namespace glslpt {

struct BvhBuildNode {
    BBox bounds;
    bool leaf = false;
    int lc = -1, rc = -1;       // internal: child node indices
    int start = 0, count = 0;   // leaf: range in BvhTree::packed
};

struct BvhTree {
    std::vector<BvhBuildNode> nodes;  // nodes[0] is the root
    std::vector<int> packed;          // primitive ids in leaf order
    BBox bounds;                      // union of all primitive boxes
    int height = 0;
};

// prims[i] is the box of primitive i. n may be 0 (empty tree with no nodes).
BvhTree BuildSahBvh(const BBox* prims, int n, float traversal_cost = 2.0f, int num_bins = 64);
BvhTree BuildMedianBvh(const BBox* prims, int n);

}  // namespace glslpt
