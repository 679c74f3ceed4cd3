// Host BVH construction and wide-node flattening.
#pragma once

#include <stdint.h>

#include <vector>

namespace ilm {

struct Bvh2Node {
    float lo[3], hi[3];
    int32_t left = -1, right = -1;  // children (internal)
    int32_t first = 0, count = 0;   // primitive range in Bvh2::order (leaf if count > 0 or left < 0)
    bool leaf() const { return left < 0; }
};

struct Bvh2 {
    std::vector<Bvh2Node> nodes;   // nodes[0] is the root
    std::vector<int32_t> order;    // primitive index of each leaf slot (kd-trees repeat primitives)
};

// Binned SAH over centroids (32 bins, 3 axes), leaf size <= max_leaf.
Bvh2 build_sah(const float* pos, const int32_t* idx, int64_t nt, int max_leaf, int threads);

// The reference BVHMesh topology (Shape/BVHMesh.cpp:78-107,203-269 with
// Shape/TreeMesh.cpp:44-85,159-212): 128-bin count heuristic over the node
// box, round-robin axes after the largest-extent first axis, midpoint
// distribution, leaf when n <= max_leaf, depth == max_depth or the node
// radius <= root radius / 1000.
Bvh2 build_reference(const float* pos, const int32_t* idx, int64_t nt, int max_leaf, int max_depth);

// The reference kd-tree topologies as trees of cells with shared triangle
// references: KDTreeMesh (count-heuristic planes; Shape/KDTreeMesh.cpp:114-135,
// 941-1005) when median is false, KDTreeMeshEx (cell-centre planes;
// Shape/KDTreeMeshEx.cpp:59-146) when it is true.
Bvh2 build_kdtree(const float* pos, const int32_t* idx, int64_t nt, int max_leaf, int max_depth, bool median);

// A kd-tree stored in PersistentMesh's tree file (Shape/PersistentMesh.cpp:
// 171-245, written by KDTreeMesh.cpp:167-257): nodes of an 8-byte header
// (float partition; uint32 axis in bits 0-1, item count in bits 2-31), an
// interior node (axis 0-2) followed by the byte offsets of its below / above
// children, a leaf (axis 3) by its triangle indices. Returns the tree of
// cells: the root cell is the vertices' bounds widened by 1e-4 (the
// KDTreeMesh build's root cell, KDTreeMesh.cpp:951), a child's cell is the
// parent's cut at the partition (clamped to the cell). Throws
// std::invalid_argument for a malformed file (offsets outside the file or
// misaligned, a node reached twice, triangle indices out of range, more
// than 64 levels).
Bvh2 import_kdtree(const uint8_t* tre, size_t bytes, const float* pos, int64_t nv, int64_t nt);

struct WideBvh {
    int width = 4;
    std::vector<uint8_t> nodes;    // WideNode<width> array, root at 0
    int64_t nnodes = 0;
    std::vector<int32_t> order;    // original triangle index per triangle slot
    int depth = 0, max_leaf = 0, stack_need = 0;
    double sah_cost = 0;
};

// Collapses a binary tree into a wide tree: each wide node repeatedly opens
// its largest-area internal child until it holds `width` children.
WideBvh collapse(const Bvh2& b, int width);

// collapse() for the subtree at `root`; the result's order is left empty
// (leaf slots index the caller's order array).
WideBvh collapse_from(const Bvh2& b, int32_t root, int width);

// The topology of BVH4 trees (nodes: WideNode<4> array, trees at `roots`)
// as binary trees over the same leaf slots: each node's valid children are
// paired (c0 c1)(c2 c3) and boxes are unions, so collapsing them to another
// width keeps every leaf's triangle range. bin_roots[i] = binary root of roots[i].
Bvh2 binary_from_bvh4(const uint8_t* nodes, const std::vector<uint32_t>& roots, std::vector<int32_t>& bin_roots);

// Triangle array in leaf order for the device: 3 float4 per triangle,
// (v0, original index as int bits), (v1, 0), (v2, 0).
void pack_triangles(const float* pos, const int32_t* idx, const std::vector<int32_t>& order,
                    std::vector<float>& out);

}  // namespace ilm
