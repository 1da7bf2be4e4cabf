// Host binned-SAH BVH builder (replaces the SAH kd-tree of
// src/lib/Graphics/Bling/Primitive/KdTree.hs:107-203; the result of a query
// does not depend on the structure, only the work does).
#pragma once
#include <cstdint>
#include <vector>

namespace bgpu {

struct Box {
  float lo[3], hi[3];
};

// Binary BVH node as built on the host.  Children are inner-node indices
// (>= 0) or leaves encoded as ~((first << kLeafCountBits) | count).
struct BNode {
  Box child[2];
  int32_t link[2];
};

constexpr int kLeafCountBits = 4;
constexpr int kMaxLeaf = (1 << kLeafCountBits) - 1;  // 15
constexpr int kMaxDepth = 60;                         // bounds the traversal stack

struct BuildParams {
  int bins = 12;          // binned SAH, 12 bins
  int max_leaf = 4;       // SAH leaf threshold
  float cost_node = 1.0f; // traversal cost relative to one primitive test
  int threads = 0;        // 0 = hardware concurrency
};

struct BuildStats {
  int nodes = 0, leaves = 0, max_depth = 0, max_leaf_size = 0;
  int width = 2, stack_bound = 0;
  int device = 0;  // 1: the triangle tree was built on the GPU
  double sah = 0.0, build_ms = 0.0;
};

// Builds over per-primitive bounds; `order` receives primitive ids in leaf
// order.  Node boxes are padded outward so that the float slab test on the
// device is conservative for every primitive hit.
void build_bvh2(const std::vector<Box>& prim_bounds, const BuildParams& p,
                std::vector<BNode>* nodes, std::vector<int32_t>* order, BuildStats* stats);

// A node of a BVH of width w: up to w children, inner links index the output
// array, leaves keep the BVH2 encoding.
constexpr int kMaxWidth = 8;
struct WNode {
  Box child[kMaxWidth];
  int32_t link[kMaxWidth];
  int n = 0;  // children in use
};
// Appends the subtree of `bin` rooted at `root` to `out` as a BVH of width
// `width` (2, 4 or 8), opening the largest inner child until each node is
// full; returns the index of the new root.
int collapse_bvh(const std::vector<BNode>& bin, int root, int width, std::vector<WNode>* out);
// Exact worst case of a traversal stack that pushes every hit child but the
// one it descends into: the maximum over root-to-leaf paths of the summed
// (children - 1).
int stack_bound(const std::vector<WNode>& w, int root);

Box pad_box(const Box& b);

}  // namespace bgpu
