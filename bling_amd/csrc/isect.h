// Device-side scene layout and the traversal kernel launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "dev_geom.h"

namespace bgpu {

// BVH node of width W (2, 4 or 8), 8 * W floats (W = 4: one 128-byte line):
//   lo.x[W] hi.x[W] lo.y[W] hi.y[W] lo.z[W] hi.z[W] link[W] pad[W]
//   link >= 0: inner node; link < 0: leaf ~((first<<4)|count); kNoChild: unused
// Leaf records, 48 B each, in leaf order:
//   r0 = (p1.xyz, prim id bits), r1 = (e1.xyz, kind bits), r2 = (e2.xyz, 0)
//   kind >= 0: triangle (e1 = p2-p1, e2 = p3-p1); kind < 0: shape index ~kind.
//
// Triangles and analytic shapes live in two BVHs sharing these arrays: the
// triangle tree is rooted at node tri_root, the shape tree at shape_root (-1
// when absent).  The reference's binary32 sphere/cylinder quadratic
// (Shape.hs:113-187) evaluates c = |o|^2 - r^2 with cancellation and so
// reports hits up to about 3*sqrt(2^-24)*|o - centre| outside the true
// surface (|.| in object space); a box padded only by rounding would cull
// those hits, which the reference does report.  The children of a
// shape-tree node are therefore tested grown by shape_pad_k times the L1
// distance from the ray origin to the farthest corner of the node's box,
// which bounds |o - centre| for every shape below it (shape_pad_k includes
// the anisotropy of the shapes' transforms).
// Traversal stack entries available per lane (LDS part plus private spill);
// scene creation checks the tree's worst case against it.
constexpr int kStackCapacity = 64;
// link of an unused child slot in a node
constexpr int32_t kNoChild = 0x7fffffff;

struct DevScene {
  const float4* nodes;  // 8 * width floats per node
  const float4* leaf;
  const DShape* shapes;
  float root_lo[3], root_hi[3];
  int32_t tri_root;    // -1 if no triangles
  int32_t shape_root;  // -1 if no shapes
  float shape_pad_k;
  int32_t num_nodes;
  int32_t width;       // 2, 4 or 8 children per node
  int32_t pad;
};

struct DevRays {
  const float *ox, *oy, *oz, *dx, *dy, *dz, *tmin, *tmax;
};
struct DevHits {
  float *t, *b1, *b2;
  int32_t* prim;
};

// Launch-shape knobs chosen by the autotuner.
struct TraceConfig {
  int block = 128;
  int persistent = 1;  // lanes refill with new rays (1) or one ray per thread (0)
};

// `work`: two zeroed-on-launch counters for the persistent kernels (may be
// null, which selects one ray per thread).
hipError_t launch_closest(const DevScene& sc, const DevRays& r, int n, const DevHits& h,
                          const TraceConfig& cfg, hipStream_t s, unsigned long long* counters,
                          unsigned int* work);
hipError_t launch_any(const DevScene& sc, const DevRays& r, int n, uint8_t* occ,
                      const TraceConfig& cfg, hipStream_t s, unsigned long long* counters,
                      unsigned int* work);

}  // namespace bgpu
