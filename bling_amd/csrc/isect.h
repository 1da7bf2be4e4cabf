// Device-side scene layout and the traversal kernel launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "dev_geom.h"

namespace bgpu {

// BVH node of width W (2, 4 or 8), 8 * W floats (W = 4: one 128-byte line):
//   lo.x[W] hi.x[W] lo.y[W] hi.y[W] lo.z[W] hi.z[W] link[W] h[W]
//   link >= 0: inner node; link < 0: leaf ~((first<<4)|count); kNoChild: unused
//   h: the shape tree's per-child margin factor (below); 0 in the triangle tree
// Leaf records, 48 B each, in leaf order:
//   r0 = (p1.xyz, prim id bits), r1 = (e1.xyz, kind bits), r2 = (e2.xyz, 0)
//   kind >= 0: triangle (e1 = p2-p1, e2 = p3-p1); kind < 0: shape index ~kind.
//
// Triangles and analytic shapes live in two BVHs sharing these arrays: the
// triangle tree is rooted at node tri_root, the shape tree at shape_root (-1
// when absent).
//
// The reference's binary32 sphere quadratic (Shape.hs:173-187, solveQuadric
// Math.hs:124-139) reports hits slightly outside the true surface, and a
// shape-tree box must contain every point it reports.  In object space, with
// u = 2^-24, the computed a = |d|^2, b = 2 o.d, c = |o|^2 - r^2 and
// b*b - (4a)*c carry an absolute error of at most 64 u |o|^2 |d|^2 in
// total, while the exact discriminant is 4 |d|^2 (r^2 - dist^2) (dist: the
// ray line's distance from the centre).  A reported hit therefore has
// dist^2 <= r^2 + 16 u |o|^2, and its t lies within 4 sqrt(u) |o| / |d| of
// the closest approach, so the hit point is within
// sqrt(r^2 + 32 u |o|^2) of the centre: an excess of at most
//   min(K0 |o|, K0^2 |o|^2 / (2 r)),  K0 = kSphereK0 >= sqrt(32 u).
// Mapped to world space with sigma(w2o) |o_world - c| >= |o| and a factor
// sigma(o2w) on distances, the children of a shape-tree node are tested
// grown by the margin computed in trav::node_margins: shape_pad_k
// (= K0 * max sigma(o2w) sigma(w2o)) bounds the first form for every shape
// kind, the per-child factor h = sigma(o2w) sigma(w2o)^2 / (2 r) (the largest
// over the child's subtree) the second for spheres, and shape_floor_k covers the
// rounding of the ray's transform and of t.  Cylinders (the same quadratic
// in two dimensions, then range tests on the hit point), disks, quads and
// boxes keep the first form.
constexpr float kSphereK0 = 2e-3f;
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
  float shape_pad_k;    // K0 * max sigma(o2w) sigma(w2o) over the shapes
  int32_t num_nodes;
  int32_t width;        // 2, 4 or 8 children per node
  float shape_floor_k;  // rounding term of the shape margins
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
  int block = 256;
  int persistent = 1;  // lanes refill with new rays (1) or one ray per thread (0)
  int stack = 12;      // traversal stack entries per lane kept in LDS
  int refill = 32;     // lane refill: idle lanes that trigger it (1..64)
  int walk = 16;       // walk inner nodes while at least this many lanes do (0: until all are at leaves)
};
// The range of TraceConfig::stack (and of the bounce loop's pt.stack); deeper
// entries spill to private memory.
constexpr int kMinStackLds = 8, kMaxStackLds = 24;

// Words of the per-GPU scratch array of the allocation-free ray batches
// (256 bytes); scratch_address returns its device address on the current
// device.
constexpr int kScratchWords = 64;
hipError_t scratch_address(void** ptr);
// Zeroes n 32-bit words on stream s with a kernel.
hipError_t zero_words(unsigned int* p, int n, hipStream_t s);

// `work`: two zeroed-on-launch counters for the persistent kernels (may be
// null, which selects one ray per thread).
hipError_t launch_closest(const DevScene& sc, const DevRays& r, int n, const DevHits& h,
                          const TraceConfig& cfg, hipStream_t s, unsigned long long* counters,
                          unsigned int* work);
hipError_t launch_any(const DevScene& sc, const DevRays& r, int n, uint8_t* occ,
                      const TraceConfig& cfg, hipStream_t s, unsigned long long* counters,
                      unsigned int* work);

}  // namespace bgpu
