// Device build of the triangle BVH (bvh_device.hip): the host builder's
// binned-SAH tree, built and collapsed on the GPU.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace bgpu {

struct Ctx;

// Ranges of at most this many primitives are built by one workgroup each;
// scenes with no more triangles than this are built on the host.
constexpr int kSmall = 1024;

class DeviceBvh {
 public:
  DeviceBvh();
  ~DeviceBvh();
  DeviceBvh(const DeviceBvh&) = delete;
  DeviceBvh& operator=(const DeviceBvh&) = delete;
  // Binary tree over `ntri` triangles (device vertices and indices, as the
  // scene holds them).  *fallback: a range needs the host's median split,
  // a primitive box is not finite or the scene is small; build on the host.
  int build(Ctx* ctx, const float* d_verts, const int32_t* d_idx, int ntri, bool* fallback);
  // Collapses to `width` children per node into an internal buffer (node
  // layout of the traversal, 8 * width floats per node, root 0); returns the
  // node count and the exact worst-case traversal stack.
  int collapse(Ctx* ctx, int width, int* count, int* stack_bound);
  const float* wide_nodes() const;
  // Triangle leaf records (3 float4 per triangle, in leaf order) into dst.
  int leaves_to(Ctx* ctx, const float* d_verts, const int32_t* d_idx, float4* dst);
  void release();

  int nodes2_used = 0, leaves = 0, max_depth = 0, fallback_reason = 0;
  double build_ms = 0.0;

 private:
  struct Impl;
  Impl* impl_;
};

}  // namespace bgpu
