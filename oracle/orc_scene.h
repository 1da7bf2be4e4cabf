// ORACLE -- test infrastructure only (see orc_math.h).
// The flattened scene as the oracle sees it, plus closest-hit / any-hit
// queries with the semantics of the reference accelerator primitive
// (src/lib/Graphics/Bling/Primitive/KdTree.hs:236-246):
//   intersect r  = intersectAABB rootBounds r >>= nearest hit over all prims
//   intersects r = intersectAABB rootBounds r && any prim `intersects` r
// The "nearest hit over all prims" is defined structure-independently as the
// minimum t over all primitives whose own test accepts the ray, ties broken
// by the lowest primitive id (the reference's winner at an exact tie depends
// on kd-tree visit order, Primitive.hs:29-43 with TriangleMesh.hs:166).
#pragma once
#include <vector>

#include "../include/bgpu.h"
#include "orc_geom.h"
#include "orc_texture.h"

namespace orc {

struct ShapePrim {
  Shape shape;
  M4 o2w, w2o;
  int material, emissive, light_id;
  float emission[16];
};

struct Scene {
  std::vector<V3> verts, normals;
  std::vector<float> uvs;
  std::vector<int> idx;       // 3 per triangle
  std::vector<int> tri_mesh;  // per triangle
  std::vector<bgpu_mesh> meshes;
  std::vector<ShapePrim> shapes;
  std::vector<bgpu_material> materials;
  int ntri = 0;
  AABB bounds;                // fold of primitive worldBounds (KdTree.hs:111)
  // oracle-private acceleration: a median-split BVH with double-precision,
  // padded boxes.  Only ever used to prune; results equal brute force.
  struct Node {
    double lo[3], hi[3];
    int left, right;  // inner: children; leaf: right = -1, left = first
    int count;
    int has_shape;    // subtree holds an analytic shape: ray-dependent padding
  };
  std::vector<Node> nodes;
  std::vector<int> order;     // primitive ids in leaf order
  // lights and camera for the path tracer
  std::vector<bgpu_light> lights;
  std::vector<std::vector<float>> light_maps;
  bgpu_camera camera;
  bgpu_film film;
  tex::Textures textures;  // the texture node list and its images

  int nprims() const { return ntri + (int)shapes.size(); }
  V3 v(int i) const { return verts[i]; }
  void tri(int t, V3* p1, V3* p2, V3* p3) const {
    *p1 = verts[idx[3 * t]];
    *p2 = verts[idx[3 * t + 1]];
    *p3 = verts[idx[3 * t + 2]];
  }
  AABB primBounds(int id) const;
};

struct Hit {
  float t, b1, b2;
  int prim;  // -1 = miss
  float eps;
  int axis;
};

// Single-primitive tests, world-space ray.
bool primIntersect(const Scene& s, int id, const Ray& r, Hit* h);
bool primIntersects(const Scene& s, int id, const Ray& r);

// Whole-scene queries (root-box test + brute force or the pruning BVH).
bool sceneIntersect(const Scene& s, const Ray& r, Hit* h, bool brute, int* ties);
bool sceneOccluded(const Scene& s, const Ray& r, bool brute);

Scene* buildScene(const bgpu_scene_desc* d, const char** err);

}  // namespace orc
