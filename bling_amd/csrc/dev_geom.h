// Device geometry for gfx950: vectors, transforms and ray/primitive tests in
// IEEE binary32 with the reference's operation order (compiled with
// -ffp-contract=off; '/' and sqrt are correctly rounded in HIP by default),
// so that hit/miss, t and barycentrics are bit-identical to the reference's
// Float arithmetic:
//   Math.hs:124-139,328-353   AABB.hs:79-94   Transform.hs:247-280
//   Primitive/TriangleMesh.hs:140-207   Shape.hs:81-284
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bgpu {

struct f3 {
  float x, y, z;
};
__device__ __forceinline__ f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 add3(f3 a, f3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ f3 sub3(f3 a, f3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ f3 mul3(f3 a, f3 b) { return mk3(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ f3 neg3(f3 a) { return mk3(-a.x, -a.y, -a.z); }
__device__ __forceinline__ f3 scl3(f3 v, float f) { return mk3(v.x * f, v.y * f, v.z * f); }
__device__ __forceinline__ float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float sqlen3(f3 v) { return v.x * v.x + v.y * v.y + v.z * v.z; }
__device__ __forceinline__ f3 cross3(f3 u, f3 v) {
  return mk3(u.y * v.z - u.z * v.y, -(u.x * v.z - u.z * v.x), u.x * v.y - u.y * v.x);
}
__device__ __forceinline__ f3 norm3(f3 v) {
  if (sqlen3(v) != 0.0f) return scl3(v, 1.0f / __builtin_sqrtf(sqlen3(v)));
  return mk3(0.0f, 1.0f, 0.0f);
}
__device__ __forceinline__ float comp3(f3 v, int d) { return d == 0 ? v.x : (d == 1 ? v.y : v.z); }
// Haskell Ord defaults: max x y = if x <= y then y else x
__device__ __forceinline__ float hmax(float x, float y) { return (x <= y) ? y : x; }
__device__ __forceinline__ float hmin(float x, float y) { return (x <= y) ? x : y; }

struct DRay {
  f3 o, d;
  float tmin, tmax;
};
__device__ __forceinline__ f3 ray_at(const DRay& r, float t) { return add3(r.o, scl3(r.d, t)); }

// Row-major 4x4.
__device__ __forceinline__ f3 xform_point(const float* m, f3 p) {
  float xp = m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3];
  float yp = m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7];
  float zp = m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11];
  float wp = m[12] * p.x + m[13] * p.y + m[14] * p.z + m[15];
  if (wp == 1.0f) return mk3(xp, yp, zp);
  return mk3(xp / wp, yp / wp, zp / wp);
}
__device__ __forceinline__ f3 xform_vector(const float* m, f3 v) {
  return mk3(m[0] * v.x + m[1] * v.y + m[2] * v.z, m[4] * v.x + m[5] * v.y + m[6] * v.z,
             m[8] * v.x + m[9] * v.y + m[10] * v.z);
}
// transNormal with the inverse matrix `inv`
__device__ __forceinline__ f3 xform_normal(const float* inv, f3 n) {
  return mk3(inv[0] * n.x + inv[4] * n.y + inv[8] * n.z, inv[1] * n.x + inv[5] * n.y + inv[9] * n.z,
             inv[2] * n.x + inv[6] * n.y + inv[10] * n.z);
}

__device__ __forceinline__ bool solve_quadric(float a, float b, float c, float* t0o, float* t1o) {
  float discrim = b * b - 4.0f * a * c;
  if (discrim < 0.0f) return false;
  float rd = __builtin_sqrtf(discrim);
  float q = (b < 0.0f) ? -(0.5f * (b - rd)) : -(0.5f * (b + rd));
  float t0 = q / a, t1 = c / q;
  *t0o = hmin(t0, t1);
  *t1o = hmax(t0, t1);
  return true;
}

__device__ __forceinline__ float atan2p(float y, float x) {
  float a = atan2f(y, x);
  return a < 0.0f ? a + 6.2831854820251465f : a;  // twoPi :: Float
}

// Root box test with the reference's slab semantics (AABB.hs:79-94).
__device__ __forceinline__ bool slab_haskell(const float* lo, const float* hi, const DRay& r) {
  float nearT = r.tmin, farT = r.tmax;
#pragma unroll
  for (int dim = 0; dim < 3; ++dim) {
    if (nearT > farT) return false;
    float dInv = 1.0f / comp3(r.d, dim);
    float oc = comp3(r.o, dim);
    float tNear = (lo[dim] - oc) * dInv;
    float tFar = (hi[dim] - oc) * dInv;
    float n2 = tNear, f2 = tFar;
    if (tNear > tFar) {
      n2 = tFar;
      f2 = tNear;
    }
    nearT = hmax(nearT, n2);
    farT = hmin(farT, f2);
  }
  return !(nearT > farT);
}

// Moeller-Trumbore exactly as triangleIntersect (TriangleMesh.hs:160-207),
// with e1 = p2 - p1 and e2 = p3 - p1 precomputed on the host (same IEEE op).
__device__ __forceinline__ bool tri_test(f3 p1, f3 e1, f3 e2, f3 ro, f3 rd, float tmin, float tmax,
                                         float* to, float* b1o, float* b2o) {
  f3 s1 = cross3(rd, e2);
  float divisor = dot3(s1, e1);
  if (divisor == 0.0f) return false;
  float invDiv = 1.0f / divisor;
  f3 d = sub3(ro, p1);
  float b1 = dot3(d, s1) * invDiv;
  if (b1 < 0.0f || b1 > 1.0f) return false;
  f3 s2 = cross3(d, e1);
  float b2 = dot3(rd, s2) * invDiv;
  if (b2 < 0.0f || b1 + b2 > 1.0f) return false;
  float t = dot3(e2, s2) * invDiv;
  if (t < tmin || t > tmax) return false;
  *to = t;
  *b1o = b1;
  *b2o = b2;
  return true;
}

enum { SH_BOX = 0, SH_CYL = 1, SH_DISK = 2, SH_QUAD = 3, SH_SPHERE = 4 };

// Device shape record: 32 floats (128 B).
struct DShape {
  int32_t type;
  int32_t material;
  int32_t light;  // index into the device light list, -1 if none
  int32_t pad;
  float p[6];
  float pad2[2];
  float w2o[16];
};
struct DShapeX {  // shading-side data, indexed like DShape
  float o2w[16];
};

// The quad's intersect (Shape.hs:157-171): the z = 0 plane within
// |x| <= sx, |y| <= sy.
__device__ __forceinline__ bool quad_intersect(const DShape& s, const DRay& ray, float* to, int* axis) {
  float sx = s.p[0], sy = s.p[1];
  if (fabsf(ray.d.z) < 1e-7f) return false;
  float t = -ray.o.z / ray.d.z;
  if (t < ray.tmin || t > ray.tmax) return false;
  f3 p = ray_at(ray, t);
  if (fabsf(p.x) > sx || fabsf(p.y) > sy) return false;
  *to = t;
  *axis = 0;
  return true;
}

// The other shapes' intersect (Shape.hs:86-229) in object space; each
// returns t, and the box its slab axis.
__device__ __forceinline__ bool box_intersect(const DShape& s, const DRay& ray, float* to, int* axis) {
  float nearT = -INFINITY, farT = INFINITY;
  int dd = 0;
#pragma unroll
  for (int dim = 0; dim < 3; ++dim) {
    if (nearT > farT) return false;
    float dInv = 1.0f / comp3(ray.d, dim);
    float oc = comp3(ray.o, dim);
    float t1p = (s.p[3 + dim] - oc) * dInv;
    float t2p = (s.p[dim] - oc) * dInv;
    float t1 = t1p, t2 = t2p;
    if (t1p > t2p) {
      t1 = t2p;
      t2 = t1p;
    }
    int ndd = (nearT < t1) ? dim : dd;
    nearT = hmax(nearT, t1);
    farT = hmin(farT, t2);
    dd = ndd;
  }
  if (nearT > farT) return false;
  float t0 = hmin(nearT, farT);
  float t1 = hmax(nearT, farT);
  if (t0 > ray.tmax || t0 < ray.tmin) return false;
  float t = (t0 < ray.tmin) ? t1 : t0;
  if (t > ray.tmax) return false;
  *to = t;
  *axis = dd;
  return true;
}
__device__ __forceinline__ bool cyl_intersect(const DShape& s, const DRay& ray, float* to, int* axis) {
  float r = s.p[0], zmin = s.p[1], zmax = s.p[2], phimax = s.p[3];
  float a = ray.d.x * ray.d.x + ray.d.y * ray.d.y;
  float b = 2.0f * (ray.d.x * ray.o.x + ray.d.y * ray.o.y);
  float c = ray.o.x * ray.o.x + ray.o.y * ray.o.y - r * r;
  float t0, t1;
  if (!solve_quadric(a, b, c, &t0, &t1)) return false;
  if (t0 > ray.tmax) return false;
  if (t1 < ray.tmin) return false;
  f3 h0 = ray_at(ray, t0);
  if (t0 > ray.tmin && h0.z > zmin && h0.z < zmax && atan2p(h0.y, h0.x) <= phimax) {
    *to = t0;
    *axis = 0;
    return true;
  }
  f3 h1 = ray_at(ray, t1);
  if (t1 <= ray.tmax && h1.z > zmin && h1.z < zmax && atan2p(h1.y, h1.x) <= phimax) {
    *to = t1;
    *axis = 0;
    return true;
  }
  return false;
}
__device__ __forceinline__ bool disk_intersect(const DShape& s, const DRay& ray, float* to, int* axis) {
  float h = s.p[0], rad = s.p[1], irad = s.p[2], phimax = s.p[3];
  if (fabsf(ray.d.z) < 1e-7f) return false;
  float t = (h - ray.o.z) / ray.d.z;
  if (t < ray.tmin || t > ray.tmax) return false;
  f3 p = ray_at(ray, t);
  float d2 = p.x * p.x + p.y * p.y;
  if (d2 > rad * rad || d2 < irad * irad) return false;
  if (atan2p(p.y, p.x) > phimax) return false;
  *to = t;
  *axis = 0;
  return true;
}
__device__ __forceinline__ bool sphere_intersect(const DShape& s, const DRay& ray, float* to, int* axis) {
  float r = s.p[0];
  float a = sqlen3(ray.d);
  float b = 2.0f * dot3(ray.o, ray.d);
  float c = sqlen3(ray.o) - r * r;
  float t1, t2;
  if (!solve_quadric(a, b, c, &t1, &t2)) return false;
  if (t1 > ray.tmax) return false;
  if (t2 < ray.tmin) return false;
  float t = (t1 < ray.tmin) ? t2 : t1;
  if (t > ray.tmax) return false;
  *to = t;
  *axis = 0;
  return true;
}

// Shape.intersect in object space; returns t and the slab axis for boxes.
__device__ __forceinline__ bool shape_intersect(const DShape& s, const DRay& ray, float* to, int* axis) {
  switch (s.type) {
    case SH_BOX:
      return box_intersect(s, ray, to, axis);
    case SH_CYL:
      return cyl_intersect(s, ray, to, axis);
    case SH_DISK:
      return disk_intersect(s, ray, to, axis);
    case SH_QUAD:
      return quad_intersect(s, ray, to, axis);
    default:
      return sphere_intersect(s, ray, to, axis);
  }
}

// Shape kinds as a bit set (bit SH_x), for code compiled for a subset.
constexpr int kAllShapes = (1 << SH_BOX) | (1 << SH_CYL) | (1 << SH_DISK) | (1 << SH_QUAD) | (1 << SH_SPHERE);
template <int SK>
__device__ __forceinline__ constexpr bool one_shape_kind() {
  return SK != 0 && (SK & (SK - 1)) == 0;
}
// shape_intersect for shapes whose kind is in SK (the caller guarantees it):
// only those kinds' code is compiled, and no dispatch for a single kind.
template <int SK>
__device__ __forceinline__ bool shape_intersect_k(const DShape& s, const DRay& ray, float* to, int* axis) {
  constexpr bool one = one_shape_kind<SK>();
  if ((SK >> SH_QUAD & 1) && (one || s.type == SH_QUAD)) return quad_intersect(s, ray, to, axis);
  if ((SK >> SH_SPHERE & 1) && (one || s.type == SH_SPHERE)) return sphere_intersect(s, ray, to, axis);
  if ((SK >> SH_BOX & 1) && (one || s.type == SH_BOX)) return box_intersect(s, ray, to, axis);
  if ((SK >> SH_DISK & 1) && (one || s.type == SH_DISK)) return disk_intersect(s, ray, to, axis);
  if ((SK >> SH_CYL & 1) && (one || s.type == SH_CYL)) return cyl_intersect(s, ray, to, axis);
  return false;
}

// Shape.intersects (Shape.hs:231-284).
__device__ __forceinline__ bool shape_intersects(const DShape& s, const DRay& ray) {
  switch (s.type) {
    case SH_BOX:
      return slab_haskell(&s.p[0], &s.p[3], ray);
    case SH_CYL: {
      float r = s.p[0], zmin = s.p[1], zmax = s.p[2], phimax = s.p[3];
      float a = ray.d.x * ray.d.x + ray.d.y * ray.d.y;
      float b = 2.0f * (ray.d.x * ray.o.x + ray.d.y * ray.o.y);
      float c = ray.o.x * ray.o.x + ray.o.y * ray.o.y - r * r;
      float t0, t1;
      if (!solve_quadric(a, b, c, &t0, &t1)) return false;
      if (t0 > ray.tmax) return false;
      if (t1 < ray.tmin) return false;
      f3 h0 = ray_at(ray, t0);
      if (t0 > ray.tmin && h0.z > zmin && h0.z < zmax && atan2p(h0.y, h0.x) <= phimax) return true;
      f3 h1 = ray_at(ray, t1);
      return t1 < ray.tmax && h1.z > zmin && h1.z < zmax && atan2p(h1.y, h1.x) <= phimax &&
             t1 <= ray.tmax;
    }
    case SH_SPHERE: {
      float rad = s.p[0];
      float a = sqlen3(ray.d);
      float b = 2.0f * dot3(ray.o, ray.d);
      float c = sqlen3(ray.o) - rad * rad;
      float t0, t1;
      if (!solve_quadric(a, b, c, &t0, &t1)) return false;
      if (t0 > ray.tmax || t1 < ray.tmin) return false;
      if (t0 < ray.tmin) return t1 < ray.tmax;
      return true;
    }
    default: {  // disk and quad: same test as intersect
      float t;
      int ax;
      return shape_intersect(s, ray, &t, &ax);
    }
  }
}

}  // namespace bgpu
