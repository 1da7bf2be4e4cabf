// ORACLE -- test infrastructure only (see orc_math.h).
// Restatement of bling's ray/primitive intersection routines:
//   triangles  src/lib/Graphics/Bling/Primitive/TriangleMesh.hs:140-207
//   shapes     src/lib/Graphics/Bling/Shape.hs:81-284
//   mkGeom     src/lib/Graphics/Bling/Primitive/Geometry.hs:14-37
#pragma once
#include "orc_math.h"

namespace orc {

struct TriHit {
  float t, b1, b2;
};

// triangleIntersect (TriangleMesh.hs:160-207): the hit test part.
inline bool triangleIntersect(V3 p1, V3 p2, V3 p3, const Ray& r, TriHit* h) {
  V3 e1 = p2 - p1, e2 = p3 - p1;
  V3 s1 = cross(r.d, e2);
  float divisor = dot(s1, e1);
  if (divisor == 0.0f) return false;
  float invDiv = 1.0f / divisor;
  V3 d = r.o - p1;
  float b1 = dot(d, s1) * invDiv;
  if (b1 < 0.0f || b1 > 1.0f) return false;
  V3 s2 = cross(d, e1);
  float b2 = dot(r.d, s2) * invDiv;
  if (b2 < 0.0f || b1 + b2 > 1.0f) return false;
  float t = dot(e2, s2) * invDiv;
  if (t < r.tmin || t > r.tmax) return false;
  h->t = t;
  h->b1 = b1;
  h->b2 = b2;
  return true;
}

// triangleIntersects (TriangleMesh.hs:140-158): same arithmetic, Bool.
inline bool triangleIntersects(V3 p1, V3 p2, V3 p3, const Ray& r) {
  TriHit h;
  return triangleIntersect(p1, p2, p3, r, &h);
}

enum { SH_BOX = 0, SH_CYL = 1, SH_DISK = 2, SH_QUAD = 3, SH_SPHERE = 4 };

struct Shape {
  int type;
  float p[6];
};

// Result of Shape.intersect: distance, epsilon and which branch produced it
// (enough to rebuild the differential geometry).
struct ShapeHit {
  float t, eps;
  int axis;  // box: slab axis of the entry face
};

// Box intersect (Shape.hs:86-111).
inline bool boxIntersect(const Shape& s, const Ray& ray, ShapeHit* h) {
  V3 pmin = mk(s.p[0], s.p[1], s.p[2]), pmax = mk(s.p[3], s.p[4], s.p[5]);
  float nearT = -kInf, farT = kInf;
  int dd = 0;
  for (int dim = 0; dim < 3; ++dim) {
    if (nearT > farT) return false;
    float dInv = 1.0f / comp(ray.d, dim);
    float oc = comp(ray.o, dim);
    float t1p = (comp(pmax, dim) - oc) * dInv;
    float t2p = (comp(pmin, dim) - oc) * dInv;
    float t1 = t1p, t2 = t2p;
    if (t1p > t2p) { t1 = t2p; t2 = t1p; }
    int ndd = (nearT < t1) ? dim : dd;
    nearT = hmax(nearT, t1);
    farT = hmin(farT, t2);
    dd = ndd;
  }
  if (nearT > farT) return false;
  float t0 = hmin(nearT, farT);
  float t1 = hmax(nearT, farT);
  (void)t1;
  if (t0 > ray.tmax || t0 < ray.tmin) return false;
  float t = (t0 < ray.tmin) ? t1 : t0;
  if (t > ray.tmax) return false;
  h->t = t;
  h->eps = 5e-4f * t;
  h->axis = dd;
  return true;
}

// Cylinder intersect (Shape.hs:113-140).
inline bool cylinderIntersect(const Shape& s, const Ray& ray, ShapeHit* h) {
  float r = s.p[0], zmin = s.p[1], zmax = s.p[2], phimax = s.p[3];
  V3 ro = ray.o, rd = ray.d;
  float a = rd.x * rd.x + rd.y * rd.y;
  float b = 2.0f * (rd.x * ro.x + rd.y * ro.y);
  float c = ro.x * ro.x + ro.y * ro.y - r * r;
  float t0, t1;
  if (!solveQuadric(a, b, c, &t0, &t1)) return false;
  if (t0 > ray.tmax) return false;
  if (t1 < ray.tmin) return false;
  V3 ph0 = rayAt(ray, t0);
  float phi0 = atan2p(ph0.y, ph0.x);
  if (t0 > ray.tmin && ph0.z > zmin && ph0.z < zmax && phi0 <= phimax) {
    h->t = t0; h->eps = 5e-4f * t0; h->axis = 0;
    return true;
  }
  V3 ph1 = rayAt(ray, t1);
  float phi1 = atan2p(ph1.y, ph1.x);
  if (t1 <= ray.tmax && ph1.z > zmin && ph1.z < zmax && phi1 <= phimax) {
    h->t = t1; h->eps = 5e-4f * t1; h->axis = 0;
    return true;
  }
  return false;
}

// Disk intersect (Shape.hs:142-155).
inline bool diskIntersect(const Shape& s, const Ray& ray, ShapeHit* h) {
  float hgt = s.p[0], rad = s.p[1], irad = s.p[2], phimax = s.p[3];
  if (std::fabs(ray.d.z) < 1e-7f) return false;
  float t = (hgt - ray.o.z) / ray.d.z;
  if (t < ray.tmin || t > ray.tmax) return false;
  V3 p = rayAt(ray, t);
  float d2 = p.x * p.x + p.y * p.y;
  if (d2 > rad * rad || d2 < irad * irad) return false;
  float phi = atan2p(p.y, p.x);
  if (phi > phimax) return false;
  h->t = t; h->eps = 5e-4f * t; h->axis = 0;
  return true;
}

// Quad intersect (Shape.hs:157-171).
inline bool quadIntersect(const Shape& s, const Ray& ray, ShapeHit* h) {
  float sx = s.p[0], sy = s.p[1];
  if (std::fabs(ray.d.z) < 1e-7f) return false;
  float t = -ray.o.z / ray.d.z;
  if (t < ray.tmin || t > ray.tmax) return false;
  V3 p = rayAt(ray, t);
  if (std::fabs(p.x) > sx || std::fabs(p.y) > sy) return false;
  h->t = t; h->eps = 5e-4f * t; h->axis = 0;
  return true;
}

// Sphere intersect (Shape.hs:173-187), hit part.
inline bool sphereIntersect(const Shape& s, const Ray& ray, ShapeHit* h) {
  float r = s.p[0];
  float a = sqLen(ray.d);
  float b = 2.0f * dot(ray.o, ray.d);
  float c = sqLen(ray.o) - r * r;
  float t1, t2;
  if (!solveQuadric(a, b, c, &t1, &t2)) return false;
  if (t1 > ray.tmax) return false;
  if (t2 < ray.tmin) return false;
  float t = (t1 < ray.tmin) ? t2 : t1;
  if (t > ray.tmax) return false;
  h->t = t; h->eps = 5e-4f * t; h->axis = 0;
  return true;
}

inline bool shapeIntersect(const Shape& s, const Ray& r, ShapeHit* h) {
  switch (s.type) {
    case SH_BOX: return boxIntersect(s, r, h);
    case SH_CYL: return cylinderIntersect(s, r, h);
    case SH_DISK: return diskIntersect(s, r, h);
    case SH_QUAD: return quadIntersect(s, r, h);
    case SH_SPHERE: return sphereIntersect(s, r, h);
  }
  return false;
}

// Shape.intersects (Shape.hs:231-284): the Bool variants differ slightly.
inline bool shapeIntersects(const Shape& s, const Ray& ray) {
  switch (s.type) {
    case SH_BOX: {  // Shape.hs:233
      AABB b{mk(s.p[0], s.p[1], s.p[2]), mk(s.p[3], s.p[4], s.p[5])};
      return intersectAABB(b, ray, nullptr, nullptr);
    }
    case SH_CYL: {  // Shape.hs:235-251
      float r = s.p[0], zmin = s.p[1], zmax = s.p[2], phimax = s.p[3];
      V3 ro = ray.o, rd = ray.d;
      float a = rd.x * rd.x + rd.y * rd.y;
      float b = 2.0f * (rd.x * ro.x + rd.y * ro.y);
      float c = ro.x * ro.x + ro.y * ro.y - r * r;
      float t0, t1;
      if (!solveQuadric(a, b, c, &t0, &t1)) return false;
      if (t0 > ray.tmax) return false;
      if (t1 < ray.tmin) return false;
      V3 ph0 = rayAt(ray, t0);
      float phi0 = atan2p(ph0.y, ph0.x);
      if (t0 > ray.tmin && ph0.z > zmin && ph0.z < zmax && phi0 <= phimax) return true;
      V3 ph1 = rayAt(ray, t1);
      float phi1 = atan2p(ph1.y, ph1.x);
      return t1 < ray.tmax && ph1.z > zmin && ph1.z < zmax && phi1 <= phimax && t1 <= ray.tmax;
    }
    case SH_DISK: {  // Shape.hs:253-264
      ShapeHit h;
      return diskIntersect(s, ray, &h);
    }
    case SH_QUAD: {  // Shape.hs:266-273
      ShapeHit h;
      return quadIntersect(s, ray, &h);
    }
    case SH_SPHERE: {  // Shape.hs:275-284
      float rad = s.p[0];
      float a = sqLen(ray.d);
      float b = 2.0f * dot(ray.o, ray.d);
      float c = sqLen(ray.o) - rad * rad;
      float t0, t1;
      if (!solveQuadric(a, b, c, &t0, &t1)) return false;
      if (t0 > ray.tmax || t1 < ray.tmin) return false;
      if (t0 < ray.tmin) return t1 < ray.tmax;
      return true;
    }
  }
  return false;
}

// objectBounds (Shape.hs:295-311)
inline AABB objectBounds(const Shape& s) {
  switch (s.type) {
    case SH_BOX: return AABB{mk(s.p[0], s.p[1], s.p[2]), mk(s.p[3], s.p[4], s.p[5])};
    case SH_CYL: return AABB{mk(-s.p[0], -s.p[0], s.p[1]), mk(s.p[0], s.p[0], s.p[2])};
    case SH_DISK: return AABB{mk(-s.p[1], -s.p[1], s.p[0]), mk(s.p[1], s.p[1], s.p[0])};
    case SH_QUAD: return AABB{mk(-s.p[0], -s.p[1], 0.0f), mk(s.p[0], s.p[1], 0.0f)};
    default: return AABB{mk(-s.p[0], -s.p[0], -s.p[0]), mk(s.p[0], s.p[0], s.p[0])};
  }
}

// transBox (Transform.hs:281-296)
inline AABB transBox(const M4& t, const AABB& b) {
  float mx = b.lo.x, my = b.lo.y, mz = b.lo.z, nx = b.hi.x, ny = b.hi.y, nz = b.hi.z;
  V3 ps[8] = {mk(mx, my, mz), mk(mx, my, nz), mk(mx, ny, mz), mk(mx, ny, nz),
              mk(nx, my, mz), mk(nx, my, nz), mk(nx, ny, mz), mk(nx, ny, nz)};
  AABB r = emptyAABB();
  for (int i = 0; i < 8; ++i) r = extendP(r, transPoint(t, ps[i]));
  return r;
}

// triangleBounds (TriangleMesh.hs:136-138)
inline AABB triangleBounds(V3 p1, V3 p2, V3 p3) {
  return extendP(extendP(extendP(emptyAABB(), p1), p2), p3);
}

}  // namespace orc
