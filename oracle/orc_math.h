// ORACLE -- test infrastructure only.  CPU restatement of bling's FP32 math
// (src/lib/Graphics/Bling/Math.hs, AABB.hs, Transform.hs), used by tests/,
// the smoke check and the bench's CPU baseline as the checker.  Never linked
// into the product library.
//
// Every function keeps the reference's operation order so that, compiled with
// -ffp-contract=off and IEEE binary32, results are bit-identical to GHC's
// Float arithmetic (SSE scalar, no FMA contraction).
#pragma once
#include <cmath>
#include <cstdint>
#include <limits>

namespace orc {

static const float kInf = std::numeric_limits<float>::infinity();
static const float kPi = 3.14159265358979323846f;      // `pi :: Float`
static const float kTwoPi = 2.0f * kPi;                  // Math.hs:57-59
static const float kInvPi = 1.0f / kPi;                  // Math.hs:49-51
static const float kInvTwoPi = 1.0f / (2.0f * kPi);      // Math.hs:53-55

// Haskell's default Ord methods for Float: max x y = if x <= y then y else x.
inline float hmax(float x, float y) { return (x <= y) ? y : x; }
inline float hmin(float x, float y) { return (x <= y) ? x : y; }
// clamp (Math.hs:78-87)
inline float clampf(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }
// lerp (Math.hs:108-110)
inline float lerpf(float t, float a, float b) { return (1.0f - t) * a + t * b; }

struct V3 {
  float x, y, z;
};
inline V3 mk(float x, float y, float z) { return V3{x, y, z}; }
inline V3 operator+(V3 a, V3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 operator-(V3 a, V3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 operator*(V3 a, V3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
inline V3 operator-(V3 a) { return mk(-a.x, -a.y, -a.z); }
// v * vpromote f, and f *# v = vpromote f * v (Math.hs:207-209)
inline V3 scale(V3 v, float f) { return mk(v.x * f, v.y * f, v.z * f); }
inline V3 sscale(float f, V3 v) { return mk(f * v.x, f * v.y, f * v.z); }
// Math.hs:328-353
inline float sqLen(V3 v) { return v.x * v.x + v.y * v.y + v.z * v.z; }
inline float len(V3 v) { return std::sqrt(sqLen(v)); }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float absDot(V3 a, V3 b) { return std::fabs(dot(a, b)); }
inline V3 cross(V3 u, V3 v) {
  return mk(u.y * v.z - u.z * v.y, -(u.x * v.z - u.z * v.x), u.x * v.y - u.y * v.x);
}
inline V3 normalize(V3 v) {
  if (sqLen(v) != 0.0f) return scale(v, 1.0f / len(v));
  return mk(0.0f, 1.0f, 0.0f);
}
inline float comp(V3 v, int d) { return d == 0 ? v.x : (d == 1 ? v.y : v.z); }
inline V3 setComp(int d, float t, V3 v) {
  if (d == 0) return mk(t, v.y, v.z);
  if (d == 1) return mk(v.x, t, v.z);
  return mk(v.x, v.y, t);
}
inline V3 faceForward(V3 v, V3 v2) { return dot(v, v2) < 0.0f ? -v : v; }

struct Ray {
  V3 o, d;
  float tmin, tmax;
};
inline V3 rayAt(const Ray& r, float t) { return r.o + scale(r.d, t); }  // Math.hs:388-390

// solveQuadric (Math.hs:124-139)
inline bool solveQuadric(float a, float b, float c, float* t0out, float* t1out) {
  float discrim = b * b - 4.0f * a * c;
  if (discrim < 0.0f) return false;
  float rootDiscrim = std::sqrt(discrim);
  float q = (b < 0.0f) ? -(0.5f * (b - rootDiscrim)) : -(0.5f * (b + rootDiscrim));
  float t0 = q / a, t1 = c / q;
  *t0out = hmin(t0, t1);
  *t1out = hmax(t0, t1);
  return true;
}

// atan2' (Math.hs:69-75): atan2 remapped to [0, 2pi]
inline float atan2p(float y, float x) {
  float a = std::atan2(y, x);
  return a < 0.0f ? a + kTwoPi : a;
}

// LocalCoordinates and coordinateSystem (Math.hs:408-449)
struct Frame {
  V3 s, t, n;
};
inline Frame coordinateSystem(V3 v) {
  if (std::fabs(v.x) > std::fabs(v.y)) {
    float invLen = 1.0f / std::sqrt(v.x * v.x + v.z * v.z);
    V3 v2 = mk(-v.z * invLen, 0.0f, v.x * invLen);
    return Frame{v2, cross(v, v2), v};
  }
  float invLen = 1.0f / std::sqrt(v.y * v.y + v.z * v.z);
  V3 v2 = mk(0.0f, v.z * invLen, -v.y * invLen);
  return Frame{v2, cross(v, v2), v};
}
inline V3 worldToLocal(const Frame& f, V3 v) { return mk(dot(v, f.s), dot(v, f.t), dot(v, f.n)); }
inline V3 localToWorld(const Frame& f, V3 v) {
  return mk(f.s.x * v.x + f.t.x * v.y + f.n.x * v.z,
            f.s.y * v.x + f.t.y * v.y + f.n.y * v.z,
            f.s.z * v.x + f.t.z * v.y + f.n.z * v.z);
}

// Row-major 4x4 matrix of a Transform (Transform.hs:123-136,218-221).
struct M4 {
  float m[16];
  float at(int r, int c) const { return m[r * 4 + c]; }
};
// transPoint (Transform.hs:247-258)
inline V3 transPoint(const M4& M, V3 p) {
  float xp = M.at(0, 0) * p.x + M.at(0, 1) * p.y + M.at(0, 2) * p.z + M.at(0, 3);
  float yp = M.at(1, 0) * p.x + M.at(1, 1) * p.y + M.at(1, 2) * p.z + M.at(1, 3);
  float zp = M.at(2, 0) * p.x + M.at(2, 1) * p.y + M.at(2, 2) * p.z + M.at(2, 3);
  float wp = M.at(3, 0) * p.x + M.at(3, 1) * p.y + M.at(3, 2) * p.z + M.at(3, 3);
  if (wp == 1.0f) return mk(xp, yp, zp);
  return mk(xp / wp, yp / wp, zp / wp);
}
// transVector (Transform.hs:259-266)
inline V3 transVector(const M4& M, V3 v) {
  return mk(M.at(0, 0) * v.x + M.at(0, 1) * v.y + M.at(0, 2) * v.z,
            M.at(1, 0) * v.x + M.at(1, 1) * v.y + M.at(1, 2) * v.z,
            M.at(2, 0) * v.x + M.at(2, 1) * v.y + M.at(2, 2) * v.z);
}
// transNormal (Transform.hs:267-274): uses the *inverse* matrix, transposed.
inline V3 transNormal(const M4& Inv, V3 n) {
  return mk(Inv.at(0, 0) * n.x + Inv.at(1, 0) * n.y + Inv.at(2, 0) * n.z,
            Inv.at(0, 1) * n.x + Inv.at(1, 1) * n.y + Inv.at(2, 1) * n.z,
            Inv.at(0, 2) * n.x + Inv.at(1, 2) * n.y + Inv.at(2, 2) * n.z);
}
inline Ray transRay(const M4& M, const Ray& r) {  // Transform.hs:275-280
  return Ray{transPoint(M, r.o), transVector(M, r.d), r.tmin, r.tmax};
}

struct AABB {
  V3 lo, hi;
};
inline AABB emptyAABB() { return AABB{mk(kInf, kInf, kInf), mk(-kInf, -kInf, -kInf)}; }
// extendAABBP (AABB.hs:43-46) with Haskell min/max
inline AABB extendP(AABB b, V3 p) {
  return AABB{mk(hmin(b.lo.x, p.x), hmin(b.lo.y, p.y), hmin(b.lo.z, p.z)),
              mk(hmax(b.hi.x, p.x), hmax(b.hi.y, p.y), hmax(b.hi.z, p.z))};
}
// extendAABB (AABB.hs:34-41): mappend
inline AABB extend(AABB a, AABB b) {
  return AABB{mk(hmin(a.lo.x, b.lo.x), hmin(a.lo.y, b.lo.y), hmin(a.lo.z, b.lo.z)),
              mk(hmax(a.hi.x, b.hi.x), hmax(a.hi.y, b.hi.y), hmax(a.hi.z, b.hi.z))};
}
// intersectAABB (AABB.hs:79-94)
inline bool intersectAABB(const AABB& b, const Ray& r, float* t0, float* t1) {
  float nearT = r.tmin, farT = r.tmax;
  for (int dim = 0; dim < 3; ++dim) {
    if (nearT > farT) return false;
    float dInv = 1.0f / comp(r.d, dim);
    float oc = comp(r.o, dim);
    float tNear = (comp(b.lo, dim) - oc) * dInv;
    float tFar = (comp(b.hi, dim) - oc) * dInv;
    float n2 = tNear, f2 = tFar;
    if (tNear > tFar) { n2 = tFar; f2 = tNear; }
    nearT = hmax(nearT, n2);
    farT = hmin(farT, f2);
  }
  if (nearT > farT) return false;
  if (t0) *t0 = nearT;
  if (t1) *t1 = farT;
  return true;
}

}  // namespace orc
