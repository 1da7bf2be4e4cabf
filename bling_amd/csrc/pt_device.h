// Device shading for the path tracer (gfx950).  Restates, per path vertex:
//   Integrator/Path.hs:41-87, Scene.hs:61-118 (NEE with MIS), Reflection.hs:201-335,
//   Reflection/{Diffuse,Specular,Microfacet}.hs, Fresnel.hs:31-69, Material.hs:32-96,
//   Light.hs:72-229, Shape.hs:314-410, Montecarlo.hs:40-186, Camera.hs:49-68,
//   DifferentialGeometry.hs:41-81, Primitive/TriangleMesh.hs:119-207.
// Spectra are 16 binary32 bands in registers; material and light spectra are
// read from HBM where they are used.
#pragma once
#include "../../include/bgpu.h"
#include "pt.h"
#include "pt_sky.h"
#include "spectral_tables.h"

namespace bgpu {
namespace sh {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kTwoPi = 2.0f * kPi;
constexpr float kInvPi = 1.0f / kPi;
constexpr float kInvTwoPi = 1.0f / (2.0f * kPi);

__constant__ float cCieX[16], cCieY[16], cCieZ[16];

// ------------------------------------------------------------- sampler
// Counter-based stratified sampler (spec shared with the oracle): splitmix64
// finaliser, keyed cycle-walking permutation standing in for `shuffle`.
__device__ __forceinline__ uint64_t mix64(uint64_t x) {
  x ^= x >> 30;
  x *= 0xbf58476d1ce4e5b9ULL;
  x ^= x >> 27;
  x *= 0x94d049bb133111ebULL;
  x ^= x >> 31;
  return x;
}
__device__ __forceinline__ uint64_t pixel_key(uint64_t seed, uint32_t pass, int ix, int iy) {
  uint64_t c = ((uint64_t)pass << 42) | ((uint64_t)(uint32_t)(iy + 1024) << 21) |
               (uint64_t)(uint32_t)(ix + 1024);
  return mix64(mix64(seed) ^ mix64(c + 0x9e3779b97f4a7c15ULL));
}
__device__ __forceinline__ uint64_t rng_bits(uint64_t key, uint32_t dim, uint32_t idx) {
  return mix64(key + 0x9e3779b97f4a7c15ULL * (((uint64_t)dim << 24) + (uint64_t)idx + 1ULL));
}
__device__ __forceinline__ float uni(uint64_t b) { return (float)(uint32_t)(b >> 40) * 0x1p-24f; }
__device__ __forceinline__ uint32_t permute(uint32_t i, uint32_t n, uint32_t p) {
  if (n <= 1) return 0;
  uint32_t w = n - 1;
  w |= w >> 1; w |= w >> 2; w |= w >> 4; w |= w >> 8; w |= w >> 16;
  do {
    i = (i ^ p) & w;
    i = (i * 0x9e3779b1u) & w;
    i ^= i >> 5;
    i ^= (i << 7) & w;
    i = (i * 0x85ebca77u) & w;
    i ^= (p >> 16) & w;
    i ^= i >> 9;
    i = (i * 0xc2b2ae3du) & w;
    i ^= i >> 11;
  } while (i >= n);
  return i;
}
enum : uint32_t {
  D_CAM_X = 1, D_CAM_Y = 2, D_LENS_X = 3, D_LENS_Y = 4, D_LENS_P = 5,
  D_T1_J = 0x1000, D_T1_P = 0x2000, D_T2_JX = 0x3000, D_T2_JY = 0x4000, D_T2_P = 0x5000,
  D_R1 = 0x6000, D_R2X = 0x7000, D_R2Y = 0x8000
};
constexpr float kAlmostOne = 0x1.fffffep-1f;

struct Sampler {
  uint64_t key;
  uint32_t nu, nv, N, n1d, n2d;
  __device__ __forceinline__ void strat2(uint32_t i, float ju, float jv, float* u, float* v) const {
    float du = 1.0f / (float)nu, dv = 1.0f / (float)nv;
    float a = ((float)(i / nu) + ju) * du, b = ((float)(i % nu) + jv) * dv;
    *u = a < kAlmostOne ? a : kAlmostOne;
    *v = b < kAlmostOne ? b : kAlmostOne;
  }
  __device__ __forceinline__ void camera(uint32_t n, float* ox, float* oy) const {
    strat2(n, uni(rng_bits(key, D_CAM_X, n)), uni(rng_bits(key, D_CAM_Y, n)), ox, oy);
  }
  __device__ __forceinline__ void lens(uint32_t n, float* u, float* v) const {
    uint32_t i = permute(n, N, (uint32_t)rng_bits(key, D_LENS_P, 0));
    strat2(i, uni(rng_bits(key, D_LENS_X, i)), uni(rng_bits(key, D_LENS_Y, i)), u, v);
  }
  __device__ __forceinline__ float get1(uint32_t n, uint32_t k) const {
    if (k < n1d) {
      uint32_t i = permute(n, N, (uint32_t)rng_bits(key, D_T1_P + k, 0));
      float du = 1.0f / (float)N;
      float a = ((float)i + uni(rng_bits(key, D_T1_J + k, i))) * du;
      return a < kAlmostOne ? a : kAlmostOne;
    }
    return uni(rng_bits(key, D_R1 + k, n));
  }
  __device__ __forceinline__ void get2(uint32_t n, uint32_t k, float* u, float* v) const {
    if (k < n2d) {
      uint32_t i = permute(n, N, (uint32_t)rng_bits(key, D_T2_P + k, 0));
      strat2(i, uni(rng_bits(key, D_T2_JX + k, i)), uni(rng_bits(key, D_T2_JY + k, i)), u, v);
      return;
    }
    *u = uni(rng_bits(key, D_R2X + k, n));
    *v = uni(rng_bits(key, D_R2Y + k, n));
  }
};

// ------------------------------------------------------------- spectra
struct Sp {
  float v[16];
};
__device__ __forceinline__ Sp sp_const(float c) {
  Sp s;
#pragma unroll
  for (int i = 0; i < 16; ++i) s.v[i] = c;
  return s;
}
__device__ __forceinline__ Sp sp_load(const float* p) {
  Sp s;
#pragma unroll
  for (int i = 0; i < 16; ++i) s.v[i] = p[i];
  return s;
}
__device__ __forceinline__ bool sp_black(const Sp& s) {
  bool b = true;
#pragma unroll
  for (int i = 0; i < 16; ++i) b = b && (s.v[i] == 0.0f);
  return b;
}
__device__ __forceinline__ float sp_y(const Sp& s) {  // sY (Spectrum.hs:371-373)
  float a = 0.0f;
#pragma unroll
  for (int i = 0; i < 16; ++i) a = a + s.v[i] * cCieY[i];
  return a / BGPU_CIE_Y_SUM;
}

// ------------------------------------------------------------- helpers
__device__ __forceinline__ float clampf_(float v, float lo, float hi) {
  return v < lo ? lo : (v > hi ? hi : v);
}
__device__ __forceinline__ float lerpf_(float t, float a, float b) { return (1.0f - t) * a + t * b; }
__device__ __forceinline__ f3 sscale3(float f, f3 v) { return mk3(f * v.x, f * v.y, f * v.z); }
struct Frame {
  f3 s, t, n;
};
__device__ __forceinline__ Frame coord_system(f3 v) {  // Math.hs:413-425
  if (fabsf(v.x) > fabsf(v.y)) {
    float il = 1.0f / __builtin_sqrtf(v.x * v.x + v.z * v.z);
    f3 v2 = mk3(-v.z * il, 0.0f, v.x * il);
    return Frame{v2, cross3(v, v2), v};
  }
  float il = 1.0f / __builtin_sqrtf(v.y * v.y + v.z * v.z);
  f3 v2 = mk3(0.0f, v.z * il, -v.y * il);
  return Frame{v2, cross3(v, v2), v};
}
__device__ __forceinline__ f3 to_local(const Frame& f, f3 v) {
  return mk3(dot3(v, f.s), dot3(v, f.t), dot3(v, f.n));
}
__device__ __forceinline__ f3 to_world(const Frame& f, f3 v) {
  return mk3(f.s.x * v.x + f.t.x * v.y + f.n.x * v.z, f.s.y * v.x + f.t.y * v.y + f.n.y * v.z,
             f.s.z * v.x + f.t.z * v.y + f.n.z * v.z);
}
__device__ __forceinline__ f3 set_comp(int d, float t, f3 v) {
  if (d == 0) return mk3(t, v.y, v.z);
  if (d == 1) return mk3(v.x, t, v.z);
  return mk3(v.x, v.y, t);
}

// ------------------------------------------------------- differential geom
struct DG {
  f3 p, n, dpdu, dpdv;
  float b1, b2;
  float u = 0.0f, v = 0.0f;  // surface parameters (0 where mkDg' builds the DG)
};
__device__ __forceinline__ DG mk_dg(f3 p, f3 dpdu, f3 dpdv) {
  return DG{p, norm3(cross3(dpdu, dpdv)), dpdu, dpdv, 0.0f, 0.0f};
}
__device__ __forceinline__ DG mk_dg1(f3 p, f3 n) {
  Frame f = coord_system(n);
  return DG{p, n, f.s, f.t, 0.0f, 0.0f};
}

// The quad's DG at p (Shape.hs:157-171).
__device__ __forceinline__ DG quad_dg(const DShape& s, f3 p) {
  DG d = mk_dg(p, mk3(s.p[0], 0.0f, 0.0f), mk3(0.0f, s.p[1], 0.0f));
  d.u = (s.p[0] + p.x) / (2.0f * s.p[0]);
  d.v = (s.p[1] + p.y) / (2.0f * s.p[1]);
  return d;
}

// The other shapes' DG at p (Shape.hs:86-229); u/v as the sphere sets
// them, 0 for the shapes built with mkDg'.
__device__ __forceinline__ DG box_dg(const DShape& s, f3 p, int axis) {
  float half = (s.p[axis] + s.p[3 + axis]) / 2.0f;
  float dir = comp3(p, axis) > half ? 1.0f : -1.0f;
  return mk_dg1(p, norm3(set_comp(axis, dir, mk3(0, 0, 0))));
}
__device__ __forceinline__ DG cyl_dg(const DShape& s, f3 p) {
  float phimax = s.p[3];
  f3 dpdu = mk3(-(phimax * p.y), phimax * p.x, 0.0f);
  f3 dpdv = mk3(0.0f, 0.0f, s.p[2] - s.p[1]);
  return mk_dg1(p, norm3(cross3(dpdu, dpdv)));
}
__device__ __forceinline__ DG disk_dg(f3 p) { return mk_dg1(p, mk3(0.0f, 0.0f, -1.0f)); }
__device__ __forceinline__ DG sphere_dg(const DShape& s, f3 p) {
  float r = s.p[0];
  const float thetaMin = kPi, thetaMax = 0.0f, phiMax = kTwoPi;
  float theta = acosf(clampf_(p.z / r, -1.0f, 1.0f));
  float zr = __builtin_sqrtf(p.x * p.x + p.y * p.y);
  float iz = 1.0f / zr;
  float cosphi = p.x * iz, sinphi = p.y * iz;
  f3 dpdu = mk3(-(phiMax * p.y), phiMax * p.x, 0.0f);
  f3 dpdv = scl3(mk3(p.z * cosphi, p.z * sinphi, -(r * sinf(theta))), thetaMax - thetaMin);
  DG d = mk_dg(p, dpdu, dpdv);
  d.u = atan2p(p.y, p.x) / phiMax;  // Shape.hs:189-192
  d.v = (theta - thetaMin) / (thetaMax - thetaMin);
  return d;
}

// Object-space DG of a shape hit (Shape.hs:86-229).
__device__ __forceinline__ DG shape_dg(const DShape& s, const DRay& ray, float t, int axis) {
  f3 p = ray_at(ray, t);
  switch (s.type) {
    case SH_BOX:
      return box_dg(s, p, axis);
    case SH_CYL:
      return cyl_dg(s, p);
    case SH_DISK:
      return disk_dg(p);
    case SH_QUAD:
      return quad_dg(s, p);
    default:
      return sphere_dg(s, p);
  }
}
// shape_dg for shapes whose kind is in SK (see shape_intersect_k).
template <int SK>
__device__ __forceinline__ DG shape_dg_k(const DShape& s, const DRay& ray, float t, int axis) {
  constexpr bool one = one_shape_kind<SK>();
  f3 p = ray_at(ray, t);
  if ((SK >> SH_QUAD & 1) && (one || s.type == SH_QUAD)) return quad_dg(s, p);
  if ((SK >> SH_SPHERE & 1) && (one || s.type == SH_SPHERE)) return sphere_dg(s, p);
  if ((SK >> SH_BOX & 1) && (one || s.type == SH_BOX)) return box_dg(s, p, axis);
  if ((SK >> SH_DISK & 1) && (one || s.type == SH_DISK)) return disk_dg(p);
  return cyl_dg(s, p);
}

// Geometric (dgg) and shading (dgs) DG at a hit, world space.
__device__ __forceinline__ void hit_dg(const DevPtScene& S, const DRay& ray, float t, float b1, float b2,
                                       int prim, DG* dgg, DG* dgs, float* eps, int* material,
                                       int* light) {
  if (prim < S.ntri) {
    int i0 = S.tri_idx[3 * prim], i1 = S.tri_idx[3 * prim + 1], i2 = S.tri_idx[3 * prim + 2];
    f3 p1 = mk3(S.verts[3 * i0], S.verts[3 * i0 + 1], S.verts[3 * i0 + 2]);
    f3 p2 = mk3(S.verts[3 * i1], S.verts[3 * i1 + 1], S.verts[3 * i1 + 2]);
    f3 p3 = mk3(S.verts[3 * i2], S.verts[3 * i2 + 1], S.verts[3 * i2 + 2]);
    int2 info = S.tri_info[prim];
    float uv00 = 0, uv01 = 0, uv10 = 1, uv11 = 0, uv20 = 1, uv21 = 1;
    if ((info.y & BGPU_MESH_UVS) && S.uvs) {
      uv00 = S.uvs[2 * i0]; uv01 = S.uvs[2 * i0 + 1];
      uv10 = S.uvs[2 * i1]; uv11 = S.uvs[2 * i1 + 1];
      uv20 = S.uvs[2 * i2]; uv21 = S.uvs[2 * i2 + 1];
    }
    f3 e1 = sub3(p2, p1), e2 = sub3(p3, p1);
    float du1 = uv00 - uv20, du2 = uv10 - uv20, dv1 = uv01 - uv21, dv2 = uv11 - uv21;
    f3 dp1 = sub3(p1, p3), dp2 = sub3(p2, p3);
    float det = du1 * dv2 - dv1 * du2;
    f3 dpdu, dpdv;
    if (det == 0.0f) {
      Frame f = coord_system(norm3(cross3(e1, e2)));
      dpdu = f.s;
      dpdv = f.t;
    } else {
      float id = 1.0f / det;
      dpdu = sscale3(id, sub3(sscale3(dv2, dp1), sscale3(dv1, dp2)));
      dpdv = sscale3(id, add3(sscale3(-du2, dp1), sscale3(du1, dp2)));
    }
    *dgg = mk_dg(ray_at(ray, t), dpdu, dpdv);
    dgg->b1 = b1;
    dgg->b2 = b2;
    {  // interpolated u, v (TriangleMesh.hs:195-198)
      const float b0 = 1.0f - b1 - b2;
      dgg->u = b0 * uv00 + b1 * uv10 + b2 * uv20;
      dgg->v = b0 * uv01 + b1 * uv11 + b2 * uv21;
    }
    *dgs = *dgg;
    if ((info.y & BGPU_MESH_NORMALS) && S.normals) {  // TriangleMesh.hs:122-134
      f3 n0 = mk3(S.normals[3 * i0], S.normals[3 * i0 + 1], S.normals[3 * i0 + 2]);
      f3 n1 = mk3(S.normals[3 * i1], S.normals[3 * i1 + 1], S.normals[3 * i1 + 2]);
      f3 n2 = mk3(S.normals[3 * i2], S.normals[3 * i2 + 1], S.normals[3 * i2 + 2]);
      float b0 = 1.0f - b1 - b2;
      f3 nsp = add3(add3(sscale3(b0, n0), sscale3(b1, n1)), sscale3(b2, n2));
      // transNormal by the identity: (1*x + 0*y) + 0*z per component
      nsp = mk3(1.0f * nsp.x + 0.0f * nsp.y + 0.0f * nsp.z, 0.0f * nsp.x + 1.0f * nsp.y + 0.0f * nsp.z,
                0.0f * nsp.x + 0.0f * nsp.y + 1.0f * nsp.z);
      f3 ns = norm3(nsp);
      f3 ss0 = norm3(dgg->dpdu);
      f3 ts0 = cross3(ss0, ns);
      dgs->n = ns;
      if (sqlen3(ts0) > 0.0f) {
        dgs->dpdu = cross3(norm3(ts0), ns);
        dgs->dpdv = norm3(ts0);
      } else {
        Frame f = coord_system(ns);
        dgs->dpdu = f.s;
        dgs->dpdv = f.t;
      }
    }
    *eps = 1e-3f * t;
    *material = info.x;
    *light = -1;
    return;
  }
  int si = prim - S.ntri;
  const DShape& s = S.geo.shapes[si];
  const float* o2w = S.shapex[si].o2w;
  DRay ro;
  ro.o = xform_point(s.w2o, ray.o);
  ro.d = xform_vector(s.w2o, ray.d);
  ro.tmin = ray.tmin;
  ro.tmax = ray.tmax;
  int axis = 0;
  if (s.type == SH_BOX) {  // recover the entry face axis (Shape.hs:100-111)
    float nearT = -INFINITY, farT = INFINITY;
    int dd = 0;
#pragma unroll
    for (int dim = 0; dim < 3; ++dim) {
      if (!(nearT > farT)) {
        float dInv = 1.0f / comp3(ro.d, dim);
        float oc = comp3(ro.o, dim);
        float t1p = (s.p[3 + dim] - oc) * dInv, t2p = (s.p[dim] - oc) * dInv;
        float ta = t1p > t2p ? t2p : t1p, tb = t1p > t2p ? t1p : t2p;
        int ndd = (nearT < ta) ? dim : dd;
        nearT = hmax(nearT, ta);
        farT = hmin(farT, tb);
        dd = ndd;
      }
    }
    axis = dd;
  }
  DG d = shape_dg(s, ro, t, axis);
  // transDg o2w (DifferentialGeometry.hs:72-81); the normal uses the inverse (w2o)
  dgg->p = xform_point(o2w, d.p);
  dgg->n = norm3(xform_normal(s.w2o, d.n));
  dgg->dpdu = xform_vector(o2w, d.dpdu);
  dgg->dpdv = xform_vector(o2w, d.dpdv);
  dgg->b1 = dgg->b2 = 0.0f;
  dgg->u = d.u;
  dgg->v = d.v;
  *dgs = *dgg;
  *eps = 5e-4f * t;
  *material = s.material;
  *light = s.light;
}

// ------------------------------------------------------------- warps
__device__ __forceinline__ float power_heuristic(float fp, float gp) {
  float f = 1.0f * fp, g = 1.0f * gp;
  return (f * f) / (f * f + g * g);
}
__device__ __forceinline__ void concentric_disk(float u1, float u2, float* ox, float* oy) {
  float sx = u1 * 2.0f - 1.0f, sy = u2 * 2.0f - 1.0f;
  if (sx == 0.0f && sy == 0.0f) {
    *ox = 0.0f;
    *oy = 0.0f;
    return;
  }
  float r, th;
  if (sx >= -sy) {
    if (sx > sy) {
      r = sx;
      th = sy > 0.0f ? sy / sx : 8.0f + sy / sx;
    } else {
      r = sy;
      th = 2.0f - sx / sy;
    }
  } else if (sx <= sy) {
    r = -sx;
    th = 4.0f - sy / (-sx);
  } else {
    r = -sy;
    th = 6.0f + sx / (-sy);
  }
  float theta = th * kPi / 4.0f;
  *ox = r * cosf(theta);
  *oy = r * sinf(theta);
}
__device__ __forceinline__ f3 cos_hemisphere(float u1, float u2) {
  float x, y;
  concentric_disk(u1, u2, &x, &y);
  return mk3(x, y, __builtin_sqrtf(hmax(0.0f, 1.0f - x * x - y * y)));
}
__device__ __forceinline__ f3 uniform_sphere(float u1, float u2) {
  float u = u1 * 2.0f - 1.0f;
  float s = __builtin_sqrtf(1.0f - u * u);
  float om = u2 * 2.0f * kPi;
  return mk3(s * cosf(om), s * sinf(om), u);
}

// ---------------------------------------------------------------- BSDF
struct Bsdf {
  const DMaterial* m;
  Frame cs;
  f3 p, ng;
};
__device__ __forceinline__ float cos_t(f3 v) { return v.z; }
__device__ __forceinline__ float acos_t(f3 v) { return fabsf(v.z); }
__device__ __forceinline__ float sin_t2(f3 v) { return hmax(0.0f, 1.0f - v.z * v.z); }
__device__ __forceinline__ float sin_t(f3 v) { return __builtin_sqrtf(sin_t2(v)); }
__device__ __forceinline__ bool same_hemi(f3 a, f3 b) { return a.z * b.z > 0.0f; }

__device__ __forceinline__ float fr_diel_scalar(float etai, float etat, float cosi) {
  float c = hmax(0.0f, 1.0f - cosi * cosi);
  float costp = cosi > 0.0f ? c / (etat * etat) : c * (etat * etat);
  float cost = __builtin_sqrtf(1.0f - clampf_(costp, 0.0f, 1.0f));
  float ac = fabsf(cosi);
  float eta = etat / etai;
  float rpl = eta * ac;
  float rParl = (cost - rpl) / (cost + rpl);
  float rpp = eta * cost;
  float rPerp = (ac - rpp) / (ac + rpp);
  return (rParl * rParl + rPerp * rPerp) * 0.5f;
}
// Spectral BSDF values are carried as at most two scalar-weighted terms and
// expanded one band at a time where they are consumed, so that no 16-band
// temporaries stay live across the shading code.  Each term reproduces the
// per-band arithmetic of the reference expression it stands for.
enum {
  T_RS = 0,   // r[b] * a                      (Lambert eval r*s, Lambert/Oren sample)
  T_ROS = 1,  // (r[b] * a) * c                (Oren-Nayar eval (r*o)*s)
  T_RFX = 2,  // (r[b] * c) * a                (r times a scalar Fresnel c, then scaled)
  T_RCX = 3,  // (r[b] * Fcond_b(c)) * a       (r times the conductor Fresnel at cosine c)
  T_TRN = 4,  // (a * r[b]) * c                (specular transmission ((1-fr)*r)*eta2)
  // FresnelBlend's two terms, expanded only by fb_band (value mode V_FB):
  T_FBD = 5,  // ((A_b * rd[b]) * (1 - rs[b])) * a, A_b = exp(ra[b] * c) under a coating, else 1
  T_FBS = 6   // (rs[b] + (1 - rs[b]) * c) * a   (the Schlick-weighted specular)
};
struct Term {
  const DBxdf* bx;
  int op;
  float a, c;
};
// value_b = V_SUM:   (0 + t0) + t1            (terms present: n)
//           V_SCALE: t0 * k1
//           V_MIX:   (t0 * k1 + fo) * k2      fo = n > 1 ? 0 + t1 : 0
//           V_FB:    (t0 + t1) * k1           FresnelBlend's diffuse + specular terms
// term_at knows ops 0-4 only; V_FB keeps FresnelBlend out of the per-band
// dispatch that every other material runs.
enum { V_SUM = 0, V_SCALE = 1, V_MIX = 2, V_FB = 3 };
struct SpecVal {
  Term t[2];
  int n, mode;
  float k1, k2;
};

// BxDF kinds a shading variant compiles (bit k: kind k).  The BSDF routines
// below take the set as a template argument and leave out the code of the
// other kinds, so that a kernel shading only diffuse materials carries none
// of the microfacet, FresnelBlend and specular code and its registers; the
// caller guarantees that every component it passes is of a kind in the set.
constexpr int kAllKinds = (1 << K_LAMBERT) | (1 << K_OREN) | (1 << K_MF) | (1 << K_SPECR) | (1 << K_SPECT) |
                          (1 << K_FB);
constexpr int kDiffuseKinds = (1 << K_LAMBERT) | (1 << K_OREN);
template <int KM>
__device__ __forceinline__ constexpr bool has_kind(int k) {
  return ((KM >> k) & 1) != 0;
}

__device__ __forceinline__ float fr_cond_band(const DBxdf& b, float cosi, int i) {  // Fresnel.hs:58-70
  float ac = fabsf(cosi);
  float eta = b.eta[i], k = b.k[i];
  float tmpF = eta * eta + k * k;
  float tmp = tmpF * (ac * ac);
  float ec2 = eta * (2.0f * ac);
  float c2 = ac * ac;
  float rPer2 = (tmpF - ec2 + c2) / (tmpF + ec2 + c2);
  float rPar2 = (tmp - ec2 + 1.0f) / (tmp + ec2 + 1.0f);
  return (rPer2 + rPar2) / 2.0f;
}
template <int KM = kAllKinds>
__device__ __forceinline__ float term_at(const Term& t, int b) {
  const DBxdf& x = *t.bx;
  // diffuse kinds make only T_RS and T_ROS terms
  if constexpr ((KM & ~kDiffuseKinds) == 0) return t.op == T_ROS ? (x.r[b] * t.a) * t.c : x.r[b] * t.a;
  switch (t.op) {
    case T_RS: return x.r[b] * t.a;
    case T_ROS: return (x.r[b] * t.a) * t.c;
    case T_RFX: return (x.r[b] * t.c) * t.a;
    case T_RCX: return (x.r[b] * fr_cond_band(x, t.c, b)) * t.a;
    default: return (t.a * x.r[b]) * t.c;
  }
}
template <int KM = kAllKinds>
__device__ __forceinline__ float spec_at(const SpecVal& v, int b) {
  if (v.mode == V_SUM) {
    float f = 0.0f;
    if (v.n > 0) f = f + term_at<KM>(v.t[0], b);
    if (v.n > 1) f = f + term_at<KM>(v.t[1], b);
    return f;
  }
  if (v.mode == V_SCALE) return term_at<KM>(v.t[0], b) * v.k1;
  if (has_kind<KM>(K_FB) && v.mode == V_FB) {  // t0: T_FBD (a = diffuse scale, c = absorption exponent), t1: T_FBS
    const DBxdf& x = *v.t[0].bx;
    const float ab = x.et > 0.0f ? expf(x.k[b] * v.t[0].c) : 1.0f;
    const float diff = ((ab * x.r[b]) * (1.0f - x.eta[b])) * v.t[0].a;
    const float spec = (x.eta[b] + (1.0f - x.eta[b]) * v.t[1].c) * v.t[1].a;
    return (diff + spec) * v.k1;
  }
  float fo = v.n > 1 ? 0.0f + term_at<KM>(v.t[1], b) : 0.0f;
  return (term_at<KM>(v.t[0], b) * v.k1 + fo) * v.k2;
}
__device__ __forceinline__ SpecVal spec_zero() {
  SpecVal v;
  v.n = 0;
  v.mode = V_SUM;
  v.k1 = v.k2 = 0.0f;
  return v;
}
// r * fresnel(cosi) (Fresnel.hs), scaled by a, as a term
__device__ __forceinline__ Term rf_term(const DBxdf& b, float cosi, float a) {
  if (b.fr == FR_COND) return Term{&b, T_RCX, a, cosi};
  float f = b.fr == FR_DIEL ? fr_diel_scalar(b.ei, b.et, cosi) : 1.0f;
  return Term{&b, T_RFX, a, f};
}

__device__ __forceinline__ float mf_g(f3 wo, f3 wi, f3 wh) {
  float nh = acos_t(wh), no = acos_t(wo), ni = acos_t(wi), oh = fabsf(dot3(wo, wh));
  return hmin(1.0f, hmin(2.0f * nh * no / oh, 2.0f * nh * ni / oh));
}
// The anisotropic microfacet distribution with exponents ex, ey
// (Microfacet.hs:136-196): D, its sampling pdf, and the sampler.
__device__ __forceinline__ float aniso_exp(float ex, float ey, f3 wh, float den) {
  return (ex * wh.x * wh.x + ey * wh.y * wh.y) / den;
}
__device__ __forceinline__ float aniso_d(float ex, float ey, f3 wh) {
  const float costh = acos_t(wh), d = 1.0f - costh * costh;
  if (d == 0.0f) return 0.0f;
  return __builtin_sqrtf((ex + 2.0f) * (ey + 2.0f)) * kInvTwoPi * powf(costh, aniso_exp(ex, ey, wh, d));
}
__device__ __forceinline__ float aniso_pdf(float ex, float ey, f3 wh) {
  const float costh = acos_t(wh);
  const float e = aniso_exp(ex, ey, wh, hmax(0.0f, 1.0f - costh * costh));
  return __builtin_sqrtf((ex + 1.0f) * (ey + 1.0f)) * kInvTwoPi * powf(costh, e);
}
// The azimuth is drawn in the first quadrant and mirrored into the quadrant
// u1 selects; the returned pdf is aniso_pdf at the returned half vector.
__device__ __forceinline__ f3 aniso_sample(float ex, float ey, float u1, float u2, float* pdf) {
  float up;
  int q;
  if (u1 < 0.25f) {
    up = 4.0f * u1, q = 0;
  } else if (u1 < 0.5f) {
    up = 4.0f * (0.5f - u1), q = 1;
  } else if (u1 < 0.75f) {
    up = 4.0f * (u1 - 0.5f), q = 2;
  } else {
    up = 4.0f * (1.0f - u1), q = 3;
  }
  const float p = ex == ey ? kPi * up * 0.5f
                           : atanf(__builtin_sqrtf((ex + 1.0f) / (ey + 1.0f)) * tanf(kPi * up * 0.5f));
  const float cp = cosf(p), sp = sinf(p);
  const float cost = powf(u2, 1.0f / (ex * cp * cp + ey * sp * sp + 1.0f));
  const float phi = q == 0 ? p : (q == 1 ? kPi - p : (q == 2 ? p + kPi : kTwoPi - p));
  const float sint = __builtin_sqrtf(hmax(0.0f, 1.0f - cost * cost));
  const f3 wh = mk3(sint * cosf(phi), sint * sinf(phi), cost);
  const float f = kInvTwoPi * powf(cost, aniso_exp(ex, ey, wh, 1.0f - cost * cost));
  *pdf = __builtin_sqrtf((ex + 1.0f) * (ey + 1.0f)) * f;
  return wh;
}
// FresnelBlend (Microfacet.hs:56-104), a material's only component: e wo wi
// as its diffuse and specular terms, its pdf, and its sample with adj False
// (f = e wo wi / pdf; false when the distribution's pdf is zero).
__device__ __forceinline__ f3 fb_half(f3 wi, f3 wo) {
  const f3 wh = norm3(add3(wi, wo));
  return wh.z < 0.0f ? neg3(wh) : wh;
}
__device__ __forceinline__ void fb_terms(const DBxdf& b, f3 wo, f3 wi, Term* td, Term* ts) {
  const float costi = acos_t(wi), costo = acos_t(wo);
  const float ka = -b.et * (costi + costo) / (costi * costo);
  const float dd = (costo * 28.0f / 23.0f * kPi) * (1.0f - powf(1.0f - 0.5f * costi, 5.0f)) *
                   (1.0f - powf(1.0f - 0.5f * costo, 5.0f));
  const f3 wh = fb_half(wi, wo);
  const float costih = fabsf(dot3(wi, wh));
  const float ds = aniso_d(b.e, b.ei, wh) * costo / (4.0f * costih * hmax(costi, costo));
  *td = Term{&b, T_FBD, dd, ka};
  *ts = Term{&b, T_FBS, ds, powf(1.0f - costih, 5.0f)};
}
__device__ __forceinline__ float fb_pdf(const DBxdf& b, f3 wo, f3 wi) {
  if (!same_hemi(wo, wi)) return 0.0f;
  const f3 wh = fb_half(wi, wo);
  return 0.5f * (acos_t(wi) * kInvPi + aniso_pdf(b.e, b.ei, wh) / (4.0f * fabsf(dot3(wo, wh))));
}
__device__ __forceinline__ bool fb_sample(const DBxdf& b, f3 wo, float u1, float u2, Term* td, Term* ts, f3* wi,
                                          float* pdf) {
  f3 w, wh;
  float pdfp;
  if (u1 < 0.5f) {
    w = cos_hemisphere(u1 * 2.0f, u2);
    if (wo.z < 0.0f) w.z = -w.z;  // toSameHemisphere
    wh = fb_half(w, wo);
    pdfp = aniso_pdf(b.e, b.ei, wh);
  } else {
    wh = aniso_sample(b.e, b.ei, 2.0f * (u1 - 0.5f), u2, &pdfp);
    w = sub3(sscale3(2.0f * dot3(wo, wh), wh), wo);
  }
  *wi = w;
  if (pdfp == 0.0f) return false;
  *pdf = 0.5f * (acos_t(w) * kInvPi + pdfp / (4.0f * fabsf(dot3(wo, wh))));
  fb_terms(b, wo, w, td, ts);
  return true;
}
__device__ __forceinline__ float oren_scale(const DBxdf& b, f3 wo, f3 wi) {  // Diffuse.hs:53-65
  float sinti = sin_t(wi), sinto = sin_t(wo);
  float maxcos = 0.0f;
  if (sinti > 1e-4f && sinto > 1e-4f) {
    float cpi = sinti == 0.0f ? 1.0f : clampf_(wi.x / sinti, -1.0f, 1.0f);
    float spi = sinti == 0.0f ? 0.0f : clampf_(wi.y / sinti, -1.0f, 1.0f);
    float cpo = sinto == 0.0f ? 1.0f : clampf_(wo.x / sinto, -1.0f, 1.0f);
    float spo = sinto == 0.0f ? 0.0f : clampf_(wo.y / sinto, -1.0f, 1.0f);
    maxcos = hmax(0.0f, cpi * cpo + spi * spo);
  }
  float sina, tanb;
  if (acos_t(wi) > acos_t(wo)) {
    sina = sinto;
    tanb = sinti / acos_t(wi);
  } else {
    sina = sinti;
    tanb = sinto / acos_t(wo);
  }
  return b.A + b.B * maxcos * sina * tanb;
}

// bxdfEval b wo wi as a term; false when the component contributes nothing
template <int KM = kAllKinds>
__device__ __forceinline__ bool bx_eval_term(const DBxdf& b, f3 wo, f3 wi, Term* t) {
  if (b.flip) wi.z = -wi.z;  // brdfToBtdf: e wo wi = bxdfEval brdf wo (otherHemisphere wi)
  if (has_kind<KM>(K_LAMBERT) && b.kind == K_LAMBERT) {
    *t = Term{&b, T_RS, kInvPi * acos_t(wo), 0.0f};
    return true;
  }
  if (has_kind<KM>(K_OREN) && b.kind == K_OREN) {
    *t = Term{&b, T_ROS, oren_scale(b, wo, wi), kInvPi * acos_t(wo)};
    return true;
  }
  if (has_kind<KM>(K_MF) && b.kind == K_MF) {  // Microfacet.hs:21-32
    float costo = acos_t(wo), costi = acos_t(wi);
    if (costi == 0.0f || costo == 0.0f) return false;
    f3 whp = add3(wi, wo);
    if (whp.x == 0.0f && whp.y == 0.0f && whp.z == 0.0f) return false;
    f3 wh = norm3(whp);
    if (cos_t(wh) < 0.0f) return false;
    float costh = dot3(wi, wh);
    float D = (b.e + 2.0f) * kInvTwoPi * powf(acos_t(wh), b.e);
    float x = D * mf_g(wo, wi, wh) / (4.0f * costi);
    *t = rf_term(b, costh, x);
    return true;
  }
  return false;  // specular components evaluate to zero
}
template <int KM = kAllKinds>
__device__ __forceinline__ float bx_pdf(const DBxdf& b, f3 wo, f3 wi) {
  if (b.flip) wi.z = -wi.z;
  if ((has_kind<KM>(K_LAMBERT) && b.kind == K_LAMBERT) || (has_kind<KM>(K_OREN) && b.kind == K_OREN))
    return same_hemi(wo, wi) ? kInvPi * acos_t(wi) : 0.0f;
  if (has_kind<KM>(K_FB) && b.kind == K_FB) return fb_pdf(b, wo, wi);
  if (has_kind<KM>(K_MF) && b.kind == K_MF) {
    f3 whp = add3(wo, wi);
    if (sqlen3(whp) == 0.0f) return 0.0f;
    f3 wh = norm3(whp);
    if (cos_t(wh) < 0.0f) return 0.0f;
    return (b.e + 1.0f) * powf(acos_t(wh), b.e) * kInvTwoPi / (4.0f * fabsf(dot3(wo, wh)));
  }
  return 0.0f;
}
// bxdfSample with adj = False: the sampled value as a term (zero: *zero set)
template <int KM = kAllKinds>
__device__ __forceinline__ void bx_sample(const DBxdf& b, f3 wo, float u1, float u2, Term* f, bool* zero,
                                          f3* wi, float* pdf) {
  *zero = false;
  if ((has_kind<KM>(K_LAMBERT) && b.kind == K_LAMBERT) || (has_kind<KM>(K_OREN) && b.kind == K_OREN)) {
    f3 w = cos_hemisphere(u1, u2);
    if (wo.z < 0.0f) w.z = -w.z;  // toSameHemisphere
    if (same_hemi(wo, w)) {
      *f = Term{&b, T_RS, (has_kind<KM>(K_OREN) && b.kind == K_OREN) ? oren_scale(b, wo, w) : 1.0f, 0.0f};
      *wi = w;
      *pdf = kInvPi * acos_t(w);
    } else {
      *zero = true;
      *wi = b.kind == K_LAMBERT ? wo : w;
      *pdf = 0.0f;
    }
    if (b.flip) wi->z = -wi->z;  // brdfToBtdf: the sampled direction's other hemisphere
    return;
  }
  if (has_kind<KM>(K_MF) && b.kind == K_MF) {  // Microfacet.hs:42-54,175-182
    float cost = powf(u1, 1.0f / (b.e + 1.0f));
    float sint = __builtin_sqrtf(hmax(0.0f, 1.0f - cost * cost));
    float phi = u2 * 2.0f * kPi;
    f3 whp = mk3(sint * cosf(phi), sint * sinf(phi), cost);
    float ff = powf(cost, b.e) * kInvTwoPi;
    float d = (b.e + 2.0f) * ff, pd = (b.e + 1.0f) * ff;
    f3 wh = cos_t(whp) < 0.0f ? neg3(whp) : whp;
    f3 w = sub3(sscale3(2.0f * dot3(wo, wh), wh), wo);
    float costH = dot3(wo, wh);
    if (same_hemi(wo, w)) {
      float fact = d * fabsf(costH) / pd * mf_g(wo, w, wh);
      *f = rf_term(b, costH, fact / acos_t(w));
      *wi = w;
      *pdf = pd / (4.0f * fabsf(costH));
    } else {
      *zero = true;
      *wi = wo;
      *pdf = 0.0f;
    }
    return;
  }
  if (has_kind<KM>(K_SPECR) && b.kind == K_SPECR) {  // Specular.hs:11-20
    *f = rf_term(b, cos_t(wo), 1.0f);
    *wi = mk3(-wo.x, -wo.y, wo.z);
    *pdf = 1.0f;
    return;
  }
  if (!has_kind<KM>(K_SPECT)) {  // not reached: every component is of a kind in KM
    *zero = true;
    *wi = wo;
    *pdf = 0.0f;
    return;
  }
  // specTrans (Specular.hs:33-58)
  bool entering = cos_t(wo) > 0.0f;
  float ei = entering ? b.ei : b.et, et = entering ? b.et : b.ei;
  float eta = ei / et, eta2 = eta * eta;
  float sint2 = eta2 * sin_t2(wo);
  if (sint2 >= 1.0f) {
    *zero = true;
    *wi = wo;
    *pdf = 0.0f;
    return;
  }
  float c = __builtin_sqrtf(hmax(0.0f, 1.0f - sint2));
  float cost = entering ? -c : c;
  *wi = mk3(eta * (-wo.x), eta * (-wo.y), cost);
  float fr = fr_diel_scalar(ei, et, cost);
  *f = Term{&b, T_TRN, 1.0f - fr, eta2};
  *pdf = 1.0f;
}

// Selector textures (Texture.hs:166-219) on the shading DG: true where the
// material takes its tex2 spectra (the alternate material entry).
__device__ __forceinline__ bool tex_selects_alt(const DMaterial& m, const DG& dg) {
  const float* q = m.tex;
  if (m.tex_kind == BGPU_TEX_CHECKER) {  // checkerBoard: tex1 where the sum is even
    const long long s = (long long)floorf(dg.p.x * q[0]) + (long long)floorf(dg.p.y * q[1]) + (long long)floorf(dg.p.z * q[2]);
    return (s & 1) != 0;
  }
  float x, z;  // graphPaper over its mapping
  if (m.tex_map == BGPU_MAP_PLANAR) {
    x = dot3(dg.p, mk3(q[1], q[2], q[3])) + q[7];
    z = dot3(dg.p, mk3(q[4], q[5], q[6])) + q[8];
  } else {
    x = q[1] * dg.u + q[3];
    z = q[2] * dg.v + q[4];
  }
  const float xf = fabsf(x - truncf(x)), zf = fabsf(z - truncf(z));  // properFraction
  const float lo = q[0] / 2.0f, hi = 1.0f - lo;
  return xf < lo || zf < lo || xf > hi || zf > hi;
}

// `pm`: the path's BxDF records evaluated by the texture kernel, taken for a
// material whose inputs are node textures.
__device__ __forceinline__ Bsdf make_bsdf(const DevPtScene& S, int material, const DG& dgg, const DG& dgs,
                                          const DMaterial* pm) {
  Bsdf b;
  b.m = (material >= 0 && material < S.nmats) ? &S.mats[material] : nullptr;
  if (b.m && b.m->tex_kind == BGPU_TEX_NODES) b.m = pm;
  else if (b.m && b.m->tex_kind != BGPU_TEX_NONE && tex_selects_alt(*b.m, dgs)) b.m = &S.mats[b.m->alt];
  f3 nn = dgs.n, sn = norm3(dgs.dpdu);
  if (b.m && b.m == pm && pm->tex_map == 1) {  // bumpMapped (Reflection.hs:344-345): the bumped shading DG
    nn = mk3(pm->tex[0], pm->tex[1], pm->tex[2]);
    sn = norm3(mk3(pm->tex[3], pm->tex[4], pm->tex[5]));
  }
  b.cs = Frame{sn, cross3(nn, sn), nn};
  b.p = dgs.p;
  b.ng = dgg.n;
  return b;
}
__device__ __forceinline__ int bsdf_n(const Bsdf& b) { return b.m ? b.m->n : 0; }

template <int KM = kAllKinds>
__device__ __forceinline__ SpecVal eval_bsdf(const Bsdf& b, f3 woW, f3 wiW) {  // Reflection.hs:318-332
  SpecVal f = spec_zero();
  float cosWo = dot3(woW, b.ng);
  float side = dot3(wiW, b.ng) / cosWo;
  if (side == 0.0f || fabsf(cosWo) < 1e-5f) return f;
  int need = side < 0.0f ? BX_T : BX_R;
  f3 wo = to_local(b.cs, woW), wi = to_local(b.cs, wiW);
  int n = bsdf_n(b);
  for (int i = 0; i < n; ++i) {
    const DBxdf& c = b.m->c[i];
    if (!(c.type & need)) continue;
    if (has_kind<KM>(K_FB) && c.kind == K_FB) {  // the only component: 0 + (diff + spec) = (diff + spec) * 1
      fb_terms(c, wi, wo, &f.t[0], &f.t[1]);
      f.n = 2;
      f.mode = V_FB;
      f.k1 = 1.0f;
    } else if (bx_eval_term<KM>(c, wi, wo, &f.t[f.n])) {
      f.n++;
    }
  }
  return f;
}
template <int KM = kAllKinds>
__device__ __forceinline__ float bsdf_pdf(const Bsdf& b, f3 woW, f3 wiW) {  // Reflection.hs:251-257
  int n = bsdf_n(b);
  if (n == 0) return 0.0f;
  f3 wo = to_local(b.cs, woW), wi = to_local(b.cs, wiW);
  float s = 0.0f;
  for (int i = 0; i < n; ++i) s = s + bx_pdf<KM>(b.m->c[i], wo, wi);
  return s / (float)n;
}
// sampleBsdf (Reflection.hs:278-316), adj = False, all flags; returns pdf.
template <int KM = kAllKinds>
__device__ __forceinline__ float sample_bsdf(const Bsdf& b, f3 woW, float uComp, float u1, float u2,
                                             SpecVal& f, f3* wiW, int* type) {
  f = spec_zero();
  *wiW = mk3(0.0f, 1.0f, 0.0f);
  *type = BX_R | BX_D;
  int cnt = bsdf_n(b);
  if (cnt == 0) return 0.0f;
  f3 wo = to_local(b.cs, woW);
  float cntf = (float)cnt, invCnt = 1.0f / cntf;
  int sNum = (int)floorf(uComp * cntf);
  sNum = cnt - 1 < sNum ? cnt - 1 : sNum;
  sNum = sNum < 0 ? 0 : sNum;
  const DBxdf& bx = b.m->c[sNum];
  if (has_kind<KM>(K_FB) && bx.kind == K_FB) {  // the only component, a reflection: f = e wo wi / pdf = (t0 + t1) * (1 / pdf)
    f3 wi;
    float pdfp;
    if (!fb_sample(bx, wo, u1, u2, &f.t[0], &f.t[1], &wi, &pdfp)) return 0.0f;
    const f3 ww = to_world(b.cs, wi);
    const float side = dot3(ww, b.ng) / dot3(woW, b.ng);
    if (side == 0.0f || side < 0.0f) return 0.0f;
    *wiW = ww;
    *type = bx.type;
    f.n = 2;
    f.mode = V_FB;
    f.k1 = 1.0f / pdfp;
    return pdfp;
  }
  Term fs;
  bool fzero;
  f3 wi;
  float pdfp;
  bx_sample<KM>(bx, wo, u1, u2, &fs, &fzero, &wi, &pdfp);
  f3 ww = to_world(b.cs, wi);
  float side = dot3(ww, b.ng) / dot3(woW, b.ng);
  if (pdfp == 0.0f || side == 0.0f) return 0.0f;
  int need = side < 0.0f ? BX_T : BX_R;
  if (!(bx.type & need)) return 0.0f;
  *wiW = ww;
  *type = bx.type;
  if (fzero) fs = Term{&bx, T_RS, 0.0f, 0.0f};  // a zero spectrum (r * 0)
  f.t[0] = fs;
  f.n = 1;
  if (bx.type & BX_S) {
    f.mode = V_SCALE;
    f.k1 = cntf;
    return pdfp * invCnt;
  }
  if (cnt == 1) {
    f.mode = V_SCALE;
    f.k1 = 1.0f;
    return pdfp;
  }
  float other = 0.0f;
  for (int i = 0; i < cnt; ++i) {
    if (i == sNum) continue;
    other = other + bx_pdf<KM>(b.m->c[i], wo, wi);
    if ((b.m->c[i].type & need) && bx_eval_term<KM>(b.m->c[i], wi, wo, &f.t[1])) f.n = 2;
  }
  float pdf = (pdfp + other) * invCnt;
  f.mode = V_MIX;
  f.k1 = pdfp;
  f.k2 = 1.0f / pdf;
  return pdf;
}

// sampleBsdf' flags with uComp 0.5 and uDir (0.5, 0.5) (Reflection.hs:278-316,
// DirectLighting.hs:53): only the components whose type is a subset of
// `flags` (bxdfMatches) take part.  Returns the pdf; f as sample_bsdf's.
__device__ __forceinline__ float sample_bsdf_flt(const Bsdf& b, f3 woW, int flags, SpecVal& f, f3* wiW) {
  f = spec_zero();
  *wiW = mk3(0.0f, 1.0f, 0.0f);
  const int n = bsdf_n(b);
  int m[2], cnt = 0;
  for (int i = 0; i < n; ++i)
    if ((b.m->c[i].type & flags) == b.m->c[i].type) m[cnt++] = i;
  if (cnt == 0) return 0.0f;
  const f3 wo = to_local(b.cs, woW);
  const float cntf = (float)cnt, invCnt = 1.0f / cntf;
  int sNum = (int)floorf(0.5f * cntf);
  sNum = cnt - 1 < sNum ? cnt - 1 : sNum;
  sNum = sNum < 0 ? 0 : sNum;
  const DBxdf& bx = b.m->c[sNum == 0 ? m[0] : m[1]];
  Term fs;
  bool fzero;
  f3 wi;
  float pdfp;
  bx_sample(bx, wo, 0.5f, 0.5f, &fs, &fzero, &wi, &pdfp);
  const f3 ww = to_world(b.cs, wi);
  const float side = dot3(ww, b.ng) / dot3(woW, b.ng);
  if (pdfp == 0.0f || side == 0.0f) return 0.0f;
  const int need = side < 0.0f ? BX_T : BX_R;
  if (!(bx.type & need)) return 0.0f;
  *wiW = ww;
  if (fzero) fs = Term{&bx, T_RS, 0.0f, 0.0f};
  f.t[0] = fs;
  f.n = 1;
  if (bx.type & BX_S) {
    f.mode = V_SCALE;
    f.k1 = cntf;
    return pdfp * invCnt;
  }
  if (cnt == 1) {
    f.mode = V_SCALE;
    f.k1 = 1.0f;
    return pdfp;
  }
  const DBxdf& o = b.m->c[sNum == 0 ? m[1] : m[0]];  // the other matching component
  const float pdf = (pdfp + bx_pdf(o, wo, wi)) * invCnt;
  if ((o.type & need) && bx_eval_term(o, wi, wo, &f.t[1])) f.n = 2;
  f.mode = V_MIX;
  f.k1 = pdfp;
  f.k2 = 1.0f / pdf;
  return pdf;
}

// -------------------------------------------------------------- lights
// Light kinds a shading variant compiles (bit SH_x: area lights on shape x;
// bit kLightInf: infinite lights).  The routines below take the set as a
// template argument: the variant for scenes whose lights are all quads
// carries none of the other shapes' sampling, pdf and intersection code and
// their registers.  The caller guarantees that every light it passes is of a
// kind in the set.  kAllLights leaves out sun/sky maps: their evaluation
// runs only in the kernels compiled with kAllLightsSky, while a scene holding
// one is resident.
constexpr int kAllLights = kAllShapes | (1 << kLightInf) | (1 << kLightPoint) | (1 << kLightDir);
constexpr int kAllLightsSky = kAllLights | (1 << kLightSky);
template <int LK>
__device__ __forceinline__ constexpr int light_shapes() {
  return LK & kAllShapes;
}
template <int LK>
__device__ __forceinline__ constexpr bool has_area_light() {
  return light_shapes<LK>() != 0;
}
template <int LK>
__device__ __forceinline__ constexpr bool has_inf_light() {
  return ((LK >> kLightInf) & 1) != 0;
}
template <int LK>
__device__ __forceinline__ constexpr bool has_delta_light() {
  return ((LK >> kLightPoint) & 3) != 0;
}
template <int LK>
__device__ __forceinline__ constexpr bool has_sky_light() {
  return ((LK >> kLightSky) & 1) != 0;
}

// A light's radiance along one direction: the 16 bands at `p` (an area
// light's radiance, a delta light's intensity, a map texel, or the sun's
// spectrum where a sun/sky map sees the sun), plus for a sun/sky map the
// sky's colour as a combination of bases (SkyColor); black when both are
// absent.  Light kinds without a sun/sky map use `p` alone.
struct LRad {
  const float* p;
  const float* sky;  // the sun/sky block, nullptr without a sky term
  SkyColor c;
};
__device__ __forceinline__ LRad lrad(const float* p) {
  LRad r;
  r.p = p;
  r.sky = nullptr;
  return r;
}
template <int LK>
__device__ __forceinline__ bool lrad_black(const LRad& r) {
  return !r.p && (!has_sky_light<LK>() || !r.sky);
}
template <int LK>
__device__ __forceinline__ float lrad_at(const LRad& r, int b) {
  if constexpr (!has_sky_light<LK>()) return r.p[b];
  if (!r.sky) return r.p ? r.p[b] : 0.0f;
  return sky_band(r.sky, r.c, b);
}
// sky_color as a call: the shading kernels run at their register limit, and
// the model's arithmetic inlined into them would make the allocator spill on
// every path; as a call, only sun/sky evaluations save and restore registers.
__device__ __noinline__ SkyColor sky_color_call(const float* par, float u, float v) {
  return sky_color(par, u, v);
}
// The sun/sky map of light l at (u, v): sky term and sun as sky_color gives them.
__device__ __forceinline__ LRad sky_rad(const DevPtScene& S, const DLight& l, float u, float v) {
  const float* blk = S.maps + l.map_off;
  LRad r;
  r.c = sky_color_call(blk + SKY_PAR, u, v);
  r.sky = (r.c.sky || r.c.sun) ? blk : nullptr;
  r.p = nullptr;
  return r;
}
template <int SK = kAllShapes>
__device__ __forceinline__ float shape_area(const DShape& s) {  // Shape.hs:314-328
  constexpr bool one = one_shape_kind<SK>();
  if ((SK >> SH_QUAD & 1) && (one || s.type == SH_QUAD)) return 4.0f * s.p[0] * s.p[1];
  if ((SK >> SH_SPHERE & 1) && (one || s.type == SH_SPHERE)) return s.p[0] * s.p[0] * 4.0f * kPi;
  if ((SK >> SH_BOX & 1) && (one || s.type == SH_BOX)) {
    float h = s.p[3] - s.p[0], w = s.p[4] - s.p[1], l = s.p[5] - s.p[2];
    return 2.0f * (h * w + h * l + w * l);
  }
  if ((SK >> SH_DISK & 1) && (one || s.type == SH_DISK)) return kPi * (s.p[1] * s.p[1] - s.p[2] * s.p[2]);
  return 2.0f * kPi * s.p[0] * (s.p[2] - s.p[1]);
}
__device__ __forceinline__ float cone_pdf(float c) { return c >= 1.0f ? 0.0f : 1.0f / (kTwoPi * (1.0f - c)); }
template <int LK = kAllLights>
__device__ __forceinline__ float general_pdf(const DShape& s, f3 p, f3 wi) {  // Shape.hs:346-350
  DRay r;
  r.o = p;
  r.d = wi;
  r.tmin = 1e-3f;
  r.tmax = INFINITY;
  float t;
  int ax;
  constexpr int SK = light_shapes<LK>();
  if (!shape_intersect_k<SK>(s, r, &t, &ax)) return 0.0f;
  DG d = shape_dg_k<SK>(s, r, t, ax);
  float pd = sqlen3(sub3(p, ray_at(r, t))) / (fabsf(dot3(d.n, neg3(wi))) * shape_area<SK>(s));
  return isinf(pd) ? 0.0f : pd;
}
template <int LK = kAllLights>
__device__ __forceinline__ float shape_pdf(const DShape& s, f3 p, f3 wi) {  // Shape.hs:333-344
  constexpr int SK = light_shapes<LK>();
  constexpr bool one = one_shape_kind<SK>();
  if ((SK >> SH_SPHERE & 1) && (one || s.type == SH_SPHERE) && !(sqlen3(p) - s.p[0] * s.p[0] < 1e-4f)) {
    float r = s.p[0];
    float st2 = r * r / sqlen3(p);
    return cone_pdf(__builtin_sqrtf(hmax(0.0f, 1.0f - st2)));
  }
  return general_pdf<LK>(s, p, wi);
}
__device__ __forceinline__ void remap_rand(int segs, float u, int* seg, float* u2) {
  float fs = (float)segs;
  int sg = (int)floorf(u * fs);
  sg = segs - 1 < sg ? segs - 1 : sg;
  *seg = sg;
  *u2 = (u - (float)sg / fs) * fs;
}
// The quad's sampleShape (Shape.hs:404-407).
__device__ __forceinline__ void quad_sample(const DShape& s, float u1, float u2, f3* ps, f3* ns) {
  *ps = mk3(lerpf_(u1, -s.p[0], s.p[0]), lerpf_(u2, -s.p[1], s.p[1]), 0.0f);
  *ns = mk3(0.0f, 0.0f, -1.0f);
}
// The other shapes' sampleShape (Shape.hs:352-410).
__device__ __forceinline__ void sphere_sample(const DShape& s, f3 p, float u1, float u2, f3* ps, f3* ns) {
  float r = s.p[0];
  if (!(sqlen3(p) - r * r < 1e-4f)) {  // Shape.hs:364-375
    f3 dn = norm3(neg3(p));
    Frame cs = coord_system(dn);
    float ctm = __builtin_sqrtf(hmax(0.0f, 1.0f - (r * r) / sqlen3(p)));
    float ct = lerpf_(u1, ctm, 1.0f);
    float st = __builtin_sqrtf(1.0f - ct * ct);
    float phi = u2 * kTwoPi;
    f3 d = add3(add3(scl3(cs.s, cosf(phi) * st), scl3(cs.t, sinf(phi) * st)), scl3(cs.n, ct));
    DRay ray;
    ray.o = p;
    ray.d = d;
    ray.tmin = 0.0f;
    ray.tmax = INFINITY;
    float t;
    int ax;
    f3 q = sphere_intersect(s, ray, &t, &ax) ? ray_at(ray, t) : scl3(dn, r);
    *ps = q;
    *ns = norm3(q);
    return;
  }
  f3 q = uniform_sphere(u1, u2);
  *ps = scl3(q, r);
  *ns = q;
}
__device__ __forceinline__ void box_sample(const DShape& s, float u1, float u2, f3* ps, f3* ns) {
  int axis, nf;
  float a, c;
  remap_rand(3, u1, &axis, &a);
  remap_rand(2, u2, &nf, &c);
  *ns = set_comp(axis, (float)nf * 2.0f - 1.0f, mk3(0, 0, 0));
  int oa0 = (axis + 1) % 3, oa1 = (axis + 2) % 3;
  f3 pmin = mk3(s.p[0], s.p[1], s.p[2]), pmax = mk3(s.p[3], s.p[4], s.p[5]);
  f3 base = nf == 0 ? pmin : pmax;
  base = set_comp(oa1, lerpf_(c, comp3(pmin, oa1), comp3(pmax, oa1)), base);
  *ps = set_comp(oa0, lerpf_(a, comp3(pmin, oa0), comp3(pmax, oa0)), base);
}
__device__ __forceinline__ void cyl_sample(const DShape& s, float u1, float u2, f3* ps, f3* ns) {
  float z = lerpf_(u1, s.p[1], s.p[2]), phi = lerpf_(u2, 0.0f, kTwoPi);
  *ps = mk3(s.p[0] * cosf(phi), s.p[0] * sinf(phi), z);
  *ns = norm3(mk3(ps->x, ps->y, 0.0f));
}
__device__ __forceinline__ void disk_sample(const DShape& s, float u1, float u2, f3* ps, f3* ns) {
  float r = lerpf_(u1, s.p[2], s.p[1]), phi = lerpf_(u2, 0.0f, s.p[3]);
  *ps = mk3(r * cosf(phi), r * sinf(phi), s.p[0]);
  *ns = mk3(0.0f, 0.0f, -1.0f);
}
template <int LK = kAllLights>
__device__ __forceinline__ void sample_shape(const DShape& s, f3 p, float u1, float u2, f3* ps, f3* ns) {
  constexpr int SK = light_shapes<LK>();
  constexpr bool one = one_shape_kind<SK>();
  if ((SK >> SH_QUAD & 1) && (one || s.type == SH_QUAD)) quad_sample(s, u1, u2, ps, ns);
  else if ((SK >> SH_SPHERE & 1) && (one || s.type == SH_SPHERE)) sphere_sample(s, p, u1, u2, ps, ns);
  else if ((SK >> SH_BOX & 1) && (one || s.type == SH_BOX)) box_sample(s, u1, u2, ps, ns);
  else if ((SK >> SH_DISK & 1) && (one || s.type == SH_DISK)) disk_sample(s, u1, u2, ps, ns);
  else cyl_sample(s, u1, u2, ps, ns);
}

// Dist1D arrays at `d`: n, func[n], cdf[n+1], fi ; Dist2D: nu, nv, marginal, then nv conditionals.
__device__ __forceinline__ int d1_upper(const float* cdf, int len, float u) {
  int lo = 0, hi = len;  // first index with cdf >= u (cdf is non-decreasing)
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (cdf[mid] >= u) hi = mid;
    else lo = mid + 1;
  }
  int r = lo == len ? len - 1 : lo - 1;
  r = r < 0 ? 0 : r;
  return r < len - 2 ? r : len - 2;
}
__device__ __forceinline__ void d1_sample(const float* d, float u, float* x, float* pdf, int* off) {
  int n = __float_as_int(d[0]);
  const float* func = d + 1;
  const float* cdf = d + 1 + n;
  float fi = d[2 + 2 * n];
  int o = d1_upper(cdf, n + 1, u);
  *pdf = fi == 0.0f ? 0.0f : func[o] / fi;
  float du = (u - cdf[o]) / (cdf[o + 1] - cdf[o]);
  *x = ((float)o + du) / (float)n;
  *off = o;
}
__device__ __forceinline__ const float* d2_cond(const float* d, int v) {
  int nu = __float_as_int(d[0]), nv = __float_as_int(d[1]);
  const float* marg = d + 2;
  const float* first = marg + 3 + 2 * nv;
  return first + (size_t)v * (3 + 2 * nu);
}
// A texel map's value at (u, v), as rgbfToTexMap looks it up (IO/Bitmap.hs:21-28)
__device__ __forceinline__ const float* tex_px(const DevPtScene& S, const DLight& l, float u, float v) {
  int px = (int)floorf((1.0f - u) * (float)l.w);
  int py = (int)floorf((1.0f - v) * (float)l.h);
  px = px < 0 ? 0 : (px > l.w - 1 ? l.w - 1 : px);  // max 0 $ min (w - 1), also for non-finite input
  py = py < 0 ? 0 : (py > l.h - 1 ? l.h - 1 : py);
  return S.maps + l.map_off + 16 * ((size_t)py * l.w + px);
}
// texMapEval of an infinite light's map at (u, v)
template <int LK>
__device__ __forceinline__ LRad map_eval(const DevPtScene& S, const DLight& l, float u, float v) {
  if (has_sky_light<LK>() && l.sky) return sky_rad(S, l, u, v);
  return lrad(tex_px(S, l, u, v));
}

struct LSample {
  LRad li;          // radiance (times `scale`); for a sun/sky map see `sky`
  float scale;      // 1, or a delta light's falloff / cosine factor
  bool delta;       // a point or directional light (Light.hs:142-150)
  bool sky;         // a sun/sky map's sample: li is left black and the map is
  float su, sv;     // evaluated at (su, sv) by sample_radiance where it is used
  f3 wi;
  DRay ray;
  float pdf;
};
template <int LK = kAllLights>
__device__ __forceinline__ LSample sample_light(const DevPtScene& S, int li, f3 pW, float e, f3 n, float u1,
                                                float u2) {
  const DLight& l = S.lights[li];
  LSample ls;
  ls.li = lrad(nullptr);
  ls.sky = false;
  ls.wi = mk3(0, 1, 0);
  ls.ray.o = mk3(0, 0, 0);
  ls.ray.d = mk3(0, 1, 0);
  ls.ray.tmin = 0.0f;
  ls.ray.tmax = 1.0f;
  ls.pdf = 0.0f;
  ls.scale = 1.0f;
  ls.delta = false;
  if constexpr (has_delta_light<LK>()) {
    if (l.type == LT_POINT) {  // Light.hs:146-150: r / |pos - p|^2 toward pos, the ray along pos - p
      const f3 d = sub3(mk3(l.vec[0], l.vec[1], l.vec[2]), pW);
      ls.li = lrad(l.L);
      ls.scale = 1.0f / sqlen3(d);
      ls.wi = norm3(d);
      ls.ray.o = pW;
      ls.ray.d = d;
      ls.ray.tmin = e;
      ls.ray.tmax = INFINITY;
      ls.pdf = 1.0f;
      ls.delta = true;
      return ls;
    }
    if (l.type == LT_DIR) {  // Light.hs:142-144: r |n . d| from d; n is the shading normal
      const f3 d = mk3(l.vec[0], l.vec[1], l.vec[2]);
      ls.li = lrad(l.L);
      ls.scale = fabsf(dot3(n, d));
      ls.wi = d;
      ls.ray.o = pW;
      ls.ray.d = d;
      ls.ray.tmin = e;
      ls.ray.tmax = INFINITY;
      ls.pdf = 1.0f;
      ls.delta = true;
      return ls;
    }
  }
  if (has_inf_light<LK>() && l.type == LT_INF) {  // Light.hs:130-141
    const float* d = S.dist + l.dist_off;
    float v, pdf1, u, pdf0;
    int off, o2;
    d1_sample(d + 2, u2, &v, &pdf1, &off);
    d1_sample(d2_cond(d, off), u1, &u, &pdf0, &o2);
    float mapPdf = pdf0 * pdf1;
    if (mapPdf == 0.0f) return ls;
    float phi = u * 2.0f * kPi, theta = v * kPi;
    float sint = sinf(theta);
    if (sint == 0.0f) return ls;
    f3 dl = mk3(sinf(theta) * cosf(phi), sinf(theta) * sinf(phi), cosf(theta));
    f3 wi = xform_vector(l.l2w, dl);
    if (has_sky_light<LK>() && l.sky) {  // evaluated late: its terms would stay live across the BSDF
      ls.sky = true;
      ls.su = u;
      ls.sv = v;
    } else {
      ls.li = lrad(tex_px(S, l, u, v));
    }
    ls.wi = wi;
    ls.ray.o = pW;
    ls.ray.d = wi;
    ls.ray.tmin = e;
    ls.ray.tmax = INFINITY;
    ls.pdf = mapPdf / (2.0f * kPi * kPi * sint);
    return ls;
  }
  if constexpr (!has_area_light<LK>()) return ls;
  const DShape& s = l.sh;  // Light.hs:152-160
  const float* l2w = l.l2w;
  f3 p = xform_point(s.w2o, pW);
  f3 ps, ns;
  sample_shape<LK>(s, p, u1, u2, &ps, &ns);
  f3 wi = norm3(sub3(ps, p));
  ls.li = lrad(dot3(ns, wi) < 0.0f ? l.L : nullptr);
  ls.wi = xform_vector(l2w, wi);
  ls.pdf = shape_pdf<LK>(s, p, wi);
  ls.ray.o = xform_point(l2w, p);
  ls.ray.d = xform_vector(l2w, wi);
  ls.ray.tmin = e;
  ls.ray.tmax = __builtin_sqrtf(sqlen3(sub3(ps, p))) - e;
  return ls;
}
// A light sample's radiance, the sun/sky map's evaluated here.
template <int LK>
__device__ __forceinline__ LRad sample_radiance(const DevPtScene& S, int li, const LSample& ls) {
  if (has_sky_light<LK>() && ls.sky) return sky_rad(S, S.lights[li], ls.su, ls.sv);
  return ls.li;
}
template <int LK = kAllLights>
__device__ __forceinline__ float light_pdf(const DevPtScene& S, int li, f3 p, f3 wi) {  // Light.hs:215-229
  const DLight& l = S.lights[li];
  if (has_inf_light<LK>() && l.type == LT_INF) {
    f3 v = xform_vector(l.w2l, wi);
    float theta = acosf(hmax(-1.0f, hmin(1.0f, v.z)));
    float pp = atan2f(v.y, v.x);
    float phi = pp < 0.0f ? pp + 2.0f * kPi : pp;
    float sint = sinf(theta);
    if (sint == 0.0f) return 0.0f;
    float u = phi / (2.0f * kPi), vv = theta / kPi;
    const float* d = S.dist + l.dist_off;
    int nu = __float_as_int(d[0]), nv = __float_as_int(d[1]);
    const float* marg = d + 2;
    int iu = (int)floorf(u * (float)nu);
    iu = iu < 0 ? 0 : (iu > nu - 1 ? nu - 1 : iu);
    int iv = (int)floorf(vv * (float)nv);
    iv = iv < 0 ? 0 : (iv > nv - 1 ? nv - 1 : iv);
    const float* c = d2_cond(d, iv);
    float mfi = marg[2 + 2 * nv], cfi = c[2 + 2 * nu];
    if (mfi * cfi == 0.0f) return 0.0f;
    float pd = (c[1 + iu] * marg[1 + iv]) / (cfi * mfi);
    return pd / (2.0f * kPi * kPi * sint);
  }
  if constexpr (!has_area_light<LK>()) return 0.0f;
  const DShape& s = l.sh;
  return shape_pdf<LK>(s, xform_point(s.w2o, p), xform_vector(s.w2o, wi));
}
// le light ray (Light.hs:98-106): black for area lights
template <int LK = kAllLights>
__device__ __forceinline__ LRad light_le(const DevPtScene& S, int li, f3 dir) {
  if constexpr (!has_inf_light<LK>()) return lrad(nullptr);
  const DLight& l = S.lights[li];
  if (l.type != LT_INF) return lrad(nullptr);
  f3 wh = norm3(xform_vector(l.w2l, dir));
  float theta = acosf(hmax(-1.0f, hmin(1.0f, wh.z)));
  float pp = atan2f(wh.y, wh.x);
  float phi = pp < 0.0f ? pp + 2.0f * kPi : pp;
  return map_eval<LK>(S, l, phi / (2.0f * kPi), theta / kPi);
}

// ------------------------------------------------------------- camera
__device__ __forceinline__ DRay fire_ray(const DevView& V, float ix, float iy, float lu, float lv) {
  if (V.cam_type == BGPU_CAMERA_ENVIRONMENT) {  // Camera.hs:70-76: the sphere of directions over the raster
    const float t = (kPi * iy) / V.env_sy, p = (kTwoPi * ix) / V.env_sx;
    DRay w;
    w.o = xform_point(V.c2w, mk3(0.0f, 0.0f, 0.0f));
    w.d = xform_vector(V.c2w, mk3(sinf(t) * cosf(p), cosf(t), sinf(t) * sinf(p)));
    w.tmin = 0.0f;
    w.tmax = INFINITY;
    return w;
  }
  f3 pc = xform_point(V.r2c, mk3(ix, iy, 0.0f));  // Camera.hs:49-68
  DRay r;
  r.o = mk3(0.0f, 0.0f, 0.0f);
  r.d = norm3(pc);
  r.tmin = 0.0f;
  r.tmax = INFINITY;
  if (V.lens_radius > 0.0f) {
    float a, b;
    concentric_disk(lu, lv, &a, &b);
    f3 ro = mk3(a * V.lens_radius, b * V.lens_radius, 0.0f);
    f3 pf = ray_at(r, V.focal_distance / r.d.z);
    r.o = ro;
    r.d = norm3(sub3(pf, ro));
  }
  DRay w;
  w.o = xform_point(V.c2w, r.o);
  w.d = xform_vector(V.c2w, r.d);
  w.tmin = r.tmin;
  w.tmax = r.tmax;
  return w;
}

}  // namespace sh
}  // namespace bgpu
