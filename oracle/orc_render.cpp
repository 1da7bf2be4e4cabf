// ORACLE -- test infrastructure only (see orc_math.h).
// CPU restatement of bling's path integrator and everything it calls, per
// camera sample, in the reference's order of operations:
//   Integrator/Path.hs:41-87 (nextVertex)    Scene.hs:61-118 (direct lighting, MIS)
//   Reflection.hs:201-335 (BSDF)             Reflection/{Diffuse,Specular,Microfacet}.hs
//   Fresnel.hs:31-69   Material.hs:32-96     Light.hs:72-229   Shape.hs:314-410
//   Montecarlo.hs:40-186                     Camera.hs:49-68   Image.hs:40-61,162-168,250-299
//   DifferentialGeometry.hs:41-81            Primitive/TriangleMesh.hs:119-207
//   Rendering.hs:111-151 (passes over the sample extent)
//   Integrator/DirectLighting.hs:23-59       Material.hs:43-53 (translucentMatte)
//   Texture.hs:191-219 (checkerBoard, graphPaper selectors)
//   Material.hs:111-127 (mkSubstrate)        Reflection/Microfacet.hs:56-104,136-196
//                                            (FresnelBlend, the anisotropic distribution)
//   SunSky.hs:12-158 (the sun/sky map)       Spectrum.hs:143-159,229-251 (rgbToSpectrum,
//                                            chromaticityToXYZ)
//
// Pins (tests/test_oracle_pins.py):
//   path integrator, NEE + MIS   rho L under a uniform sky; rho Le (R/D)^2 below a
//                                sphere light; sky-only renders return the sky exactly
//   directLighting               the same closed forms, zero on escapes, and
//                                r Le w / r Le (w + 1) through a mirror (maxDepth 1 / >= 2)
//   translucentMatte             (kr + kt (1 - kr)) L inside a uniform sky
//   texture selectors            closed-form points of checkerBoard and graphPaper
//   sampler                      SplitMix64 known answers, permutation bijectivity,
//                                stratification of every table dimension
//   film, camera                 job views equal scenes built at that size
// Pins (tests/test_oracle_microfacet.py):
//   anisotropic distribution     D cos and pdf integrate to 1; ex = ey is Blinn's D,
//                                pdf and sampler in closed form; sampled moments
//   FresnelBlend                 sample = eval / pdf at its own pdf; the coating's
//                                absorption is Beer-Lambert
// Pins (tests/test_oracle_sky.py):
//   sun/sky light map (SunSky.hs)   the zenith with the sun on the horizon; the
//                                black lower half; the sun disk's centre and
//                                bound; sunR black below the horizon and falling
//                                with turbidity and the sun's height
//   texel light maps             rgbfToTexMap's lookup (IO/Bitmap.hs:21-28), in
//                                tests/test_oracle_closed_forms.py
// Parity unpinned: the bits of glibc sinf/cosf/acosf/atan2f/powf/expf as GHC
// calls them (they move sampled directions, compared statistically),
// FresnelBlend's diffuse and specular factors (the reference's own forms),
// and sunR's per-band values beyond its orderings.
#include <array>
#include <atomic>
#include <cmath>
#include <cstring>
#include <vector>

#include "../bling_amd/csrc/spectral_tables.h"
#include "orc_sampler.h"
#include "orc_scene.h"

namespace orc {

// Planted mistakes, for the tests that show each closed-form pin fails when
// the function it pins is wrong (orc_set_mutant; 0 = the restatement as is).
enum {
  MUT_NONE = 0,
  MUT_FR_DIEL_ETA = 1,     // frDielectric: Snell's sin^2 scaled by 1 / etat, not 1 / etat^2
  MUT_FR_COND_K = 2,       // frConductor: k^2 dropped from eta^2 + k^2
  MUT_BLINN_NORM = 3,      // Blinn D: (e + 1) / 2pi instead of (e + 2) / 2pi
  MUT_MF_G = 4,            // Torrance-Sparrow G: factor 2 dropped
  MUT_TRANS_ETA2 = 5,      // specTrans: f scaled by eta instead of eta^2
  MUT_TRANS_DIR = 6,       // specTrans: tangential direction not reversed
  MUT_OREN_SWAP = 7,       // Oren-Nayar: alpha / beta taken from the wrong side
  MUT_ENV_SIN = 8,         // infinite light pdf: the 1 / sin(theta) Jacobian dropped
  MUT_ENV_ROW = 9,         // Dist2D rows built from the map's next row
  MUT_PATH_PC = 10,        // path throughput: 1 / pc dropped
  MUT_PATH_T = 11,         // path throughput: t' = f (the old t dropped)
  MUT_POWER = 12,          // balance heuristic instead of the power heuristic
  MUT_FILTER_CLIP = 13,    // film splats not clipped at the tile origin
  // 14..21: the texture mistakes of orc_texture.h
  MUT_BUMP_FACE = 22,      // bump: the normal not faced toward the geometric one
  MUT_BUMP_SHIFT = 23,     // bump: the u-shifted point moved along dpdv
  MUT_SHINY_SQRT = 24,     // frApproxEta: r' = r, the square root dropped
  MUT_POINT_FALLOFF = 25,  // point light: the 1 / |pos - p|^2 falloff dropped
  MUT_DIR_COS = 26,        // directional light: the |n . d| factor dropped
  MUT_SKY_UP = 27,         // sun/sky: the sky above +z instead of -z (dz = dir.z)
  MUT_SKY_DISK = 28,       // sun/sky: the sun disk tested against sunD, z not flipped
  MUT_SKY_SUN_BELOW = 29,  // sun/sky: the sun's spectrum kept with the sun below the horizon
  MUT_LIGHT_TEXEL = 30     // texel light maps looked up as getPixel (Texture.hs:94-99) instead of rgbfToTexMap
};
static int g_mutant = MUT_NONE;

// ---------------------------------------------------------------- spectra
using Spec = std::array<float, 16>;
static inline Spec sconst(float c) {
  Spec s;
  s.fill(c);
  return s;
}
static inline Spec operator+(const Spec& a, const Spec& b) {
  Spec r;
  for (int i = 0; i < 16; ++i) r[i] = a[i] + b[i];
  return r;
}
static inline Spec operator-(const Spec& a, const Spec& b) {
  Spec r;
  for (int i = 0; i < 16; ++i) r[i] = a[i] - b[i];
  return r;
}
static inline Spec operator*(const Spec& a, const Spec& b) {
  Spec r;
  for (int i = 0; i < 16; ++i) r[i] = a[i] * b[i];
  return r;
}
static inline Spec operator/(const Spec& a, const Spec& b) {
  Spec r;
  for (int i = 0; i < 16; ++i) r[i] = a[i] / b[i];
  return r;
}
static inline Spec sscale(const Spec& a, float f) {  // sScale (Spectrum.hs:448-450)
  Spec r;
  for (int i = 0; i < 16; ++i) r[i] = a[i] * f;
  return r;
}
static inline bool isBlack(const Spec& s) {
  for (float v : s)
    if (v != 0.0f) return false;
  return true;
}
static inline float sY(const Spec& s) {  // Spectrum.hs:371-373
  float acc = 0.0f;
  for (int i = 0; i < 16; ++i) acc = acc + s[i] * BGPU_CIE_Y[i];
  return acc / BGPU_CIE_Y_SUM;
}
static inline void toXYZ(const Spec& s, float* x, float* y, float* z) {  // Spectrum.hs:349-355
  float a = 0, b = 0, c = 0;
  for (int i = 0; i < 16; ++i) {
    a = a + BGPU_CIE_X[i] * s[i];
    b = b + BGPU_CIE_Y[i] * s[i];
    c = c + BGPU_CIE_Z[i] * s[i];
  }
  *x = a / BGPU_CIE_Y_SUM;
  *y = b / BGPU_CIE_Y_SUM;
  *z = c / BGPU_CIE_Y_SUM;
}
static inline Spec specOf(const float* v) {
  Spec s;
  for (int i = 0; i < 16; ++i) s[i] = v[i];
  return s;
}
static inline Spec sclamp(const Spec& s, float lo, float hi) {  // sClamp (Spectrum.hs:453-456)
  Spec r;
  for (int i = 0; i < 16; ++i) r[i] = hmax(lo, hmin(hi, s[i]));
  return r;
}

// ------------------------------------------------------- differential geom
struct DG {
  V3 p, n;
  float u, v;
  V3 dpdu, dpdv;
  bool tri;
  float b1, b2;
};

static DG mkDg(V3 p, float u, float v, V3 dpdu, V3 dpdv) {  // DifferentialGeometry.hs:41-51
  return DG{p, normalize(cross(dpdu, dpdv)), u, v, dpdu, dpdv, false, 0, 0};
}
static DG mkDg1(V3 p, V3 n) {  // mkDg' (DifferentialGeometry.hs:53-56)
  Frame f = coordinateSystem(n);
  return DG{p, n, 0, 0, f.s, f.t, false, 0, 0};
}
static DG transDg(const M4& o2w, const M4& w2o, const DG& d) {  // DifferentialGeometry.hs:72-81
  DG r = d;
  r.p = transPoint(o2w, d.p);
  r.n = normalize(transNormal(w2o, d.n));
  r.dpdu = transVector(o2w, d.dpdu);
  r.dpdv = transVector(o2w, d.dpdv);
  return r;
}

// Object-space DG of a shape hit (Shape.hs:86-229).
static DG shapeDG(const Shape& s, const Ray& ray, const ShapeHit& h) {
  float t = h.t;
  V3 p = rayAt(ray, t);
  switch (s.type) {
    case SH_BOX: {
      int axis = h.axis;
      float half = (s.p[axis] + s.p[3 + axis]) / 2.0f;
      float dir = comp(p, axis) > half ? 1.0f : -1.0f;
      V3 n = normalize(setComp(axis, dir, mk(0, 0, 0)));
      return mkDg1(p, n);
    }
    case SH_CYL: {
      float phimax = s.p[3];
      V3 dpdu = mk(-(phimax * p.y), phimax * p.x, 0.0f);
      V3 dpdv = mk(0.0f, 0.0f, s.p[2] - s.p[1]);
      return mkDg1(p, normalize(cross(dpdu, dpdv)));
    }
    case SH_DISK:
      return mkDg1(p, mk(0.0f, 0.0f, -1.0f));
    case SH_QUAD: {
      float sx = s.p[0], sy = s.p[1];
      float u = (sx + p.x) / (2.0f * sx), v = (sy + p.y) / (2.0f * sy);
      return mkDg(p, u, v, mk(sx, 0, 0), mk(0, sy, 0));
    }
    default: {  // sphere
      float r = s.p[0];
      const float thetaMin = kPi, thetaMax = 0.0f, phiMax = kTwoPi;
      float phi = atan2p(p.y, p.x);
      float u = phi / phiMax;
      float theta = std::acos(clampf(p.z / r, -1.0f, 1.0f));
      float v = (theta - thetaMin) / (thetaMax - thetaMin);
      float zr = std::sqrt(p.x * p.x + p.y * p.y);
      float iz = 1.0f / zr;
      float cosphi = p.x * iz, sinphi = p.y * iz;
      V3 dpdu = mk(-(phiMax * p.y), phiMax * p.x, 0.0f);
      V3 dpdv = scale(mk(p.z * cosphi, p.z * sinphi, -(r * std::sin(theta))), thetaMax - thetaMin);
      return mkDg(p, u, v, dpdu, dpdv);
    }
  }
}

// ------------------------------------------------------------ intersection
struct Isect {
  float t, eps;
  int prim;
  DG dgg, dgs;  // geometric and shading DG (world space)
};

static void triUVs(const Scene& s, int t, float* uv) {
  const bgpu_mesh& m = s.meshes[s.tri_mesh[t]];
  if ((m.flags & BGPU_MESH_UVS) && !s.uvs.empty()) {
    for (int k = 0; k < 3; ++k) {
      int o = s.idx[3 * t + k];
      uv[2 * k] = s.uvs[2 * o];
      uv[2 * k + 1] = s.uvs[2 * o + 1];
    }
  } else {  // triangleDefaultUVs (TriangleMesh.hs:119-120)
    uv[0] = 0; uv[1] = 0; uv[2] = 1; uv[3] = 0; uv[4] = 1; uv[5] = 1;
  }
}

static DG triDG(const Scene& s, int t, const Ray& r, float th, float b1, float b2) {
  V3 p1, p2, p3;
  s.tri(t, &p1, &p2, &p3);
  float uv[6];
  triUVs(s, t, uv);
  float uv00 = uv[0], uv01 = uv[1], uv10 = uv[2], uv11 = uv[3], uv20 = uv[4], uv21 = uv[5];
  V3 e1 = p2 - p1, e2 = p3 - p1;
  V3 n = normalize(cross(e1, e2));
  float du1 = uv00 - uv20, du2 = uv10 - uv20, dv1 = uv01 - uv21, dv2 = uv11 - uv21;
  V3 dp1 = p1 - p3, dp2 = p2 - p3;
  float determinant = du1 * dv2 - dv1 * du2;
  V3 dpdu, dpdv;
  if (determinant == 0.0f) {
    Frame f = coordinateSystem(n);
    dpdu = f.s;
    dpdv = f.t;
  } else {
    float invDet = 1.0f / determinant;
    dpdu = sscale(invDet, sscale(dv2, dp1) - sscale(dv1, dp2));
    dpdv = sscale(invDet, sscale(-du2, dp1) + sscale(du1, dp2));
  }
  float b0 = 1.0f - b1 - b2;
  float tu = b0 * uv00 + b1 * uv10 + b2 * uv20;
  float tv = b0 * uv01 + b1 * uv11 + b2 * uv21;
  DG d = mkDg(rayAt(r, th), tu, tv, dpdu, dpdv);
  d.tri = true;
  d.b1 = b1;
  d.b2 = b2;
  return d;
}

// triangleShadingGeometry with o2w = identity (TriangleMesh.hs:122-134, Primitive.hs:65)
static DG triShading(const Scene& s, int t, const DG& dgg) {
  V3 n0 = s.normals[s.idx[3 * t]], n1 = s.normals[s.idx[3 * t + 1]], n2 = s.normals[s.idx[3 * t + 2]];
  float b0 = 1.0f - dgg.b1 - dgg.b2;
  V3 nsp = sscale(b0, n0) + sscale(dgg.b1, n1) + sscale(dgg.b2, n2);
  M4 id{{1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1}};
  V3 ns = normalize(transNormal(id, nsp));
  V3 ss0 = normalize(dgg.dpdu);
  V3 ts0 = cross(ss0, ns);
  DG r = dgg;
  r.n = ns;
  if (sqLen(ts0) > 0.0f) {
    r.dpdu = cross(normalize(ts0), ns);
    r.dpdv = normalize(ts0);
  } else {
    Frame f = coordinateSystem(ns);
    r.dpdu = f.s;
    r.dpdv = f.t;
  }
  return r;
}

static bool intersectFull(const Scene& s, const Ray& r, Isect* is) {
  Hit h;
  if (!sceneIntersect(s, r, &h, false, nullptr)) return false;
  is->t = h.t;
  is->eps = h.eps;
  is->prim = h.prim;
  if (h.prim < s.ntri) {
    is->dgg = triDG(s, h.prim, r, h.t, h.b1, h.b2);
    const bgpu_mesh& m = s.meshes[s.tri_mesh[h.prim]];
    is->dgs = ((m.flags & BGPU_MESH_NORMALS) && !s.normals.empty()) ? triShading(s, h.prim, is->dgg)
                                                                     : is->dgg;
  } else {
    const ShapePrim& sp = s.shapes[h.prim - s.ntri];
    Ray ro = transRay(sp.w2o, r);
    ShapeHit sh{h.t, h.eps, h.axis};
    is->dgg = transDg(sp.o2w, sp.w2o, shapeDG(sp.shape, ro, sh));
    is->dgs = is->dgg;
  }
  return true;
}

// ------------------------------------------------------------- monte carlo
static inline float powerHeuristic(float fPdf, float gPdf) {  // Montecarlo.hs:113-116
  float f = 1.0f * fPdf, g = 1.0f * gPdf;
  if (g_mutant == MUT_POWER) return f / (f + g);
  return (f * f) / (f * f + g * g);
}
static void concentricSampleDisk(float u1, float u2, float* ox, float* oy) {  // Montecarlo.hs:160-179
  float sx = u1 * 2.0f - 1.0f, sy = u2 * 2.0f - 1.0f;
  if (sx == 0.0f && sy == 0.0f) {
    *ox = 0.0f;
    *oy = 0.0f;
    return;
  }
  float r, th;
  if (sx >= -sy) {
    if (sx > sy) {
      if (sy > 0.0f) { r = sx; th = sy / sx; }
      else { r = sx; th = 8.0f + sy / sx; }
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
  *ox = r * std::cos(theta);
  *oy = r * std::sin(theta);
}
static V3 cosineSampleHemisphere(float u1, float u2) {  // Montecarlo.hs:146-158
  float x, y;
  concentricSampleDisk(u1, u2, &x, &y);
  return mk(x, y, std::sqrt(hmax(0.0f, 1.0f - x * x - y * y)));
}
static V3 uniformSampleCone(const Frame& f, float cosThetaMax, float u1, float u2) {  // :362-373
  float cosTheta = lerpf(u1, cosThetaMax, 1.0f);
  float sinTheta = std::sqrt(1.0f - cosTheta * cosTheta);
  float phi = u2 * kTwoPi;
  return scale(f.s, std::cos(phi) * sinTheta) + scale(f.t, std::sin(phi) * sinTheta) +
         scale(f.n, cosTheta);
}
static float uniformConePdf(float c) { return c >= 1.0f ? 0.0f : 1.0f / (kTwoPi * (1.0f - c)); }
static V3 uniformSampleSphere(float u1, float u2) {  // Montecarlo.hs:181-186
  float u = u1 * 2.0f - 1.0f;
  float s = std::sqrt(1.0f - u * u);
  float omega = u2 * 2.0f * kPi;
  return mk(s * std::cos(omega), s * std::sin(omega), u);
}

// Dist1D / Dist2D (Montecarlo.hs:34-104)
struct Dist1D {
  std::vector<float> func, cdf;
  float fi;
};
static Dist1D mkDist1D(const std::vector<float>& f) {
  Dist1D d;
  d.func = f;
  int n = (int)f.size();
  std::vector<float> i(n + 1);
  i[0] = 0.0f;
  for (int k = 0; k < n; ++k) i[k + 1] = i[k] + f[k] / (float)n;
  d.fi = i[n];
  d.cdf.resize(n + 1);
  for (int k = 0; k <= n; ++k) d.cdf[k] = d.fi != 0.0f ? i[k] / d.fi : (float)k / (float)n;
  return d;
}
static int upperBound(const std::vector<float>& v, float u) {
  int len = (int)v.size(), idx = -1;
  for (int k = 0; k < len; ++k)
    if (v[k] >= u) { idx = k; break; }
  int r = idx < 0 ? len - 1 : idx - 1;
  r = r < 0 ? 0 : r;
  return r < len - 2 ? r : len - 2;
}
static void sampleC1D(const Dist1D& d, float u, float* x, float* pdf, int* off) {
  int o = upperBound(d.cdf, u);
  *pdf = d.fi == 0.0f ? 0.0f : d.func[o] / d.fi;
  float du = (u - d.cdf[o]) / (d.cdf[o + 1] - d.cdf[o]);
  *x = ((float)o + du) / (float)d.func.size();
  *off = o;
}
struct Dist2D {
  std::vector<Dist1D> cond;
  Dist1D marg;
};

// ------------------------------------------------------------------ lights
// The Perez sun/sky model's data (SkyData and sunSpectrum', SunSky.hs:29-155)
struct SkyData {
  V3 sunDir;       // normalize (worldToLocal basis (normalize sdw))
  float sunTheta;
  float px[5], py[5], pY[5];
  float zx, zy, zY;
  Spec sunR;
  V3 diskDir;      // normalize (worldToLocal basis sdw), what sunSpectrum tests against
};

struct LightRec {
  int type;  // 0 area (shape index), 1 infinite, 2 point, 3 directional
  int shape;
  int id;
  Spec L;  // area: radiance; point, directional: intensity
  V3 vec;  // point: position; directional: normalize n
  // infinite
  int w, h;
  const float* map;  // texel maps
  bool sky;          // a sun/sky map (mkSunSkyLight)
  SkyData sd;
  M4 w2l, l2w;
  Dist2D dist;
};

// rgbToSpectrum over the illuminant bases (Spectrum.hs:143-159)
static Spec rgbIllum(float r, float g, float b) {
  const Spec wb = specOf(BGPU_RGB_ILLUM_WHITE), cb = specOf(BGPU_RGB_ILLUM_CYAN),
             mb = specOf(BGPU_RGB_ILLUM_MAGENTA), yb = specOf(BGPU_RGB_ILLUM_YELLOW),
             rb = specOf(BGPU_RGB_ILLUM_RED), gb = specOf(BGPU_RGB_ILLUM_GREEN), bb = specOf(BGPU_RGB_ILLUM_BLUE);
  if (r <= g && r <= b)
    return sscale(wb, r) + (g <= b ? sscale(cb, g - r) + sscale(bb, b - g) : sscale(cb, b - r) + sscale(gb, g - b));
  if (g <= r && g <= b)
    return sscale(wb, g) + (r <= b ? sscale(mb, r - g) + sscale(bb, b - r) : sscale(mb, b - g) + sscale(rb, r - b));
  return sscale(wb, b) + (r <= b ? sscale(yb, r - b) + sscale(gb, g - r) : sscale(yb, g - b) + sscale(rb, r - g));
}

static float evalRegular(const float* a, int n, float l0, float l1, float l) {  // Spectrum.hs:271-280
  if (l <= l0) return a[0];
  if (l >= l1) return a[n - 1];
  float d1 = 1.0f / ((l1 - l0) / (float)(n - 1));
  float x = (l - l0) * d1;
  int b0 = (int)std::floor(x);
  int b1 = b0 + 1 < n - 1 ? b0 + 1 : n - 1;
  float dx = x - (float)b0;
  return (1.0f - dx) * a[b0] + dx * a[b1];
}

static float evalIrregular(const float* ls, const float* vs, int n, float l) {  // Spectrum.hs:258-269
  if (l <= ls[0]) return vs[0];
  if (l >= ls[n - 1]) return vs[n - 1];
  // fi lo hi: binary search for the interval's index
  int lo = 0, hi = n - 1, i;
  for (;;) {
    int mid = (lo + hi) / 2;
    if (lo == mid) { i = lo; break; }
    if (ls[mid] == l) { i = mid; break; }
    if (ls[mid] < l) lo = mid;
    else hi = mid;
  }
  float t = (l - ls[i]) / (ls[i + 1] - ls[i]);
  return (1.0f - t) * vs[i] + t * vs[i + 1];
}

static SkyData initSky(V3 east, V3 sdw, float t) {  // mkSunSkyLight / initSky (SunSky.hs:12-74)
  SkyData k;
  V3 w = normalize(normalize(mk(0, 1, 0)));  // coordinateSystem' (normalize up) (normalize east)
  V3 u = normalize(cross(normalize(east), w));
  V3 v = cross(w, u);
  auto toLocal = [&](V3 d) { return mk(dot(d, u), dot(d, v), dot(d, w)); };
  k.sunDir = normalize(toLocal(normalize(sdw)));
  k.diskDir = normalize(toLocal(sdw));
  float st = std::acos(clampf(k.sunDir.z, -1.0f, 1.0f));
  k.sunTheta = st;
  float st2 = st * st, st3 = st * st * st, t2 = t * t;
  float chi = (4.0f / 9.0f - t / 120.0f) * (kPi - 2.0f * st);
  const float pY[5] = {0.17872f * t - 1.46303f, -0.35540f * t + 0.42749f, -0.02266f * t + 5.32505f,
                       0.12064f * t - 2.57705f, -0.06696f * t + 0.37027f};
  const float px[5] = {-0.01925f * t - 0.25922f, -0.06651f * t + 0.00081f, -0.00041f * t + 0.21247f,
                       -0.06409f * t - 0.89887f, -0.00325f * t + 0.04517f};
  const float py[5] = {-0.01669f * t - 0.26078f, -0.09495f * t + 0.00921f, -0.00792f * t + 0.21023f,
                       -0.04405f * t - 1.65369f, -0.01092f * t + 0.05291f};
  for (int i = 0; i < 5; ++i) k.pY[i] = pY[i], k.px[i] = px[i], k.py[i] = py[i];
  k.zY = ((4.04530f * t - 4.97100f) * std::tan(chi) - 0.2155f * t + 2.4192f) * 1000.0f;
  k.zx = (0.00165f * st3 - 0.00374f * st2 + 0.00208f * st) * t2 +
         (-0.02902f * st3 + 0.06377f * st2 - 0.03202f * st + 0.00394f) * t +
         (0.11693f * st3 - 0.21196f * st2 + 0.06052f * st + 0.25885f);
  k.zy = (0.00275f * st3 - 0.00610f * st2 + 0.00316f * st) * t2 +
         (-0.04212f * st3 + 0.08970f * st2 - 0.04153f * st + 0.00515f) * t +
         (0.15346f * st3 - 0.26756f * st2 + 0.06669f * st + 0.26688f);
  // sunSpectrum' (SunSky.hs:115-155): black with the sun below the horizon
  k.sunR = sconst(0.0f);
  if (!(k.sunDir.z < 0.0f) || g_mutant == MUT_SKY_SUN_BELOW) {
    float m = 1.0f / (std::cos(st) + 0.000940f * std::pow(1.6386f - st, -1.253f));  // relative optical mass
    auto sf = [&](float l) {
      float tR = std::exp(-m * 0.008735f * std::pow(l / 1000.0f, -4.08f));
      float beta = 0.04608365822050f * t - 0.04586025928522f;
      float tA = std::exp(-m * beta * std::pow(l / 1000.0f, -1.3f));
      float tO = std::exp(-m * evalIrregular(BGPU_SKY_KO_L, BGPU_SKY_KO_A, 64, l) * 0.35f);
      float kg = evalIrregular(BGPU_SKY_KG_L, BGPU_SKY_KG_A, 4, l);
      float tG = std::exp(-1.41f * kg * m / std::pow(1.0f + 118.93f * kg * m, 0.45f));
      float kwa = evalIrregular(BGPU_SKY_KWA_L, BGPU_SKY_KWA_A, 13, l);
      float tWA = std::exp(-0.2385f * kwa * 2.0f * m / std::pow(1.0f + 20.07f * kwa * 2.0f * m, 0.45f));
      return evalRegular(BGPU_SKY_SOL, 38, 380.0f, 750.0f, l) * tR * tA * tO * tG * tWA;
    };
    for (int i = 0; i < 16; ++i) {  // fromSpd: (f l0 + f l1) / 2 per band (Spectrum.hs:316,329-335)
      float a = (float)i / 16.0f, b = (float)(i + 1) / 16.0f;
      float l0 = (1.0f - a) * 400.0f + a * 700.0f, l1 = (1.0f - b) * 400.0f + b * 700.0f;
      k.sunR[i] = (sf(l0) + sf(l1)) * 0.5f;
    }
  }
  return k;
}

static float perez(const float* p, float sunT, float t, float g, float lvz) {  // SunSky.hs:99-104
  float num = (1.0f + p[0] * std::exp(p[1] / std::cos(t))) * (1.0f + p[2] * std::exp(p[3] * g)) +
              p[4] * std::cos(g) * std::cos(g);
  float den = (1.0f + p[0] * std::exp(p[1])) * (1.0f + p[2] * std::exp(p[3] * sunT)) +
              p[4] * std::cos(sunT) * std::cos(sunT);
  return lvz * num / den;
}

static Spec skyEval(const SkyData& k, float u, float v) {  // mkSunSkyLight's eval (SunSky.hs:17-20)
  float phi = u * 2.0f * kPi, th = v * kPi;  // cartToSph
  V3 dir = mk(std::sin(th) * std::cos(phi), std::sin(th) * std::sin(phi), std::cos(th));  // sphToDir
  Spec sky = sconst(0.0f);
  float dz = g_mutant == MUT_SKY_UP ? dir.z : -dir.z;  // skySpectrum (SunSky.hs:80-96)
  if (!(dz < 1e-4f)) {
    float theta = std::acos(dz);
    float gamma = std::acos(clampf(dot(dir, k.sunDir), -1.0f, 1.0f));
    float x = perez(k.px, k.sunTheta, theta, gamma, k.zx);
    float y = perez(k.py, k.sunTheta, theta, gamma, k.zy);
    float yy = perez(k.pY, k.sunTheta, theta, gamma, k.zY) * 1e-4f;
    // chromaticityToXYZ (Spectrum.hs:229-251)
    float den = 0.0241f + 0.2562f * x - 0.7341f * y;
    float m1 = (-1.3515f - 1.7703f * x + 5.9114f * y) / den;
    float m2 = (0.03f - 31.4424f * x + 30.0717f * y) / den;
    float cx = BGPU_S0_XYZ[0] + m1 * BGPU_S1_XYZ[0] + m2 * BGPU_S2_XYZ[0];
    float cy = BGPU_S0_XYZ[1] + m1 * BGPU_S1_XYZ[1] + m2 * BGPU_S2_XYZ[1];
    float cz = BGPU_S0_XYZ[2] + m1 * BGPU_S1_XYZ[2] + m2 * BGPU_S2_XYZ[2];
    float X = cx * yy / cy, Z = cz * yy / cy;
    // xyzToSpectrum = rgbToSpectrumIllum . xyzToRgb (Spectrum.hs:162-168,357-358)
    float r = 3.240479f * X - 1.537150f * yy - 0.498535f * Z;
    float g = -0.969256f * X + 1.875991f * yy + 0.041556f * Z;
    float b = 0.055648f * X - 0.204043f * yy + 1.057311f * Z;
    sky = rgbIllum(r, g, b);
  }
  // sunSpectrum (SunSky.hs:106-113)
  V3 sd = g_mutant == MUT_SKY_DISK ? k.diskDir : mk(k.diskDir.x, k.diskDir.y, -k.diskDir.z);
  float stm = std::sqrt(hmax(0.0f, 1.0f - 6.955e5f / 1.496e8f));  // sunThetaMax2 (SunSky.hs:42-47)
  return sky + (dot(sd, dir) > stm ? k.sunR : sconst(0.0f));
}

// texMapEval of an infinite light's map at (u, v): a sun/sky map, or a texel
// map looked up as rgbfToTexMap does (IO/Bitmap.hs:21-28)
static Spec texEval(const LightRec& l, float u, float v) {
  if (l.sky) return skyEval(l.sd, u, v);
  if (g_mutant == MUT_LIGHT_TEXEL) {
    auto md = [](int a, int b) { int r = a % b; return r < 0 ? r + b : r; };
    int px = md((int)std::floor(u * (float)l.w), l.w);
    int py = md((int)std::floor((-v) * (float)l.h), l.h);
    return specOf(l.map + 16 * ((size_t)py * l.w + px));
  }
  int px = (int)std::floor((1.0f - u) * (float)l.w);
  int py = (int)std::floor((1.0f - v) * (float)l.h);
  px = px < 0 ? 0 : (px > l.w - 1 ? l.w - 1 : px);
  py = py < 0 ? 0 : (py > l.h - 1 ? l.h - 1 : py);
  return specOf(l.map + 16 * ((size_t)py * l.w + px));
}

static Dist2D buildDist(const LightRec& l) {  // mkDist2D (texSize rmap) (sY . eval)  (Light.hs:76-77)
  Dist2D d;
  float sx = (float)l.w, sy = (float)l.h;
  std::vector<float> fi;
  for (int v = 0; v < l.h; ++v) {
    std::vector<float> row(l.w);
    const int vr = g_mutant == MUT_ENV_ROW ? (v + 1) % l.h : v;
    for (int u = 0; u < l.w; ++u) row[u] = sY(texEval(l, (float)u / sx, (float)vr / sy));
    d.cond.push_back(mkDist1D(row));
    fi.push_back(d.cond.back().fi);
  }
  d.marg = mkDist1D(fi);
  return d;
}

struct LightSample {
  Spec li;
  V3 wi;
  Ray ray;
  float pdf;
  bool delta = false;
};

static float shapeArea(const Shape& s) {  // Shape.hs:314-328
  switch (s.type) {
    case SH_BOX: {
      float h = s.p[3] - s.p[0], w = s.p[4] - s.p[1], l = s.p[5] - s.p[2];
      return 2.0f * (h * w + h * l + w * l);
    }
    case SH_CYL: return 2.0f * kPi * s.p[0] * (s.p[2] - s.p[1]);
    case SH_DISK: return kPi * (s.p[1] * s.p[1] - s.p[2] * s.p[2]);
    case SH_QUAD: return 4.0f * s.p[0] * s.p[1];
    default: return s.p[0] * s.p[0] * 4.0f * kPi;
  }
}
static bool insideSphere(float r, V3 p) { return sqLen(p) - r * r < 1e-4f; }  // Shape.hs:330-331

static float generalPdf(const Shape& s, V3 p, V3 wi) {  // Shape.hs:346-350
  Ray r{p, wi, 1e-3f, kInf};
  ShapeHit h;
  if (!shapeIntersect(s, r, &h)) return 0.0f;
  DG d = shapeDG(s, r, h);
  float pd = sqLen(p - rayAt(r, h.t)) / (absDot(d.n, -wi) * shapeArea(s));
  return std::isinf(pd) ? 0.0f : pd;
}
static float shapePdf(const Shape& s, V3 p, V3 wi) {  // Shape.hs:333-344
  if (s.type == SH_SPHERE && !insideSphere(s.p[0], p)) {
    float r = s.p[0];
    float sinThetaMax2 = r * r / sqLen(p);
    float cosThetaMax = std::sqrt(hmax(0.0f, 1.0f - sinThetaMax2));
    return uniformConePdf(cosThetaMax);
  }
  return generalPdf(s, p, wi);
}
static void remapRand(int segs, float u, int* seg, float* u2) {  // Math.hs:113-121
  float fs = (float)segs;
  int sg = (int)std::floor(u * fs);
  sg = segs - 1 < sg ? segs - 1 : sg;
  *seg = sg;
  *u2 = (u - (float)sg / fs) * fs;
}
static void sampleShape1(const Shape& s, float u1, float u2, V3* p, V3* n) {  // Shape.hs:379-409
  switch (s.type) {
    case SH_BOX: {
      int axis, nf;
      float u1p, u2p;
      remapRand(3, u1, &axis, &u1p);
      remapRand(2, u2, &nf, &u2p);
      *n = setComp(axis, (float)nf * 2.0f - 1.0f, mk(0, 0, 0));
      int oa0 = (axis + 1) % 3, oa1 = (axis + 2) % 3;
      V3 pmin = mk(s.p[0], s.p[1], s.p[2]), pmax = mk(s.p[3], s.p[4], s.p[5]);
      V3 base = nf == 0 ? pmin : pmax;
      base = setComp(oa1, lerpf(u2p, comp(pmin, oa1), comp(pmax, oa1)), base);
      *p = setComp(oa0, lerpf(u1p, comp(pmin, oa0), comp(pmax, oa0)), base);
      return;
    }
    case SH_CYL: {
      float r = s.p[0];
      float z = lerpf(u1, s.p[1], s.p[2]);
      float phi = lerpf(u2, 0.0f, kTwoPi);
      *p = mk(r * std::cos(phi), r * std::sin(phi), z);
      *n = normalize(mk(p->x, p->y, 0.0f));
      return;
    }
    case SH_DISK: {
      float r = lerpf(u1, s.p[2], s.p[1]);
      float phi = lerpf(u2, 0.0f, s.p[3]);
      *p = mk(r * std::cos(phi), r * std::sin(phi), s.p[0]);
      *n = mk(0, 0, -1);
      return;
    }
    case SH_QUAD:
      *p = mk(lerpf(u1, -s.p[0], s.p[0]), lerpf(u2, -s.p[1], s.p[1]), 0.0f);
      *n = mk(0, 0, -1);
      return;
    default: {
      V3 q = uniformSampleSphere(u1, u2);
      *p = scale(q, s.p[0]);
      *n = q;
      return;
    }
  }
}
static void sampleShape(const Shape& s, V3 p, float u1, float u2, V3* ps, V3* ns) {  // :362-376
  if (s.type == SH_SPHERE && !insideSphere(s.p[0], p)) {
    float r = s.p[0];
    V3 dn = normalize(-p);
    Frame cs = coordinateSystem(dn);
    float cosThetaMax = std::sqrt(hmax(0.0f, 1.0f - (r * r) / sqLen(p)));
    V3 d = uniformSampleCone(cs, cosThetaMax, u1, u2);
    Ray ray{p, d, 0.0f, kInf};
    ShapeHit h;
    V3 q = shapeIntersect(s, ray, &h) ? rayAt(ray, h.t) : scale(dn, r);
    *ps = q;
    *ns = normalize(q);
    return;
  }
  sampleShape1(s, u1, u2, ps, ns);
}

struct Ctx {
  const Scene* s;
  std::vector<LightRec> lights;
  std::vector<int> shapeLight;  // shape index -> light index or -1
  float filter[256];
  float fw, fh;
  std::atomic<uint64_t> rays{0};
};

static LightSample sampleLight(const Ctx& c, const LightRec& l, V3 pW, float e, V3 n, float u1, float u2) {
  LightSample ls;
  if (l.type == 2) {  // PointLight (Light.hs:146-150)
    const V3 d = l.vec - pW;
    ls.li = g_mutant == MUT_POINT_FALLOFF ? l.L : sscale(l.L, 1.0f / sqLen(d));
    ls.wi = normalize(d);
    ls.ray = Ray{pW, d, e, kInf};  // along pos - p, not normalised, to infinity
    ls.pdf = 1.0f;
    ls.delta = true;
    return ls;
  }
  if (l.type == 3) {  // Directional (Light.hs:142-144)
    ls.li = g_mutant == MUT_DIR_COS ? l.L : sscale(l.L, std::fabs(dot(n, l.vec)));
    ls.wi = l.vec;
    ls.ray = Ray{pW, l.vec, e, kInf};
    ls.pdf = 1.0f;
    ls.delta = true;
    return ls;
  }
  if (l.type == 1) {  // Light.hs:130-141
    float uv[2], mapPdf;
    float pdf1, pdf0;
    int off;
    sampleC1D(l.dist.marg, u2, &uv[1], &pdf1, &off);
    int dummy;
    sampleC1D(l.dist.cond[off], u1, &uv[0], &pdf0, &dummy);
    mapPdf = pdf0 * pdf1;
    Ray r0{mk(0, 0, 0), mk(0, 1, 0), 0.0f, 1.0f};
    if (mapPdf == 0.0f) return LightSample{sconst(0), mk(0, 1, 0), r0, 0.0f};
    float phi = uv[0] * 2.0f * kPi, theta = uv[1] * kPi;  // cartToSph (Types.hs:31-33)
    float sint = std::sin(theta);
    if (sint == 0.0f) return LightSample{sconst(0), mk(0, 1, 0), r0, 0.0f};
    V3 dl = mk(std::sin(theta) * std::cos(phi), std::sin(theta) * std::sin(phi), std::cos(theta));
    V3 wi = transVector(l.l2w, dl);
    ls.li = texEval(l, uv[0], uv[1]);
    ls.wi = wi;
    ls.ray = Ray{pW, wi, e, kInf};
    ls.pdf = g_mutant == MUT_ENV_SIN ? mapPdf / (2.0f * kPi * kPi) : mapPdf / (2.0f * kPi * kPi * sint);
    return ls;
  }
  const ShapePrim& sp = c.s->shapes[l.shape];  // Light.hs:152-160
  V3 p = transPoint(sp.w2o, pW);
  V3 ps, ns;
  sampleShape(sp.shape, p, u1, u2, &ps, &ns);
  V3 wi = normalize(ps - p);
  ls.li = dot(ns, wi) < 0.0f ? l.L : sconst(0);
  ls.wi = transVector(sp.o2w, wi);
  ls.pdf = shapePdf(sp.shape, p, wi);
  Ray ray{p, wi, e, len(ps - p) - e};
  ls.ray = transRay(sp.o2w, ray);
  return ls;
}

static float lightPdf(const Ctx& c, const LightRec& l, V3 p, V3 wi) {  // Light.hs:215-229
  if (l.type == 2 || l.type == 3) return 0.0f;  // point, directional: no direction is sampled by chance
  if (l.type == 1) {
    V3 v = transVector(l.w2l, wi);
    float theta = std::acos(hmax(-1.0f, hmin(1.0f, v.z)));  // sphericalTheta (Math.hs:159-161)
    float pp = std::atan2(v.y, v.x);                        // sphericalPhi (Math.hs:163-169)
    float phi = pp < 0.0f ? pp + 2.0f * kPi : pp;
    float sint = std::sin(theta);
    if (sint == 0.0f) return 0.0f;
    float u = phi / (2.0f * kPi), vv = theta / kPi;
    const Dist2D& d = l.dist;
    int nu = (int)d.cond[0].func.size(), nv = (int)d.marg.func.size();
    int iu = (int)std::floor(u * (float)nu);
    iu = iu < 0 ? 0 : (iu > nu - 1 ? nu - 1 : iu);
    int iv = (int)std::floor(vv * (float)nv);
    iv = iv < 0 ? 0 : (iv > nv - 1 ? nv - 1 : iv);
    float pd;
    if (d.marg.fi * d.cond[iv].fi == 0.0f) pd = 0.0f;
    else pd = (d.cond[iv].func[iu] * d.marg.func[iv]) / (d.cond[iv].fi * d.marg.fi);
    return g_mutant == MUT_ENV_SIN ? pd / (2.0f * kPi * kPi) : pd / (2.0f * kPi * kPi * sint);
  }
  const ShapePrim& sp = c.s->shapes[l.shape];
  return shapePdf(sp.shape, transPoint(sp.w2o, p), transVector(sp.w2o, wi));
}

static Spec lightLe(const LightRec& l, const Ray& r) {  // Light.hs:98-106
  if (l.type != 1) return sconst(0);
  V3 wh = normalize(transVector(l.w2l, r.d));
  float theta = std::acos(hmax(-1.0f, hmin(1.0f, wh.z)));
  float pp = std::atan2(wh.y, wh.x);
  float phi = pp < 0.0f ? pp + 2.0f * kPi : pp;
  return texEval(l, phi / (2.0f * kPi), theta / kPi);
}

// -------------------------------------------------------------------- BSDF
enum { BX_R = 1, BX_T = 2, BX_D = 4, BX_G = 8, BX_S = 16 };
enum { K_LAMBERT, K_OREN, K_MF, K_SPECR, K_SPECT, K_FB };
enum { FR_NOOP, FR_DIEL, FR_COND };

struct Bxdf {
  int kind, type;
  Spec r;
  float A, B;   // Oren-Nayar
  float e;      // Blinn exponent
  int fr;
  float ei, et; // dielectric Fresnel / transmission etas
  Spec eta, k;  // conductor
  int flip;     // brdfToBtdf (Reflection.hs:188-195): wi mirrored to the other hemisphere
  // FresnelBlend (Microfacet.hs:56-104): r is the diffuse layer rd
  Spec rs, ra;              // specular reflectance, absorption
  float ex, ey, depth;      // anisotropic exponents, coating thickness
};

struct Bsdf {
  int n;
  Bxdf c[2];
  Frame cs;
  V3 p, ng;
};

static inline float cosT(V3 v) { return v.z; }
static inline float absCosT(V3 v) { return std::fabs(v.z); }
static inline float sinT2(V3 v) { return hmax(0.0f, 1.0f - v.z * v.z); }
static inline float sinT(V3 v) { return std::sqrt(sinT2(v)); }
static inline float cosPhi(V3 v) {
  float st = sinT(v);
  return st == 0.0f ? 1.0f : clampf(v.x / st, -1.0f, 1.0f);
}
static inline float sinPhi(V3 v) {
  float st = sinT(v);
  return st == 0.0f ? 0.0f : clampf(v.y / st, -1.0f, 1.0f);
}
static inline bool sameHemi(V3 a, V3 b) { return a.z * b.z > 0.0f; }
static inline V3 toSameHemi(V3 wo, V3 wi) { return wo.z < 0.0f ? mk(wi.x, wi.y, -wi.z) : wi; }

static Spec frDielectric(float etai, float etat, float cosi) {  // Fresnel.hs:21-55
  float c = hmax(0.0f, 1.0f - cosi * cosi);
  const float e2 = g_mutant == MUT_FR_DIEL_ETA ? etat : etat * etat;
  float costp = cosi > 0.0f ? c / e2 : c * e2;
  float cost = std::sqrt(1.0f - clampf(costp, 0.0f, 1.0f));
  float ac = std::fabs(cosi);
  float eta = etat / etai;
  float rParlp = eta * ac;
  float rParl = (cost - rParlp) / (cost + rParlp);
  float rPerpp = eta * cost;
  float rPerp = (ac - rPerpp) / (ac + rPerpp);
  return sconst((rParl * rParl + rPerp * rPerp) * 0.5f);
}
static Spec frConductor(const Spec& eta, const Spec& k, float cosi) {  // Fresnel.hs:58-70
  float ac = std::fabs(cosi);
  Spec tmpF = g_mutant == MUT_FR_COND_K ? eta * eta : eta * eta + k * k;
  Spec tmp = sscale(tmpF, ac * ac);
  Spec ec2 = sscale(eta, 2.0f * ac);
  Spec c2 = sconst(ac * ac);
  Spec one = sconst(1.0f);
  Spec rPer2 = (tmpF - ec2 + c2) / (tmpF + ec2 + c2);
  Spec rPar2 = (tmp - ec2 + one) / (tmp + ec2 + one);
  return (rPer2 + rPar2) / sconst(2.0f);
}
static Spec fresnel(const Bxdf& b, float cosi) {
  if (b.fr == FR_DIEL) return frDielectric(b.ei, b.et, cosi);
  if (b.fr == FR_COND) return frConductor(b.eta, b.k, cosi);
  return sconst(1.0f);
}

static float cosPdf(V3 wo, V3 wi) { return sameHemi(wo, wi) ? kInvPi * absCosT(wi) : 0.0f; }

static float mfG(V3 wo, V3 wi, V3 wh) {  // Microfacet.hs:113-120
  float nDotWh = absCosT(wh), nDotWo = absCosT(wo), nDotWi = absCosT(wi), woDotWh = absDot(wo, wh);
  const float two = g_mutant == MUT_MF_G ? 1.0f : 2.0f;
  return hmin(1.0f, hmin(two * nDotWh * nDotWo / woDotWh, two * nDotWh * nDotWi / woDotWh));
}
static float blinnD(float e, V3 wh) {
  return (g_mutant == MUT_BLINN_NORM ? e + 1.0f : e + 2.0f) * kInvTwoPi * std::pow(absCosT(wh), e);
}
static float blinnPdf(float e, V3 wh) { return (e + 1.0f) * std::pow(absCosT(wh), e) * kInvTwoPi; }

// The anisotropic distribution (Microfacet.hs:136-196).
static float anisoExponent(float ex, float ey, V3 wh, float den) {
  return (ex * wh.x * wh.x + ey * wh.y * wh.y) / den;
}
static float anisoD(float ex, float ey, V3 wh) {  // mfDistD
  float costh = absCosT(wh);
  float d = 1.0f - costh * costh;
  if (d == 0.0f) return 0.0f;
  return std::sqrt((ex + 2.0f) * (ey + 2.0f)) * kInvTwoPi * std::pow(costh, anisoExponent(ex, ey, wh, d));
}
static float anisoPdf(float ex, float ey, V3 wh) {  // mfDistPdf
  float costh = absCosT(wh);
  float e = anisoExponent(ex, ey, wh, hmax(0.0f, 1.0f - costh * costh));
  return std::sqrt((ex + 1.0f) * (ey + 1.0f)) * kInvTwoPi * std::pow(costh, e);
}
static V3 anisoSample(float ex, float ey, float u1, float u2, float* dOut, float* pdfOut) {  // mfDistSample
  // smpFirstQuadrand, mirrored into the quadrant that u1 selects
  auto first = [&](float u, float* p, float* c) {
    *p = ex == ey ? kPi * u * 0.5f : std::atan(std::sqrt((ex + 1.0f) / (ey + 1.0f)) * std::tan(kPi * u * 0.5f));
    float cp = std::cos(*p), sp = std::sin(*p);
    *c = std::pow(u2, 1.0f / (ex * cp * cp + ey * sp * sp + 1.0f));
  };
  float p, cost, phi;
  if (u1 < 0.25f) {
    first(4.0f * u1, &p, &cost);
    phi = p;
  } else if (u1 < 0.5f) {
    first(4.0f * (0.5f - u1), &p, &cost);
    phi = kPi - p;
  } else if (u1 < 0.75f) {
    first(4.0f * (u1 - 0.5f), &p, &cost);
    phi = p + kPi;
  } else {
    first(4.0f * (1.0f - u1), &p, &cost);
    phi = kTwoPi - p;
  }
  float sint = std::sqrt(hmax(0.0f, 1.0f - cost * cost));
  V3 wh = mk(sint * std::cos(phi), sint * std::sin(phi), cost);
  float f = kInvTwoPi * std::pow(cost, anisoExponent(ex, ey, wh, 1.0f - cost * cost));
  *dOut = std::sqrt((ex + 2.0f) * (ey + 2.0f)) * f;
  *pdfOut = std::sqrt((ex + 1.0f) * (ey + 1.0f)) * f;
  return wh;
}
// FresnelBlend's half vector: normalize (wi + wo), turned into the upper hemisphere
static V3 fbHalf(V3 wi, V3 wo) {
  V3 wh = normalize(wi + wo);
  return wh.z < 0.0f ? -wh : wh;
}
static Spec fbEval(const Bxdf& b, V3 wo, V3 wi) {  // e (Microfacet.hs:65-86)
  float costi = absCosT(wi), costo = absCosT(wo);
  Spec white = sconst(1.0f);
  Spec a = white;
  if (b.depth > 0.0f) {  // absorption in the coating
    Spec x = sscale(b.ra, -b.depth * (costi + costo) / (costi * costo));
    for (int k = 0; k < 16; ++k) a[k] = std::exp(x[k]);
  }
  Spec diff = sscale(a * b.r * (white - b.rs), (costo * 28.0f / 23.0f * kPi) *
                                                   (1.0f - std::pow(1.0f - 0.5f * costi, 5.0f)) *
                                                   (1.0f - std::pow(1.0f - 0.5f * costo, 5.0f)));
  V3 wh = fbHalf(wi, wo);
  float costih = absDot(wi, wh);
  Spec schlick = b.rs + sscale(white - b.rs, std::pow(1.0f - costih, 5.0f));
  Spec spec = sscale(schlick, anisoD(b.ex, b.ey, wh) * costo / (4.0f * costih * hmax(costi, costo)));
  return diff + spec;
}
static float fbPdf(const Bxdf& b, V3 wo, V3 wi) {  // p (Microfacet.hs:101-104)
  if (!sameHemi(wo, wi)) return 0.0f;
  V3 wh = fbHalf(wi, wo);
  return 0.5f * (absCosT(wi) * kInvPi + anisoPdf(b.ex, b.ey, wh) / (4.0f * absDot(wo, wh)));
}

static Spec orenNayar(const Bxdf& b, V3 wo, V3 wi) {  // Diffuse.hs:53-65
  float sinti = sinT(wi), sinto = sinT(wo);
  float maxcos = 0.0f;
  if (sinti > 1e-4f && sinto > 1e-4f)
    maxcos = hmax(0.0f, cosPhi(wi) * cosPhi(wo) + sinPhi(wi) * sinPhi(wo));
  float sina, tanb;
  if ((absCosT(wi) > absCosT(wo)) != (g_mutant == MUT_OREN_SWAP)) { sina = sinto; tanb = sinti / absCosT(wi); }
  else { sina = sinti; tanb = sinto / absCosT(wo); }
  return sscale(b.r, b.A + b.B * maxcos * sina * tanb);
}

static Spec bxEval(const Bxdf& b, V3 wo, V3 wi) {
  if (b.flip) wi.z = -wi.z;
  switch (b.kind) {
    case K_LAMBERT: return sscale(b.r, kInvPi * absCosT(wo));
    case K_OREN: return sscale(orenNayar(b, wo, wi), kInvPi * absCosT(wo));
    case K_MF: {  // Microfacet.hs:21-32
      float costo = absCosT(wo), costi = absCosT(wi);
      if (costi == 0.0f || costo == 0.0f) return sconst(0);
      V3 whp = wi + wo;
      if (whp.x == 0.0f && whp.y == 0.0f && whp.z == 0.0f) return sconst(0);
      V3 wh = normalize(whp);
      if (cosT(wh) < 0.0f) return sconst(0);
      float costh = dot(wi, wh);
      float x = blinnD(b.e, wh) * mfG(wo, wi, wh) / (4.0f * costi);
      return sscale(b.r * fresnel(b, costh), x);
    }
    case K_FB: return fbEval(b, wo, wi);
    default: return sconst(0);
  }
}
static float bxPdf(const Bxdf& b, V3 wo, V3 wi) {
  if (b.flip) wi.z = -wi.z;
  switch (b.kind) {
    case K_LAMBERT:
    case K_OREN: return cosPdf(wo, wi);
    case K_MF: {  // Microfacet.hs:34-40
      V3 whp = wo + wi;
      if (sqLen(whp) == 0.0f) return 0.0f;
      V3 wh = normalize(whp);
      if (cosT(wh) < 0.0f) return 0.0f;
      return blinnPdf(b.e, wh) / (4.0f * absDot(wo, wh));
    }
    case K_FB: return fbPdf(b, wo, wi);
    default: return 0.0f;
  }
}
// bxdfSample with adj = False
static void bxSample(const Bxdf& b, V3 wo, float u1, float u2, Spec* f, V3* wi, float* pdf) {
  switch (b.kind) {
    case K_LAMBERT:
    case K_OREN: {  // Diffuse.hs:14-22,38-42
      V3 w = toSameHemi(wo, cosineSampleHemisphere(u1, u2));
      if (sameHemi(wo, w)) {
        *f = b.kind == K_LAMBERT ? b.r : orenNayar(b, wo, w);
        *wi = w;
        *pdf = cosPdf(wo, w);
      } else {
        *f = sconst(0);
        *wi = b.kind == K_LAMBERT ? wo : w;
        *pdf = 0.0f;
      }
      if (b.flip) wi->z = -wi->z;
      return;
    }
    case K_MF: {  // Microfacet.hs:42-54,175-182
      float cost = std::pow(u1, 1.0f / (b.e + 1.0f));
      float sint = std::sqrt(hmax(0.0f, 1.0f - cost * cost));
      float phi = u2 * 2.0f * kPi;
      V3 whp = mk(sint * std::cos(phi), sint * std::sin(phi), cost);
      float ff = std::pow(cost, b.e) * kInvTwoPi;
      float d = (b.e + 2.0f) * ff;
      float pd = (b.e + 1.0f) * ff;
      V3 wh = cosT(whp) < 0.0f ? -whp : whp;
      V3 w = sscale(2.0f * dot(wo, wh), wh) - wo;
      float costH = dot(wo, wh);
      float fact = d * std::fabs(costH) / pd * mfG(wo, w, wh);
      Spec fp = b.r * fresnel(b, costH);
      if (sameHemi(wo, w)) {
        *f = sscale(fp, fact / absCosT(w));
        *wi = w;
        *pdf = pd / (4.0f * std::fabs(costH));
      } else {
        *f = sconst(0);
        *wi = wo;
        *pdf = 0.0f;
      }
      return;
    }
    case K_FB: {  // s False (Microfacet.hs:88-99): the sampled value is e wo wi / pdf
      V3 w, wh;
      float pdfp;
      if (u1 < 0.5f) {
        w = toSameHemi(wo, cosineSampleHemisphere(u1 * 2.0f, u2));
        wh = fbHalf(w, wo);
        pdfp = anisoPdf(b.ex, b.ey, wh);
      } else {
        float d;
        wh = anisoSample(b.ex, b.ey, 2.0f * (u1 - 0.5f), u2, &d, &pdfp);
        w = sscale(2.0f * dot(wo, wh), wh) - wo;
      }
      *wi = w;
      if (pdfp == 0.0f) {
        *f = sconst(0);
        *pdf = 0.0f;
        return;
      }
      *pdf = 0.5f * (absCosT(w) * kInvPi + pdfp / (4.0f * absDot(wo, wh)));
      *f = sscale(fbEval(b, wo, w), 1.0f / *pdf);
      return;
    }
    case K_SPECR:  // Specular.hs:11-20
      *f = b.r * fresnel(b, cosT(wo));
      *wi = mk(-wo.x, -wo.y, wo.z);
      *pdf = 1.0f;
      return;
    default: {  // specTrans (Specular.hs:33-58), adj = False
      bool entering = cosT(wo) > 0.0f;
      float ei = entering ? b.ei : b.et, et = entering ? b.et : b.ei;
      float sini2 = sinT2(wo);
      float eta = ei / et;
      float eta2 = eta * eta;
      float sint2 = eta2 * sini2;
      if (sint2 >= 1.0f) {
        *f = sconst(0);
        *wi = wo;
        *pdf = 0.0f;
        return;
      }
      float c = std::sqrt(hmax(0.0f, 1.0f - sint2));
      float cost = entering ? -c : c;
      *wi = g_mutant == MUT_TRANS_DIR ? mk(eta * wo.x, eta * wo.y, cost) : mk(eta * (-wo.x), eta * (-wo.y), cost);
      Spec fr = frDielectric(ei, et, cost);
      *f = sscale((sconst(1.0f) - fr) * b.r, g_mutant == MUT_TRANS_ETA2 ? eta : eta2);
      *pdf = 1.0f;
      return;
    }
  }
}

// Selector textures on the shading DG, the material's `tex dgs`: checkerBoard
// (Texture.hs:209-219) and graphPaper (Texture.hs:191-207) over uvMapping /
// planarMapping (Texture.hs:166-179).  True where tex2 is taken.
static bool texSelectsTex2(const bgpu_texture& t, const DG& dg) {
  const float* q = t.params;
  if (t.kind == BGPU_TEX_CHECKER) {
    long long s = (long long)std::floor(dg.p.x * q[0]) + (long long)std::floor(dg.p.y * q[1]) +
                  (long long)std::floor(dg.p.z * q[2]);
    long long m = ((s % 2) + 2) % 2;  // Haskell `mod`
    return m != 0;
  }
  float x, z;
  if (t.mapping == BGPU_MAP_PLANAR) {
    x = dot(dg.p, mk(q[1], q[2], q[3])) + q[7];
    z = dot(dg.p, mk(q[4], q[5], q[6])) + q[8];
  } else {
    x = q[1] * dg.u + q[3];
    z = q[2] * dg.v + q[4];
  }
  float xf = std::fabs(x - std::trunc(x)), zf = std::fabs(z - std::trunc(z));  // properFraction
  float lo = q[0] / 2.0f, hi = 1.0f - lo;
  return xf < lo || zf < lo || xf > hi || zf > hi;
}

static void addComponents(const bgpu_material& m, const Spec& s0, const Spec& s1, Bsdf& b);

// bump d dgg dgs (Reflection.hs:347-378): the shading DG displaced along its
// normal by the scalar texture rooted at `root`, sampled at dgs and at dgs
// shifted by 0.01 in u and in v; the new normal faces the geometric one.
static DG bumpDG(const Ctx& c, int root, const DG& dgg, const DG& dgs) {
  const float du = 0.01f, dv = 0.01f;
  const tex::TexDG t0{dgs.p, dgs.u, dgs.v};
  const tex::TexDG tu{dgs.p + scale(g_mutant == MUT_BUMP_SHIFT ? dgs.dpdv : dgs.dpdu, du), dgs.u + du, dgs.v};
  const tex::TexDG tv{dgs.p + scale(dgs.dpdv, dv), dgs.u, dgs.v + dv};
  const float uDisp = tex::scalar(c.s->textures, root, tu);
  const float vDisp = tex::scalar(c.s->textures, root, tv);
  const float disp = tex::scalar(c.s->textures, root, t0);
  const float vscale = (vDisp - disp) / dv;
  const V3 dpdv = dgs.dpdv + scale(dgs.n, vscale);
  const float uscale = (uDisp - disp) / du;
  const V3 dpdu = dgs.dpdu + scale(dgs.n, uscale);
  V3 nn = normalize(cross(dpdu, dpdv));
  if (dot(nn, dgg.n) < 0.0f && g_mutant != MUT_BUMP_FACE)
    nn = mk(0.0f - nn.x, 0.0f - nn.y, 0.0f - nn.z);  // faceForward, Num's 0 - v
  DG r = dgs;
  r.n = nn;
  r.dpdu = dpdu;
  r.dpdv = dpdv;
  return r;
}

static Bsdf makeBsdf(const Ctx& c, int material, const DG& dgg, const DG& dgs_in) {
  Bsdf b;
  b.n = 0;
  const bool valid = material >= 0 && material < (int)c.s->materials.size();
  // bumpMapped d mat dgg dgs = mat dgg (bump d dgg dgs) (Reflection.hs:344-345)
  const bool bumped = valid && c.s->materials[material].tex.kind == BGPU_TEX_NODES &&
                      ((c.s->materials[material].tex.slots >> 8) & 1);
  const DG dgs = bumped ? bumpDG(c, c.s->materials[material].tex.roots[8], dgg, dgs_in) : dgs_in;
  // mkBsdf (Reflection.hs:209-218)
  V3 nn = dgs.n, sn = normalize(dgs.dpdu);
  b.cs = Frame{sn, cross(nn, sn), nn};
  b.p = dgs.p;
  b.ng = dgg.n;
  if (!valid) return b;
  const bgpu_material& m = c.s->materials[material];
  if (m.tex.kind == BGPU_TEX_NODES) {
    // textured inputs evaluated at the shading DG (`tex dgs`), the others
    // constant; the material is then built from them as from constants
    bgpu_material mt = m;
    const tex::TexDG tdg{dgs.p, dgs.u, dgs.v};
    for (int k = 0; k < 3; ++k)
      if ((m.tex.slots >> k) & 1) {
        const tex::Spec16 v = tex::spectrum(c.s->textures, m.tex.roots[k], tdg);
        for (int j = 0; j < 16; ++j) mt.spec[k][j] = v[j];
      }
    for (int k = 0; k < 4; ++k)
      if ((m.tex.slots >> (4 + k)) & 1) mt.scalar[k] = tex::scalar(c.s->textures, m.tex.roots[4 + k], tdg);
    addComponents(mt, specOf(mt.spec[0]), specOf(mt.spec[1]), b);
    return b;
  }
  const float* sp0 = m.spec[0];
  const float* sp1 = m.spec[1];
  if (m.tex.kind != BGPU_TEX_NONE && texSelectsTex2(m.tex, dgs)) {
    if (m.tex.slots & 1) sp0 = m.tex.alt[0];
    if (m.tex.slots & 2) sp1 = m.tex.alt[1];
  }
  addComponents(m, specOf(sp0), specOf(sp1), b);
  return b;
}

// The BxDFs of material m (Material.hs:32-127), with spectra s0 and s1 (the
// texture's choice for the first two slots).
static void addComponents(const bgpu_material& m, const Spec& s0, const Spec& s1, Bsdf& b) {
  auto blinnE = [](float rough) {  // mkBlinn (1 / rough), fixExponent (Microfacet.hs:127-134)
    float e = 1.0f / rough;
    return (e > 10000.0f || std::isnan(e)) ? 10000.0f : e;
  };
  switch (m.type) {
    case BGPU_MAT_MATTE: {  // Material.hs:32-41
      float sig = m.scalar[0];
      Bxdf x{};
      x.r = s0;
      if (sig == 0.0f) {
        x.kind = K_LAMBERT;
      } else {
        x.kind = K_OREN;
        float sc = clampf(sig, 0.0f, 1.0f);
        float sig2 = sc * sc;
        x.A = 1.0f - (sig2 / (2.0f * (sig2 + 0.33f)));
        x.B = 0.45f * sig2 / (sig2 + 0.09f);
      }
      x.type = BX_R | BX_D;
      b.c[b.n++] = x;
      break;
    }
    case BGPU_MAT_PLASTIC: {  // Material.hs:76-87
      Bxdf d{};
      d.kind = K_LAMBERT;
      d.type = BX_R | BX_D;
      d.r = s0;
      Bxdf g{};
      g.kind = K_MF;
      g.type = BX_R | BX_G;
      g.r = s1;
      g.e = blinnE(m.scalar[0]);
      g.fr = FR_DIEL;
      g.ei = 1.0f;
      g.et = 1.5f;
      b.c[b.n++] = d;
      b.c[b.n++] = g;
      break;
    }
    case BGPU_MAT_METAL: {  // Material.hs:89-96
      Bxdf g{};
      g.kind = K_MF;
      g.type = BX_R | BX_G;
      g.r = sconst(1.0f);
      g.e = blinnE(m.scalar[0]);
      g.fr = FR_COND;
      g.eta = s0;
      g.k = s1;
      b.c[b.n++] = g;
      break;
    }
    case BGPU_MAT_GLASS: {  // Material.hs:56-66
      float ior = m.scalar[0];
      Bxdf r{};
      r.kind = K_SPECR;
      r.type = BX_R | BX_S;
      r.r = sclamp(s0, 0.0f, 1.0f);
      r.fr = FR_DIEL;
      r.ei = 1.0f;
      r.et = ior;
      Bxdf t{};
      t.kind = K_SPECT;
      t.type = BX_T | BX_S;
      t.r = sclamp(s1, 0.0f, 1.0f);
      t.ei = 1.0f;
      t.et = ior;
      b.c[b.n++] = r;
      b.c[b.n++] = t;
      break;
    }
    case BGPU_MAT_SUBSTRATE: {  // mkSubstrate (Material.hs:111-127)
      Bxdf x{};
      x.kind = K_FB;
      x.type = BX_R | BX_G;
      x.r = sclamp(s0, 0.0f, 1.0f);
      x.rs = sclamp(s1, 0.0f, 1.0f);
      x.ra = sclamp(specOf(m.spec[2]), 0.0f, 1.0f);
      x.ex = blinnE(hmax(0.0f, m.scalar[0]));  // mkAnisotropic (1 / u) (1 / v): fixExponent of each
      x.ey = blinnE(hmax(0.0f, m.scalar[1]));
      x.depth = m.scalar[2];
      b.c[b.n++] = x;
      break;
    }
    case BGPU_MAT_TRANSMATTE: {  // translucentMatte (Material.hs:43-53)
      float sig = m.scalar[0];
      Spec r = sclamp(s0, 0.0f, 1.0f);
      Spec t = sclamp(s1, 0.0f, 1.0f);
      for (int k = 0; k < 16; ++k) t[k] = t[k] * (1.0f - r[k]);
      for (int side = 0; side < 2; ++side) {
        Bxdf x{};
        x.r = side ? t : r;
        if (sig == 0.0f) {
          x.kind = K_LAMBERT;
        } else {
          x.kind = K_OREN;
          float sc = clampf(sig, 0.0f, 1.0f);
          float sig2 = sc * sc;
          x.A = 1.0f - (sig2 / (2.0f * (sig2 + 0.33f)));
          x.B = 0.45f * sig2 / (sig2 + 0.09f);
        }
        x.type = side ? (BX_T | BX_D) : (BX_R | BX_D);
        x.flip = side;
        b.c[b.n++] = x;
      }
      break;
    }
    case BGPU_MAT_SHINYMETAL: {  // mkShinyMetal (Material.hs:98-109): [diff, spec]
      // frApproxEta / frApproxK (Fresnel.hs:72-78), band by band
      auto approxEta = [](const Spec& r) {
        Spec o;
        for (int i = 0; i < 16; ++i) {
          float c = hmax(0.0f, hmin(0.999f, r[i]));
          float rp = g_mutant == MUT_SHINY_SQRT ? c : std::sqrt(c);
          o[i] = (1.0f + rp) / (1.0f - rp);
        }
        return o;
      };
      auto approxK = [](const Spec& r) {
        Spec o;
        for (int i = 0; i < 16; ++i) {
          float refl = hmax(0.0f, hmin(0.999f, r[i]));
          o[i] = std::sqrt(refl / (1.0f - refl)) * 2.0f;
        }
        return o;
      };
      Bxdf g{};
      g.kind = K_MF;
      g.type = BX_R | BX_G;
      g.r = sconst(1.0f);
      g.e = blinnE(m.scalar[0]);
      g.fr = FR_COND;
      g.eta = approxEta(s1);
      g.k = approxK(s1);
      b.c[b.n++] = g;
      Bxdf r{};
      r.kind = K_SPECR;
      r.type = BX_R | BX_S;
      r.r = sconst(1.0f);
      r.fr = FR_COND;
      r.eta = approxEta(s0);
      r.k = approxK(s0);
      b.c[b.n++] = r;
      break;
    }
    case BGPU_MAT_MIRROR: {  // Material.hs:68-74
      Bxdf r{};
      r.kind = K_SPECR;
      r.type = BX_R | BX_S;
      r.r = sclamp(s0, 0.0f, 1.0f);
      r.fr = FR_NOOP;
      b.c[b.n++] = r;
      break;
    }
    default:
      break;
  }
}

static Spec evalBsdf(const Bsdf& b, V3 woW, V3 wiW) {  // Reflection.hs:318-332, adj = False
  float cosWo = dot(woW, b.ng);
  float sideTest = dot(wiW, b.ng) / cosWo;
  if (sideTest == 0.0f) return sconst(0);
  if (std::fabs(cosWo) < 1e-5f) return sconst(0);
  int need = sideTest < 0.0f ? BX_T : BX_R;
  V3 wo = worldToLocal(b.cs, woW), wi = worldToLocal(b.cs, wiW);
  Spec f = sconst(0);
  for (int i = 0; i < b.n; ++i)
    if (b.c[i].type & need) f = f + bxEval(b.c[i], wi, wo);
  return f;
}
static float bsdfPdf(const Bsdf& b, V3 woW, V3 wiW) {  // Reflection.hs:251-257
  if (b.n == 0) return 0.0f;
  V3 wo = worldToLocal(b.cs, woW), wi = worldToLocal(b.cs, wiW);
  float s = 0.0f;
  for (int i = 0; i < b.n; ++i) s = s + bxPdf(b.c[i], wo, wi);
  return s / (float)b.n;
}
struct BsdfSample {
  int type;
  float pdf;
  Spec f;
  V3 wi;
};
constexpr int kBxAll = BX_R | BX_T | BX_D | BX_G | BX_S;
// sampleBsdf' flags (Reflection.hs:278-316): the components whose type is a
// subset of `flags` (bxdfMatches) take part; sampleBsdf is flags = all.
static BsdfSample sampleBsdf(const Bsdf& b, V3 woW, float uComp, float u1, float u2, int flags = kBxAll) {
  BsdfSample empty{BX_R | BX_D, 0.0f, sconst(0), mk(0, 1, 0)};
  int m[8], cnt = 0;
  for (int i = 0; i < b.n; ++i)
    if ((b.c[i].type & flags) == b.c[i].type) m[cnt++] = i;
  if (cnt == 0) return empty;
  V3 wo = worldToLocal(b.cs, woW);
  float cntf = (float)cnt, invCnt = 1.0f / cntf;
  int sNum = (int)std::floor(uComp * cntf);
  sNum = cnt - 1 < sNum ? cnt - 1 : sNum;
  sNum = sNum < 0 ? 0 : sNum;
  const Bxdf& bx = b.c[m[sNum]];
  Spec fS;
  V3 wi;
  float pdfp;
  bxSample(bx, wo, u1, u2, &fS, &wi, &pdfp);
  V3 wiW = localToWorld(b.cs, wi);
  float sideTest = dot(wiW, b.ng) / dot(woW, b.ng);
  if (pdfp == 0.0f || sideTest == 0.0f) return empty;
  int need = sideTest < 0.0f ? BX_T : BX_R;
  if (!(bx.type & need)) return empty;
  if (bx.type & BX_S) return BsdfSample{bx.type, pdfp * invCnt, sscale(fS, cntf), wiW};
  if (cnt == 1) return BsdfSample{bx.type, pdfp, fS, wiW};
  float other = 0.0f;
  Spec fOthers = sconst(0);
  for (int k = 0; k < cnt; ++k) {
    if (k == sNum) continue;
    const Bxdf& o = b.c[m[k]];
    other = other + bxPdf(o, wo, wi);
    if (o.type & need) fOthers = fOthers + bxEval(o, wi, wo);
  }
  float pdf = (pdfp + other) * invCnt;
  Spec fSum = sscale(sscale(fS, pdfp) + fOthers, 1.0f / pdf);
  return BsdfSample{bx.type, pdf, fSum, wiW};
}

// --------------------------------------------------------------- integrator
struct SampleCtx {
  const PixelSampler* ps;
  uint32_t n;
};

static Spec lEmitAt(const Ctx& c, const Isect& is, V3 w) {  // intLe (Primitive.hs:68-76), lEmit (Light.hs:85-93)
  if (is.prim < c.s->ntri) return sconst(0);
  int si = is.prim - c.s->ntri;
  int li = c.shapeLight[si];
  if (li < 0) return sconst(0);
  return dot(is.dgg.n, w) > 0.0f ? c.lights[li].L : sconst(0);
}

static Spec estimateDirect(Ctx& c, int li, V3 p, float eps, V3 n, V3 wo, const Bsdf& bsdf,
                           float ulDir1, float ulDir2, float uBC, float uBD1, float uBD2) {
  const LightRec& l = c.lights[li];
  Spec ls = sconst(0);
  {  // sampleLightMis (Scene.hs:61-69)
    LightSample s = sampleLight(c, l, p, eps, n, ulDir1, ulDir2);
    Spec f = evalBsdf(bsdf, wo, s.wi);
    if (!(s.pdf == 0.0f || isBlack(s.li) || isBlack(f))) {
      c.rays++;
      if (!sceneOccluded(*c.s, s.ray, false)) {
        if (s.delta) {
          ls = sscale(f * s.li, 1.0f / s.pdf);
        } else {
          float w = powerHeuristic(s.pdf, bsdfPdf(bsdf, wo, s.wi));
          ls = sscale(f * s.li, w / s.pdf);
        }
      }
    }
  }
  Spec bs = sconst(0);
  {  // sampleBsdfMis (Scene.hs:71-82)
    BsdfSample smp = sampleBsdf(bsdf, wo, uBC, uBD1, uBD2);
    if (!(smp.pdf == 0.0f || isBlack(smp.f))) {
      Ray ray{p, smp.wi, eps, kInf};
      float lPdf = lightPdf(c, l, p, smp.wi);
      float w = powerHeuristic(smp.pdf, lPdf);
      Isect is;
      c.rays++;
      if (intersectFull(*c.s, ray, &is)) {
        int si = is.prim - c.s->ntri;
        int hitLight = (is.prim >= c.s->ntri) ? c.shapeLight[si] : -1;
        // AreaLight equality by id (Light.hs:47-50); other lights never compare equal
        if (hitLight >= 0 && l.type == 0 && c.lights[hitLight].id == l.id)
          bs = sscale(smp.f * lEmitAt(c, is, -smp.wi), w);
      } else {
        bs = sscale(smp.f * lightLe(l, ray), w);
      }
    }
  }
  return ls + bs;
}

static Spec li(Ctx& c, const Ray& camRay, const PixelSampler& ps, uint32_t n, int md, int rrDepth) {
  Spec L = sconst(0), t = sconst(1.0f);
  bool spec = true;
  int depth = 0;
  Ray ray = camRay;
  Isect is;
  c.rays++;
  bool hit = intersectFull(*c.s, ray, &is);
  for (;;) {
    if (!hit) {  // Path.hs:44-47
      if (spec) {
        Spec sum = sconst(0);
        for (const LightRec& l : c.lights) sum = sum + lightLe(l, ray);
        L = L + t * sum;
      }
      return L;
    }
    if (depth == md) return L;  // Path.hs:51
    uint32_t d = (uint32_t)depth;
    float lNumU = ps.get1(n, 1 + 4 * d);
    float lDir1, lDir2, bDir1, bDir2, ud1, ud2;
    ps.get2(n, 1 + 3 * d, &lDir1, &lDir2);
    float lBsdfCompU = ps.get1(n, 2 + 4 * d);
    ps.get2(n, 2 + 3 * d, &bDir1, &bDir2);
    Spec intl = spec ? lEmitAt(c, is, ray.d) : sconst(0);
    V3 wo = -ray.d;
    const Scene& s = *c.s;
    int mat = is.prim < s.ntri ? s.meshes[s.tri_mesh[is.prim]].material
                               : s.shapes[is.prim - s.ntri].material;
    Bsdf bsdf = makeBsdf(c, mat, is.dgg, is.dgs);
    V3 nrm = bsdf.cs.n, p = bsdf.p;
    float eps = is.eps;
    // sampleOneLight (Scene.hs:110-118)
    Spec ld = sconst(0);
    int lc = (int)c.lights.size();
    if (lc == 1) {
      ld = estimateDirect(c, 0, p, eps, nrm, wo, bsdf, lDir1, lDir2, lBsdfCompU, bDir1, bDir2);
    } else if (lc > 1) {
      int ln = (int)std::floor(lNumU * (float)lc);
      ln = ln < lc - 1 ? ln : lc - 1;
      ld = sscale(estimateDirect(c, ln, p, eps, nrm, wo, bsdf, lDir1, lDir2, lBsdfCompU, bDir1, bDir2),
                  (float)lc);
    }
    Spec lHere = intl + ld;
    L = L + t * lHere;
    float pc = depth <= rrDepth ? 1.0f : hmin(0.75f, sY(t));  // Path.hs:68
    float x = ps.get1(n, 3 + 4 * d);
    if (x > pc) return L;
    float uc = ps.get1(n, 0 + 4 * d);
    ps.get2(n, 0 + 3 * d, &ud1, &ud2);
    BsdfSample smp = sampleBsdf(bsdf, wo, uc, ud1, ud2);
    if (smp.pdf == 0.0f || isBlack(smp.f)) return L;
    ray = Ray{p, smp.wi, eps, kInf};
    spec = (smp.type & BX_S) != 0;
    if (g_mutant == MUT_PATH_PC) t = smp.f * t;
    else if (g_mutant == MUT_PATH_T) t = sscale(smp.f, 1.0f / pc);
    else t = sscale(smp.f * t, 1.0f / pc);
    depth++;
    c.rays++;
    hit = intersectFull(*c.s, ray, &is);
  }
}

// DirectLighting (Integrator/DirectLighting.hs:23-59): at every hit the
// emission towards wo, sampleOneLight with the sample dimensions 2d / 2d + 1,
// and the specular reflection and transmission continuations (uComp 0.5,
// uDir (0.5, 0.5)) up to maxDepth, each weighted by its sampled f; escaping
// rays are black.
static Spec sampleOneLightAt(Ctx& c, V3 p, float eps, V3 nrm, V3 wo, const Bsdf& bsdf, float lNumU, float lDir1,
                             float lDir2, float uBC, float uBD1, float uBD2) {  // Scene.hs:110-118
  int lc = (int)c.lights.size();
  if (lc == 1) return estimateDirect(c, 0, p, eps, nrm, wo, bsdf, lDir1, lDir2, uBC, uBD1, uBD2);
  if (lc > 1) {
    int ln = (int)std::floor(lNumU * (float)lc);
    ln = ln < lc - 1 ? ln : lc - 1;
    return sscale(estimateDirect(c, ln, p, eps, nrm, wo, bsdf, lDir1, lDir2, uBC, uBD1, uBD2), (float)lc);
  }
  return sconst(0);
}

static Spec liDirect(Ctx& c, const Ray& ray, const PixelSampler& ps, uint32_t n, int d, int md) {
  Isect is;
  c.rays++;
  if (!intersectFull(*c.s, ray, &is)) return sconst(0);
  uint32_t k = 2 * (uint32_t)d;
  float uln = ps.get1(n, k), ubc = ps.get1(n, k + 1);
  float uld1, uld2, ubd1, ubd2;
  ps.get2(n, k, &uld1, &uld2);
  ps.get2(n, k + 1, &ubd1, &ubd2);
  const Scene& s = *c.s;
  int mat = is.prim < s.ntri ? s.meshes[s.tri_mesh[is.prim]].material : s.shapes[is.prim - s.ntri].material;
  Bsdf bsdf = makeBsdf(c, mat, is.dgg, is.dgs);
  V3 p = bsdf.p, nrm = bsdf.cs.n, wo = -ray.d;
  float e = is.eps;
  Spec l = sampleOneLightAt(c, p, e, nrm, wo, bsdf, uln, uld1, uld2, ubc, ubd1, ubd2);
  auto cont = [&](int flags) {  // cont (DirectLighting.hs:45-59)
    if (d + 1 == md) return sconst(0);
    BsdfSample smp = sampleBsdf(bsdf, wo, 0.5f, 0.5f, 0.5f, flags);
    if (smp.pdf == 0.0f) return sconst(0);
    return smp.f * liDirect(c, Ray{p, smp.wi, e, kInf}, ps, n, d + 1, md);
  };
  Spec re = cont(BX_S | BX_R);
  Spec tr = cont(BX_S | BX_T);
  return l + re + tr + lEmitAt(c, is, wo);
}

// --------------------------------------------------------------------- film
static float evalFilter(const bgpu_film& f, float x, float y) {  // Filter.hs:68-96
  switch (f.filter) {
    case BGPU_FILTER_GAUSS: {
      float w = f.fparams[0], h = f.fparams[1], a = f.fparams[2];
      float ex = std::exp(-a * w * w), ey = std::exp(-a * h * h);
      float gx = hmax(0.0f, std::exp(-a * x * x) - ex), gy = hmax(0.0f, std::exp(-a * y * y) - ey);
      return gx * gy;
    }
    case BGPU_FILTER_MITCHELL: {
      float w = f.fparams[0], h = f.fparams[1], B = f.fparams[2], C = f.fparams[3];
      auto m1d = [&](float xp) {
        float xx = std::fabs(2.0f * xp);
        if (xx > 1.0f)
          return ((-B - 6 * C) * xx * xx * xx + (6 * B + 30 * C) * xx * xx + (-12 * B - 48 * C) * xx +
                  (8 * B + 24 * C)) * (1.0f / 6.0f);
        return ((12 - 9 * B - 6 * C) * xx * xx * xx + (-18 + 12 * B + 6 * C) * xx * xx + (6 - 2 * B)) *
               (1.0f / 6.0f);
      };
      return m1d(x * (1.0f / w)) * m1d(y * (1.0f / h));
    }
    case BGPU_FILTER_SINC: {
      float ix = 1.0f / f.fparams[0], iy = 1.0f / f.fparams[1], tau = f.fparams[2];
      auto s1 = [&](float v) {
        if (std::fabs(v) > 1.0f) return 0.0f;
        if (std::fabs(v) < 1e-5f) return 1.0f;
        float xp = std::fabs(v) * kPi;
        return (std::sin(xp) / xp) * (std::sin(xp * tau) / (xp * tau));
      };
      return s1(x * ix) * s1(y * iy);
    }
    case BGPU_FILTER_TRIANGLE: {
      float w = f.fparams[0], h = f.fparams[1];
      return (hmax(0.0f, w - std::fabs(x * 2.0f)) * hmax(0.0f, h - std::fabs(y * 2.0f))) / (w * h);
    }
    default:
      return (std::fabs(x) < 0.5f && std::fabs(y) < 0.5f) ? 1.0f : 0.0f;
  }
}
static void filterSize(const bgpu_film& f, float* fw, float* fh) {
  if (f.filter == BGPU_FILTER_BOX) { *fw = 0.5f; *fh = 0.5f; }
  else { *fw = f.fparams[0]; *fh = f.fparams[1]; }
}

// The tile a sample of pixel (ix, iy) is rendered into: prender splits the
// sample extent [ex0, ex1] x [ey0, ey1] into 16 x 16 windows (splitWindow,
// Sampling.hs:55-58) and gives each its own image (mkImageTile,
// Image.hs:108-120) at offset (max 0 xStart, max 0 yStart), of width
// xEnd - px + floor (0.5 + fw); addSample clips a splat to that image and
// addTile (Image.hs:178-199) drops what lies beyond the film.
struct TileClip {
  int lx, rx, ly, ry;
};
static TileClip tileClip(int ix, int iy, int ex0, int ex1, int ey0, int ey1, float fw, float fh) {
  const int xs = ex0 + ((ix - ex0) / 16) * 16, ys = ey0 + ((iy - ey0) / 16) * 16;
  const int xe = std::min(xs + 15, ex1), ye = std::min(ys + 15, ey1);
  const int px = std::max(0, xs), py = std::max(0, ys);
  const int w = xe - px + (int)std::floor(0.5f + fw), h = ye - py + (int)std::floor(0.5f + fh);
  return TileClip{px, px + w - 1, py, py + h - 1};
}

static void addSample(const Ctx& c, float* wxyz, float* sq, int W, int H, float sx, float sy, const TileClip& tc,
                      const Spec& L) {  // Image.hs:250-299 into the sample's tile
  for (float v : L)
    if (std::isnan(v) || std::isinf(v)) return;
  float X, Y, Z;
  toXYZ(L, &X, &Y, &Z);
  float fw = c.fw, fh = c.fh;
  float ifw = 1.0f / fw, ifh = 1.0f / fw;  // Image.hs:262-263 (sic)
  float dx = sx - 0.5f, dy = sy - 0.5f;
  int x0 = (int)std::ceil(dx - fw), x1 = (int)std::floor(dx + fw);
  int y0 = (int)std::ceil(dy - fh), y1 = (int)std::floor(dy + fh);
  if (g_mutant != MUT_FILTER_CLIP) {
    x0 = std::max(tc.lx, x0), x1 = std::min(tc.rx, x1);
    y0 = std::max(tc.ly, y0), y1 = std::min(tc.ry, y1);
  }
  x0 = std::max(0, x0), x1 = std::min(W - 1, x1);
  y0 = std::max(0, y0), y1 = std::min(H - 1, y1);
  if (x1 - x0 < 0 || y1 - y0 < 0) return;
  for (int y = y0; y <= y1; ++y) {
    int iy = std::min((int)std::floor(std::fabs(((float)y - dy) * ifh * 16.0f)), 15);
    for (int x = x0; x <= x1; ++x) {
      int ix = std::min((int)std::floor(std::fabs(((float)x - dx) * ifw * 16.0f)), 15);
      float w = c.filter[iy * 16 + ix];
      float* o = wxyz + 4 * ((size_t)y * W + x);
#pragma omp atomic
      o[0] += w;
#pragma omp atomic
      o[1] += X * w;
#pragma omp atomic
      o[2] += Y * w;
#pragma omp atomic
      o[3] += Z * w;
      if (sq) {
        float* q = sq + 3 * ((size_t)y * W + x);
#pragma omp atomic
        q[0] += X * X * w;
#pragma omp atomic
        q[1] += Y * Y * w;
#pragma omp atomic
        q[2] += Z * Z * w;
      }
    }
  }
}

static Ray fireRay(const bgpu_camera& cam, float ix, float iy, float lu, float lv) {  // Camera.hs:49-76
  M4 r2c, c2w;
  std::memcpy(r2c.m, cam.r2c, sizeof(r2c.m));
  std::memcpy(c2w.m, cam.c2w, sizeof(c2w.m));
  if (cam.type == BGPU_CAMERA_ENVIRONMENT) {  // fireRay (Environment c2w sx sy)
    const float t = kPi * iy / cam.yres, p = 2.0f * kPi * ix / cam.xres;
    const V3 d = mk(std::sin(t) * std::cos(p), std::cos(t), std::sin(t) * std::sin(p));
    return transRay(c2w, Ray{mk(0, 0, 0), d, 0.0f, kInf});
  }
  V3 pCamera = transPoint(r2c, mk(ix, iy, 0.0f));
  Ray ray{mk(0, 0, 0), normalize(pCamera), 0.0f, kInf};
  if (cam.lens_radius > 0.0f) {
    float a, b;
    concentricSampleDisk(lu, lv, &a, &b);
    V3 ro = mk(a * cam.lens_radius, b * cam.lens_radius, 0.0f);
    V3 pFocus = rayAt(ray, cam.focal_distance / ray.d.z);
    ray = Ray{ro, normalize(pFocus - ro), 0.0f, kInf};
  }
  return transRay(c2w, ray);
}

}  // namespace orc

extern "C" {

// The sampler's primitives, for known-answer and bijection tests.
uint64_t orc_mix64(uint64_t x) { return orc::mix64(x); }
uint32_t orc_permute(uint32_t i, uint32_t n, uint32_t p) { return orc::permute(i, n, p); }

// Whether texture `t` picks tex2 at point p with surface parameters (u, v).
int orc_tex_selects(const bgpu_texture* t, float px, float py, float pz, float u, float v) {
  orc::DG dg{};
  dg.p = orc::mk(px, py, pz);
  dg.u = u;
  dg.v = v;
  return orc::texSelectsTex2(*t, dg) ? 1 : 0;
}

// The anisotropic distribution (ex, ey) at n half vectors wh (3 n floats):
// D into d[i], its pdf into pdf[i]; with u (2 n floats) set, the sampler's
// half vectors go to wh and their pdfs to pdf instead.
void orc_aniso(float ex, float ey, int n, const float* u, float* wh, float* d, float* pdf) {
  using namespace orc;
  for (int i = 0; i < n; ++i) {
    if (u) {
      float dd;
      V3 w = anisoSample(ex, ey, u[2 * i], u[2 * i + 1], &dd, &pdf[i]);
      wh[3 * i] = w.x, wh[3 * i + 1] = w.y, wh[3 * i + 2] = w.z;
      d[i] = dd;
    } else {
      V3 w = mk(wh[3 * i], wh[3 * i + 1], wh[3 * i + 2]);
      d[i] = anisoD(ex, ey, w);
      pdf[i] = anisoPdf(ex, ey, w);
    }
  }
}

// Splats n samples at image positions (sx[i], sy[i]) with XYZ-giving
// constant spectra: sample i has the spectrum sConst v[i] and belongs to
// pixel (px[i], py[i]) of the film's sample extent (prender's tiles).
void orc_film_splat(const bgpu_film* film, int n, const float* sx, const float* sy, const int* px, const int* py,
                    const float* v, float* wxyz) {
  using namespace orc;
  Ctx c;
  c.s = nullptr;
  filterSize(*film, &c.fw, &c.fh);
  for (int y = 0; y < 16; ++y) {  // mkTableFilter (Image.hs:48-61)
    float fy = ((float)y + 0.5f) * c.fh / 16.0f;
    for (int x = 0; x < 16; ++x) {
      float fx = ((float)x + 0.5f) * c.fw / 16.0f;
      c.filter[y * 16 + x] = evalFilter(*film, fx, fy);
    }
  }
  const int W = film->width, H = film->height;
  int x0 = (int)std::floor(0.5f - c.fw), x1 = (int)std::floor(0.5f + (float)W + c.fw);
  int y0 = (int)std::floor(0.5f - c.fh), y1 = (int)std::floor(0.5f + (float)H + c.fh);
  for (int i = 0; i < n; ++i)
    addSample(c, wxyz, nullptr, W, H, sx[i], sy[i], tileClip(px[i], py[i], x0, x1, y0, y1, c.fw, c.fh),
              sconst(v[i]));
}

// Selects a planted mistake (the MUT_* list); 0 restores the restatement.
void orc_set_mutant(int m) {
  orc::g_mutant = m;
  orc::tex::texMutant() = m;
}

// Texture `node` of the scene's node list at n points (xyz, uv): 16 bands
// per point for a spectrum texture, one value for a scalar one.  Returns
// 0, or -1 for a node that is not a texture.
int orc_tex_eval(void* scene, int node, int n, const float* xyz, const float* uv, float* out) {
  const orc::Scene* s = static_cast<const orc::Scene*>(scene);
  if (node < 0 || node >= (int)s->textures.nodes.size()) return -1;
  const int op = s->textures.nodes[node].op;
  const bool spectrum = orc::tex::isSpectrumOp(op);
  if (!spectrum && !orc::tex::isScalarOp(op)) return -1;
  for (int k = 0; k < n; ++k) {
    const orc::tex::TexDG dg{orc::mk(xyz[3 * k], xyz[3 * k + 1], xyz[3 * k + 2]), uv[2 * k], uv[2 * k + 1]};
    if (spectrum) {
      const orc::tex::Spec16 v = orc::tex::spectrum(s->textures, node, dg);
      for (int b = 0; b < 16; ++b) out[16 * k + b] = v[b];
    } else {
      out[k] = orc::tex::scalar(s->textures, node, dg);
    }
  }
  return 0;
}

// The bumped shading DG of material `material` (bumpMapped, its bump texture
// at roots[8]) for a shading DG g = p, n, dpdu, dpdv, u, v (14 floats) and a
// geometric normal ng: out = n, dpdu, dpdv (9 floats).  Returns 0, or -1 when
// the material is not bump-mapped.
int orc_bump(void* scene, int material, const float* g, const float* ng, float* out) {
  using namespace orc;
  const Scene* s = static_cast<const Scene*>(scene);
  if (material < 0 || material >= (int)s->materials.size()) return -1;
  const bgpu_material& m = s->materials[material];
  if (m.tex.kind != BGPU_TEX_NODES || !((m.tex.slots >> 8) & 1)) return -1;
  Ctx c;
  c.s = s;
  DG dgs{};
  dgs.p = mk(g[0], g[1], g[2]);
  dgs.n = mk(g[3], g[4], g[5]);
  dgs.dpdu = mk(g[6], g[7], g[8]);
  dgs.dpdv = mk(g[9], g[10], g[11]);
  dgs.u = g[12];
  dgs.v = g[13];
  DG dgg = dgs;
  dgg.n = mk(ng[0], ng[1], ng[2]);
  const DG r = bumpDG(c, m.tex.roots[8], dgg, dgs);
  const V3 v[3] = {r.n, r.dpdu, r.dpdv};
  for (int k = 0; k < 3; ++k) {
    out[3 * k] = v[k].x;
    out[3 * k + 1] = v[k].y;
    out[3 * k + 2] = v[k].z;
  }
  return 0;
}

// frDielectric etai etat cosi (kind 0) or frConductor eta k cosi (kind 1) as
// 16 bands (Fresnel.hs:21-70).
void orc_fresnel(int kind, float etai, float etat, const float* eta, const float* k, float cosi, float* out) {
  using namespace orc;
  Spec r = kind == 0 ? frDielectric(etai, etat, cosi) : frConductor(specOf(eta), specOf(k), cosi);
  for (int b = 0; b < 16; ++b) out[b] = r[b];
}

// The Blinn distribution of exponent e at n half vectors wh (3 n floats):
// D and the half-vector sampling pdf (Microfacet.hs:175-196 via mkBlinn).
void orc_blinn(float e, int n, const float* wh, float* d, float* pdf) {
  using namespace orc;
  for (int i = 0; i < n; ++i) {
    V3 w = mk(wh[3 * i], wh[3 * i + 1], wh[3 * i + 2]);
    d[i] = blinnD(e, w);
    pdf[i] = blinnPdf(e, w);
  }
}

// An infinite area light with a w x h spectral map (mkInfiniteAreaLight,
// Light.hs:72-82,130-141,215-229): op 0 samples n directions with (u1, u2) =
// u[2 i..] into wi (3 n floats), their pdf and radiance li (16 n); op 1
// evaluates the pdf and le at the given wi.  w2l and l2w are row-major 4x4.
static void envlightOps(orc::LightRec& l, const float* w2l, const float* l2w, int op, int n, const float* u,
                        float* wi, float* pdf, float* li) {
  using namespace orc;
  std::memcpy(l.w2l.m, w2l, sizeof(l.w2l.m));
  std::memcpy(l.l2w.m, l2w, sizeof(l.l2w.m));
  l.dist = buildDist(l);
  Ctx c;
  c.s = nullptr;
  for (int i = 0; i < n; ++i) {
    Spec le;
    if (op == 0) {
      LightSample ls = sampleLight(c, l, mk(0, 0, 0), 0.0f, mk(0, 0, 1), u[2 * i], u[2 * i + 1]);
      wi[3 * i] = ls.wi.x, wi[3 * i + 1] = ls.wi.y, wi[3 * i + 2] = ls.wi.z;
      pdf[i] = ls.pdf;
      le = ls.li;
    } else {
      V3 d = mk(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]);
      pdf[i] = lightPdf(c, l, mk(0, 0, 0), d);
      le = lightLe(l, Ray{mk(0, 0, 0), d, 0.0f, kInf});
    }
    for (int b = 0; b < 16; ++b) li[16 * i + b] = le[b];
  }
}

void orc_envlight(int w, int h, const float* map, const float* w2l, const float* l2w, int op, int n,
                  const float* u, float* wi, float* pdf, float* li) {
  using namespace orc;
  LightRec l{};
  l.type = 1;
  l.shape = -1;
  l.id = -1;
  l.w = w;
  l.h = h;
  l.map = map;
  envlightOps(l, w2l, l2w, op, n, u, wi, pdf, li);
}

// The same for an infinite light with the sun/sky map mkSunSkyLight east
// sunDir turbidity, sky = [east xyz, sunDir xyz, turbidity].
void orc_skylight(const float* sky, const float* w2l, const float* l2w, int op, int n, const float* u, float* wi,
                  float* pdf, float* li) {
  using namespace orc;
  LightRec l{};
  l.type = 1;
  l.shape = -1;
  l.id = -1;
  l.sky = true;
  l.sd = initSky(mk(sky[0], sky[1], sky[2]), mk(sky[3], sky[4], sky[5]), sky[6]);
  l.w = 640;
  l.h = 480;
  envlightOps(l, w2l, l2w, op, n, u, wi, pdf, li);
}

// The sun/sky map itself: its spectrum at n points uv (2 n floats) into
// out (16 n), and the model's data: sunR (16), sunDir (3), sunTheta, the
// zenith x, y and Y, into data (23 floats).
void orc_sky_eval(const float* sky, int n, const float* uv, float* out, float* data) {
  using namespace orc;
  SkyData k = initSky(mk(sky[0], sky[1], sky[2]), mk(sky[3], sky[4], sky[5]), sky[6]);
  for (int i = 0; i < n; ++i) {
    Spec s = skyEval(k, uv[2 * i], uv[2 * i + 1]);
    for (int b = 0; b < 16; ++b) out[16 * i + b] = s[b];
  }
  for (int b = 0; b < 16; ++b) data[b] = k.sunR[b];
  data[16] = k.sunDir.x, data[17] = k.sunDir.y, data[18] = k.sunDir.z;
  data[19] = k.sunTheta, data[20] = k.zx, data[21] = k.zy, data[22] = k.zY;
}

// Component `comp` of material m in its local shading frame, for n
// directions wo (3 n floats): op 0 evaluates bxdfEval wo wi into f (16 n) and
// bxdfPdf wo wi into pdf; op 1 samples with (u1, u2) = u[2 i..], writing wi,
// f and pdf.  Returns -1 when the material has no such component.
int orc_bxdf_comp(const bgpu_material* m, int comp, int op, int n, const float* wo, const float* u, float* wi,
                  float* f, float* pdf) {
  using namespace orc;
  Bsdf b;
  b.n = 0;
  addComponents(*m, specOf(m->spec[0]), specOf(m->spec[1]), b);
  if (comp < 0 || comp >= b.n) return -1;
  const Bxdf& x = b.c[comp];
  for (int i = 0; i < n; ++i) {
    V3 o = mk(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]);
    Spec s;
    if (op == 0) {
      V3 w = mk(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]);
      s = bxEval(x, o, w);
      pdf[i] = bxPdf(x, o, w);
    } else {
      V3 w;
      bxSample(x, o, u[2 * i], u[2 * i + 1], &s, &w, &pdf[i]);
      wi[3 * i] = w.x, wi[3 * i + 1] = w.y, wi[3 * i + 2] = w.z;
    }
    for (int k = 0; k < 16; ++k) f[16 * i + k] = s[k];
  }
  return 0;
}

// Component 0 (orc_bxdf_comp with comp 0).
int orc_bxdf(const bgpu_material* m, int op, int n, const float* wo, const float* u, float* wi, float* f,
             float* pdf) {
  return orc_bxdf_comp(m, 0, op, n, wo, u, wi, f, pdf);
}

// Sampler values of every sample n of pixel (pix, pix) in `pass`, for the
// stratification tests: per row [camera x, camera y, 1D table dim 0,
// 2D table dim 0 (u, v), 1D dim 100 (past the tables), lens (u, v)].
void orc_sampler_table(uint64_t seed, int pass, int pix, int nu, int nv, float* out) {
  using namespace orc;
  uint32_t N = (uint32_t)(nu * nv);
  PixelSampler ps{pixelKey(seed, (uint32_t)pass, pix, pix), (uint32_t)nu, (uint32_t)nv, N, 4u, 3u};
  for (uint32_t n = 0; n < N; ++n) {
    float* o = out + 8 * n;
    ps.camera(n, &o[0], &o[1]);
    o[2] = ps.get1(n, 0);
    ps.get2(n, 0, &o[3], &o[4]);
    o[5] = ps.get1(n, 100);
    ps.lens(n, &o[6], &o[7]);
  }
}

// Renders a whole job (all passes) on the CPU.  Returns the number of ray
// queries traced via *rays.  wxyz: w*h*4, sq: w*h*3 or NULL.
int orc_render(void* sp, const bgpu_job_desc* job, float* wxyz, float* sq, int nthreads,
               uint64_t* rays, int pass_begin, int pass_end) {
  using namespace orc;
  const Scene& s = *(const Scene*)sp;
  Ctx c;
  c.s = &s;
  // light list: explicit lights then geometric lights in primitive order (Scene.hs:37-43)
  for (size_t i = 0; i < s.lights.size(); ++i) {
    LightRec l{};
    l.type = 1;
    l.shape = -1;
    l.id = -1;
    if (s.lights[i].type == BGPU_LIGHT_POINT || s.lights[i].type == BGPU_LIGHT_DIRECTIONAL) {  // Light.hs:53-61
      l.type = s.lights[i].type == BGPU_LIGHT_POINT ? 2 : 3;
      l.L = specOf(s.light_maps[i].data());
      V3 v = mk(s.lights[i].vec[0], s.lights[i].vec[1], s.lights[i].vec[2]);
      l.vec = l.type == 3 ? normalize(v) : v;  // mkDirectional s n = Directional s (normalize n)
      c.lights.push_back(l);
      continue;
    }
    l.sky = s.lights[i].map == BGPU_LIGHTMAP_SUNSKY;
    if (l.sky) {  // mkSunSkyLight east sunDir turbidity: 640 x 480 (SunSky.hs:17)
      const float* k = s.lights[i].sky;
      l.sd = initSky(mk(k[0], k[1], k[2]), mk(k[3], k[4], k[5]), k[6]);
      l.w = 640;
      l.h = 480;
      l.map = nullptr;
    } else {
      l.w = s.lights[i].width;
      l.h = s.lights[i].height;
      l.map = s.light_maps[i].data();
    }
    std::memcpy(l.w2l.m, s.lights[i].w2l, sizeof(l.w2l.m));
    std::memcpy(l.l2w.m, s.lights[i].l2w, sizeof(l.l2w.m));
    c.lights.push_back(l);
  }
  for (auto& l : c.lights)
    if (l.type == 1) l.dist = buildDist(l);
  c.shapeLight.assign(s.shapes.size(), -1);
  for (size_t i = 0; i < s.shapes.size(); ++i) {
    if (!s.shapes[i].emissive) continue;
    LightRec l{};
    l.type = 0;
    l.shape = (int)i;
    l.id = s.shapes[i].light_id;
    l.L = specOf(s.shapes[i].emission);
    c.shapeLight[i] = (int)c.lights.size();
    c.lights.push_back(l);
  }
  // a job's camera and film follow its imageSize (IO/CameraParser.hs:30-38, Rendering.hs:43-44)
  const bgpu_film& film = job->film ? *job->film : s.film;
  const bgpu_camera& camera = job->camera ? *job->camera : s.camera;
  filterSize(film, &c.fw, &c.fh);
  for (int y = 0; y < 16; ++y) {  // mkTableFilter (Image.hs:48-61)
    float fy = ((float)y + 0.5f) * c.fh / 16.0f;
    for (int x = 0; x < 16; ++x) {
      float fx = ((float)x + 0.5f) * c.fw / 16.0f;
      c.filter[y * 16 + x] = evalFilter(film, fx, fy);
    }
  }
  const int W = film.width, H = film.height;
  // sampleExtent (Image.hs:162-168)
  int x0 = (int)std::floor(0.5f - c.fw), x1 = (int)std::floor(0.5f + (float)W + c.fw);
  int y0 = (int)std::floor(0.5f - c.fh), y1 = (int)std::floor(0.5f + (float)H + c.fh);
  int ew = x1 - x0 + 1, eh = y1 - y0 + 1;
  uint32_t nu = job->spp_nu, nv = job->spp_nv, N = nu * nv;
  int md = job->max_depth, sd = job->sample_depth, rr = job->rr_start_depth;
  const bool direct = job->integrator == BGPU_INTEGRATOR_DIRECT;
  // the sampler's table sizes: SurfaceIntegrator s1d s2d (Path.hs:18-30,
  // DirectLighting.hs:14-20)
  const uint32_t n1d = direct ? 2u * (uint32_t)md : 4u * (uint32_t)sd;
  const uint32_t n2d = direct ? 2u * (uint32_t)md : 3u * (uint32_t)sd;
  for (int pass = pass_begin; pass < pass_end; ++pass) {
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
    for (long pix = 0; pix < (long)ew * eh; ++pix) {
      int ix = x0 + (int)(pix % ew), iy = y0 + (int)(pix / ew);
      PixelSampler ps{pixelKey(job->seed, (uint32_t)pass, ix, iy), nu, nv, N, n1d, n2d};
      for (uint32_t n = 0; n < N; ++n) {
        float ox, oy, lu, lv;
        ps.camera(n, &ox, &oy);
        ps.lens(n, &lu, &lv);
        float sx = (float)ix + ox, sy = (float)iy + oy;
        Ray r = fireRay(camera, sx, sy, lu, lv);
        Spec L = direct ? liDirect(c, r, ps, n, 0, md) : li(c, r, ps, n, md, rr);
        addSample(c, wxyz, sq, W, H, sx, sy, tileClip(ix, iy, x0, x1, y0, y1, c.fw, c.fh), L);
      }
    }
  }
  if (rays) *rays = c.rays.load();
  return 0;
}

}  // extern "C"
