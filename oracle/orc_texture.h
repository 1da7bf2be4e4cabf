// Textures of a scene's node list (bgpu_tex_node), restating Texture.hs:59-414
// as the material parser builds them (IO/MaterialParser.hs:106-226).  Test
// infrastructure: only the tests, smoke() and the bench's CPU baseline load
// the oracle.
//
// A node is evaluated recursively from its root, as the Haskell closures
// are, at the shading DG the material sees (`tex dgs`).
#pragma once
#include <array>
#include <cmath>
#include <cstdint>
#include <vector>

#include "../bling_amd/csrc/noise_table.h"
#include "../include/bgpu.h"
#include "orc_math.h"

namespace orc {
namespace tex {

using Spec16 = std::array<float, 16>;

// The point and surface parameters a texture reads (dgP, dgU, dgV).
struct TexDG {
  V3 p;
  float u, v;
};

struct Textures {
  std::vector<bgpu_tex_node> nodes;
  struct Image {
    int w = 0, h = 0, ch = 0;
    std::vector<float> texels;
  };
  std::vector<Image> images;
};

// The noise permutation, noisePerms (Texture.hs:400-414), twice.
constexpr const int* kNoisePerms = bgpu::kNoisePermTable;

// A planted mistake in the texture functions (the oracle's mutant switch,
// test-only): 0 none; kMutPerlinWeight noiseWeight taken as t; kMutCellOwn
// only the point's own cell searched; kMutGradSwap the gradient's weights
// swapped; kMutCrystalWrap no mirroring of odd integer parts; kMutBlendSwap
// the blend's weights swapped; kMutImageFlip the image's rows not flipped;
// kMutFbmFreq octave frequencies 2^k instead of 1.99^k; kMutCheckerTrunc the
// checker's cells truncated toward zero instead of floored.
enum { kMutPerlinWeight = 14, kMutCellOwn = 15, kMutGradSwap = 16, kMutCrystalWrap = 17, kMutBlendSwap = 18,
       kMutImageFlip = 19, kMutFbmFreq = 20, kMutCheckerTrunc = 21 };
inline int& texMutant() {
  static int m = 0;
  return m;
}

// floor to an integer; 0 for values no 64-bit integer holds (NaN included:
// the reference's `floor` of such a Float is unspecified)
inline long long floorLL(float x) { return std::fabs(x) < 4.6e18f ? (long long)std::floor(x) : 0ll; }

// lerp (Math.hs:108-110)
inline float lerpT(float t, float a, float b) { return (1.0f - t) * a + t * b; }

// noiseWeight (Texture.hs:388-391): 6 t^4 t - 15 t^4 + 10 t^3
inline float noiseWeight(float t) {
  if (texMutant() == kMutPerlinWeight) return t;
  float t3 = t * t * t;
  float t4 = t3 * t;
  return 6.0f * t4 * t - 15.0f * t4 + 10.0f * t3;
}

// grad (Texture.hs:393-400)
inline float noiseGrad(int x, int y, int z, float dx, float dy, float dz) {
  int h = kNoisePerms[kNoisePerms[kNoisePerms[x] + y] + z] & 15;
  float uu = (h < 8 || h == 12 || h == 13) ? dx : dy;
  float vv = (h < 4 || h == 12 || h == 13) ? dy : dz;
  float u = (h & 1) != 0 ? -uu : uu;
  float v = (h & 2) != 0 ? -vv : vv;
  return u + v;
}

// perlin3d (Texture.hs:358-386)
inline float perlin3d(float x, float y, float z) {
  long long fx = floorLL(x), fy = floorLL(y), fz = floorLL(z);
  float dx = x - (float)fx, dy = y - (float)fy, dz = z - (float)fz;
  int ix = (int)(fx & 255), iy = (int)(fy & 255), iz = (int)(fz & 255);
  float w000 = noiseGrad(ix, iy, iz, dx, dy, dz);
  float w100 = noiseGrad(ix + 1, iy, iz, dx - 1.0f, dy, dz);
  float w010 = noiseGrad(ix, iy + 1, iz, dx, dy - 1.0f, dz);
  float w110 = noiseGrad(ix + 1, iy + 1, iz, dx - 1.0f, dy - 1.0f, dz);
  float w001 = noiseGrad(ix, iy, iz + 1, dx, dy, dz - 1.0f);
  float w101 = noiseGrad(ix + 1, iy, iz + 1, dx - 1.0f, dy, dz - 1.0f);
  float w011 = noiseGrad(ix, iy + 1, iz + 1, dx, dy - 1.0f, dz - 1.0f);
  float w111 = noiseGrad(ix + 1, iy + 1, iz + 1, dx - 1.0f, dy - 1.0f, dz - 1.0f);
  float wx = noiseWeight(dx), wy = noiseWeight(dy), wz = noiseWeight(dz);
  float x00 = lerpT(wx, w000, w100), x10 = lerpT(wx, w010, w110);
  float x01 = lerpT(wx, w001, w101), x11 = lerpT(wx, w011, w111);
  float y0 = lerpT(wy, x00, x10), y1 = lerpT(wy, x01, x11);
  return lerpT(wz, y0, y1);
}

// fbm (Texture.hs:342-350): the sum, left to right, of omega^k perlin3d at
// the point scaled by 1.99^k, for the first `octaves` k.
inline float fbm(int octaves, float omega, V3 p) {
  float s = 0.0f, l = 1.0f, o = 1.0f;
  for (int k = 0; k < octaves; ++k) {
    s = s + o * perlin3d(p.x * l, p.y * l, p.z * l);
    l = (texMutant() == kMutFbmFreq ? 2.0f : 1.99f) * l;
    o = omega * o;
  }
  return s;
}

// Worley's cell noise (Texture.hs:253-307)
inline long long lcg(long long x) { return (1103515245LL * x + 12345LL) % 4294967296LL; }
inline long long cellHash(long long x, long long y, long long z) {
  long long h = (x * 73856093LL) ^ (y * 19349663LL) ^ (z * 83492791LL);
  return (h < 0 ? -h : h) % 4294967296LL;
}
inline int cellCount(long long v) {  // prob: a Poisson look-up
  if (v < 393325350LL) return 1;
  if (v < 1022645910LL) return 2;
  if (v < 1861739990LL) return 3;
  if (v < 2700834071LL) return 4;
  if (v < 3372109335LL) return 5;
  if (v < 3819626178LL) return 6;
  if (v < 4075350088LL) return 7;
  if (v < 4203212043LL) return 8;
  return 9;
}
inline float cellDist(int kind, V3 a, V3 b) {  // euclidianDist .. chebyshevDist
  V3 d = mk(a.x - b.x, a.y - b.y, a.z - b.z);
  switch (kind) {
    case 1: return d.x * d.x + d.y * d.y + d.z * d.z;
    case 2: return std::fabs(d.x) + std::fabs(d.y) + std::fabs(d.z);
    case 3: {
      float m = std::fabs(d.x), ay = std::fabs(d.y), az = std::fabs(d.z);
      m = m <= ay ? ay : m;  // maximum, Haskell max
      return m <= az ? az : m;
    }
    default: return std::sqrt(d.x * d.x + d.y * d.y + d.z * d.z);
  }
}
inline float cellNoise(int kind, V3 p) {
  long long ox = floorLL(p.x), oy = floorLL(p.y), oz = floorLL(p.z);
  bool first = true;
  float best = 0.0f;
  const long long r = texMutant() == kMutCellOwn ? 0 : 1;
  for (long long x = ox - r; x <= ox + r; ++x)
    for (long long y = oy - r; y <= oy + r; ++y)
      for (long long z = oz - r; z <= oz + r; ++z) {
        long long u = lcg(cellHash(x, y, z));
        int n = cellCount(u);
        for (int k = 0; k < n; ++k) {
          long long u1 = lcg(u), u2 = lcg(u1), u3 = lcg(u2);
          V3 q = mk((float)x + (float)u1 / 4294967296.0f, (float)y + (float)u2 / 4294967296.0f,
                    (float)z + (float)u3 / 4294967296.0f);
          u = u3;
          float d = cellDist(kind, p, q);
          if (first || !(best <= d)) best = d;  // minimum, Haskell min
          first = false;
        }
      }
  return best;
}

// quasiCrystal (Texture.hs:309-340): the angles 0, pi/n, ... as GHC 8.0's
// enumFromThen makes them (each the previous one doubled, less the one
// before), the waves summed left to right, then wrapped.
inline float quasiCrystal(int octaves, float x, float y) {
  const float step = 3.14159265358979323846f / (float)octaves;
  float s = 0.0f, t1 = 0.0f, t2 = 0.0f;  // the two angles before this one
  for (int k = 0; k < octaves; ++k) {
    const float th = k == 0 ? 0.0f : (k == 1 ? step : (t1 + t1) - t2);
    t2 = t1;
    t1 = th;
    const float w = (std::cos(std::cos(th) * x + std::sin(th) * y) + 1.0f) / 2.0f;
    s = s + w;
  }
  // wrap: properFraction, then odd integer parts mirror
  float tr = std::trunc(s);
  long long k = floorLL(tr);
  float v = s - tr;
  if (v < 0.0f) {
    k = k - 1;
    v = 1.0f + v;
  }
  return ((k & 1) != 0 && texMutant() != kMutCrystalWrap) ? 1.0f - v : v;
}

// The 2d mappings (Texture.hs:162-179) and the 3d identity mapping
// (Texture.hs:150-156).
inline void map2d(const bgpu_tex_node& n, const TexDG& dg, float* u, float* v) {
  const float* m = n.m;
  if (n.map == BGPU_MAP_PLANAR) {
    *u = (dg.p.x * m[0] + dg.p.y * m[1] + dg.p.z * m[2]) + m[6];
    *v = (dg.p.x * m[3] + dg.p.y * m[4] + dg.p.z * m[5]) + m[7];
  } else {
    *u = m[0] * dg.u + m[2];
    *v = m[1] * dg.v + m[3];
  }
}
inline V3 map3d(const bgpu_tex_node& n, const TexDG& dg) {
  M4 w2t;
  for (int k = 0; k < 16; ++k) w2t.m[k] = n.m[k];
  return transPoint(w2t, dg.p);
}

// getPixel / getPixelScalar (Texture.hs:89-104): the texel index at (u, v).
inline long long imageTexel(const Textures::Image& im, float u, float v) {
  auto modp = [](long long a, long long b) { long long r = a % b; return r < 0 ? r + b : r; };
  long long px = modp(floorLL(u * (float)im.w), im.w);
  long long py = modp(floorLL((texMutant() == kMutImageFlip ? v : -v) * (float)im.h), im.h);
  return py * im.w + px;
}

// Whether each node's children precede it and have the kind its operation
// takes (scalar or spectrum); nullptr if so, else the reason.
inline bool isScalarOp(int op) { return op >= BGPU_TN_SCONST && op <= BGPU_TN_SCALE; }
inline bool isSpectrumOp(int op) { return op >= BGPU_TN_CONST && op <= BGPU_TN_GRAPHPAPER; }
inline const char* validate(const Textures& T) {
  const int n = (int)T.nodes.size();
  for (int i = 0; i < n; ++i) {
    const bgpu_tex_node& t = T.nodes[i];
    auto sc = [&](int c) { return c >= 0 && c < i && isScalarOp(T.nodes[c].op); };
    auto sp = [&](int c) { return c >= 0 && c < i && isSpectrumOp(T.nodes[c].op); };
    auto map2 = [&]() { return t.map == BGPU_MAP_UV || t.map == BGPU_MAP_PLANAR; };
    auto img = [&](int ch) { return t.a >= 0 && t.a < (int)T.images.size() && T.images[t.a].ch == ch; };
    bool ok = true;
    switch (t.op) {
      case BGPU_TN_SCONST: case BGPU_TN_CONST: case BGPU_TN_GSTEP: break;
      case BGPU_TN_SIMAGE: ok = img(1) && map2(); break;
      case BGPU_TN_IMAGE: ok = img(16) && map2(); break;
      case BGPU_TN_CELL: ok = t.a >= 0 && t.a <= 3 && t.map == BGPU_MAP_IDENTITY; break;
      case BGPU_TN_FBM: ok = t.a >= 0 && t.map == BGPU_MAP_IDENTITY; break;
      case BGPU_TN_PERLIN: ok = t.map == BGPU_MAP_IDENTITY; break;
      case BGPU_TN_CRYSTAL: ok = t.a >= 1 && map2(); break;
      case BGPU_TN_SCALE: ok = sc(t.a); break;
      case BGPU_TN_BLEND: ok = sp(t.a) && sp(t.b) && sc(t.c); break;
      case BGPU_TN_CHECKER: ok = sp(t.a) && sp(t.b); break;
      case BGPU_TN_GRAPHPAPER: ok = sp(t.a) && sp(t.b) && map2(); break;
      case BGPU_TN_GRADIENT:
        ok = sc(t.a) && t.c >= 1 && t.b >= 0 && t.b + t.c <= i;
        for (int k = 0; ok && k < t.c; ++k) {
          ok = T.nodes[t.b + k].op == BGPU_TN_GSTEP;
          if (ok && k > 0) ok = T.nodes[t.b + k - 1].q[0] <= T.nodes[t.b + k].q[0];
        }
        break;
      default: ok = false;
    }
    if (!ok) return "bad texture node";
  }
  return nullptr;
}

float scalar(const Textures& T, int node, const TexDG& dg);

inline Spec16 spectrum(const Textures& T, int node, const TexDG& dg) {
  const bgpu_tex_node& n = T.nodes[node];
  Spec16 out{};
  switch (n.op) {
    case BGPU_TN_CONST:
      for (int b = 0; b < 16; ++b) out[b] = n.spec[b];
      return out;
    case BGPU_TN_IMAGE: {
      const Textures::Image& im = T.images[n.a];
      float u, v;
      map2d(n, dg, &u, &v);
      const float* t = im.texels.data() + 16 * imageTexel(im, u, v);
      for (int b = 0; b < 16; ++b) out[b] = t[b];
      return out;
    }
    case BGPU_TN_BLEND: {  // spectrumBlend (Texture.hs:133-145)
      Spec16 v1 = spectrum(T, n.a, dg), v2 = spectrum(T, n.b, dg);
      float x = scalar(T, n.c, dg);
      if (x <= 0.0f) return v1;
      if (x >= 1.0f) return v2;
      const float x1 = texMutant() == kMutBlendSwap ? 1.0f - x : x;
      for (int b = 0; b < 16; ++b) out[b] = v1[b] * (1.0f - x1) + v2[b] * x1;
      return out;
    }
    case BGPU_TN_CHECKER: {  // checkerBoard (Texture.hs:209-219)
      long long s = texMutant() == kMutCheckerTrunc
                            ? (long long)std::trunc(dg.p.x * n.q[0]) + (long long)std::trunc(dg.p.y * n.q[1]) +
                                  (long long)std::trunc(dg.p.z * n.q[2])
                            : floorLL(dg.p.x * n.q[0]) + floorLL(dg.p.y * n.q[1]) + floorLL(dg.p.z * n.q[2]);
      return ((s % 2) + 2) % 2 == 0 ? spectrum(T, n.a, dg) : spectrum(T, n.b, dg);
    }
    case BGPU_TN_GRADIENT: {  // gradient (Texture.hs:236-247)
      float f = scalar(T, n.a, dg);
      const int first = n.b, cnt = n.c;
      float lo = T.nodes[first].q[0], hi = lo;
      for (int k = 1; k < cnt; ++k) {
        float p = T.nodes[first + k].q[0];
        lo = p < lo ? p : lo;
        hi = p > hi ? p : hi;
      }
      auto col = [&](int k) {
        Spec16 c;
        for (int b = 0; b < 16; ++b) c[b] = T.nodes[first + k].spec[b];
        return c;
      };
      if (f <= lo) return col(0);
      if (f >= hi) return col(cnt - 1);
      int idx = -1;
      for (int k = 0; k < cnt; ++k)
        if (T.nodes[first + k].q[0] > f) {
          idx = k;
          break;
        }
      if (idx <= 0) return out;  // f is NaN: the reference's fromJust fails; black here
      float p0 = T.nodes[first + idx - 1].q[0], p1 = T.nodes[first + idx].q[0];
      float w = (f - p0) / (p1 - p0);
      if (texMutant() == kMutGradSwap) w = 1.0f - w;
      Spec16 c0 = col(idx - 1), c1 = col(idx);
      for (int b = 0; b < 16; ++b) out[b] = c0[b] * (1.0f - w) + c1[b] * w;
      return out;
    }
    case BGPU_TN_GRAPHPAPER: {  // graphPaper (Texture.hs:191-207)
      float x, z;
      map2d(n, dg, &x, &z);
      float xf = std::fabs(x - std::trunc(x)), zf = std::fabs(z - std::trunc(z));
      float lo = n.q[0] / 2.0f, hi = 1.0f - lo;
      bool line = xf < lo || zf < lo || xf > hi || zf > hi;
      return line ? spectrum(T, n.b, dg) : spectrum(T, n.a, dg);
    }
    default:
      return out;
  }
}

inline float scalar(const Textures& T, int node, const TexDG& dg) {
  const bgpu_tex_node& n = T.nodes[node];
  switch (n.op) {
    case BGPU_TN_SCONST: return n.q[0];
    case BGPU_TN_SIMAGE: {
      const Textures::Image& im = T.images[n.a];
      float u, v;
      map2d(n, dg, &u, &v);
      return im.texels[imageTexel(im, u, v)];
    }
    case BGPU_TN_CELL: return cellNoise(n.a, map3d(n, dg));
    case BGPU_TN_FBM: return fbm(n.a, n.q[0], map3d(n, dg));
    case BGPU_TN_PERLIN: {
      V3 p = map3d(n, dg);
      return perlin3d(p.x, p.y, p.z);
    }
    case BGPU_TN_CRYSTAL: {
      float u, v;
      map2d(n, dg, &u, &v);
      return quasiCrystal(n.a, u, v);
    }
    case BGPU_TN_SCALE: return n.q[0] + n.q[1] * scalar(T, n.a, dg);  // scaleTexture (Texture.hs:184-185)
    default: return 0.0f;
  }
}

}  // namespace tex
}  // namespace orc
