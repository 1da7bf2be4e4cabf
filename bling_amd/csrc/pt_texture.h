// Device evaluation of the texture node list (bgpu_tex_node), Texture.hs:59-414
// as the material parser builds the textures (IO/MaterialParser.hs:106-226).
//
// A texture is evaluated over the node range [lo, root] of its root (every
// child precedes its node, and scene creation bounds the range by
// kMaxTexRange), in two passes so that no spectrum is held per node:
//   pass A (once): scalar values, and per spectrum node what the band pass
//                  needs (a selection, a blend or gradient weight, a texel);
//   pass B (per band): the band's value of every spectrum node.
// Each node's arithmetic restates the reference's Float expression in its
// order, so values equal the oracle's bit for bit, up to cosf/sinf ulps in
// quasiCrystal.
#pragma once
#include "noise_table.h"
#include "pt.h"

namespace bgpu {
namespace tx {

constexpr int kMaxTexRange = 64;

// The point and surface parameters a texture reads (dgP, dgU, dgV).
struct TexPt {
  f3 p;
  float u, v;
};

// floor to an integer, 0 for values no 64-bit integer holds (NaN included)
__device__ __forceinline__ long long floor_ll(float x) {
  return fabsf(x) < 4.6e18f ? (long long)floorf(x) : 0ll;
}

// lerp (Math.hs:108-110)
__device__ __forceinline__ float lerp_t(float t, float a, float b) { return (1.0f - t) * a + t * b; }

// noiseWeight (Texture.hs:388-391)
__device__ __forceinline__ float noise_weight(float t) {
  float t3 = t * t * t;
  float t4 = t3 * t;
  return 6.0f * t4 * t - 15.0f * t4 + 10.0f * t3;
}

// grad (Texture.hs:393-400)
__device__ __forceinline__ float noise_grad(int x, int y, int z, float dx, float dy, float dz) {
  const int h = kNoisePermTable[kNoisePermTable[kNoisePermTable[x] + y] + z] & 15;
  const float uu = (h < 8 || h == 12 || h == 13) ? dx : dy;
  const float vv = (h < 4 || h == 12 || h == 13) ? dy : dz;
  return ((h & 1) != 0 ? -uu : uu) + ((h & 2) != 0 ? -vv : vv);
}

// perlin3d (Texture.hs:358-386)
__device__ __forceinline__ float perlin3d(float x, float y, float z) {
  const long long fx = floor_ll(x), fy = floor_ll(y), fz = floor_ll(z);
  const float dx = x - (float)fx, dy = y - (float)fy, dz = z - (float)fz;
  const int ix = (int)(fx & 255), iy = (int)(fy & 255), iz = (int)(fz & 255);
  const float w000 = noise_grad(ix, iy, iz, dx, dy, dz);
  const float w100 = noise_grad(ix + 1, iy, iz, dx - 1.0f, dy, dz);
  const float w010 = noise_grad(ix, iy + 1, iz, dx, dy - 1.0f, dz);
  const float w110 = noise_grad(ix + 1, iy + 1, iz, dx - 1.0f, dy - 1.0f, dz);
  const float w001 = noise_grad(ix, iy, iz + 1, dx, dy, dz - 1.0f);
  const float w101 = noise_grad(ix + 1, iy, iz + 1, dx - 1.0f, dy, dz - 1.0f);
  const float w011 = noise_grad(ix, iy + 1, iz + 1, dx, dy - 1.0f, dz - 1.0f);
  const float w111 = noise_grad(ix + 1, iy + 1, iz + 1, dx - 1.0f, dy - 1.0f, dz - 1.0f);
  const float wx = noise_weight(dx), wy = noise_weight(dy), wz = noise_weight(dz);
  const float x00 = lerp_t(wx, w000, w100), x10 = lerp_t(wx, w010, w110);
  const float x01 = lerp_t(wx, w001, w101), x11 = lerp_t(wx, w011, w111);
  return lerp_t(wz, lerp_t(wy, x00, x10), lerp_t(wy, x01, x11));
}

// fbm (Texture.hs:342-350)
__device__ __forceinline__ float fbm(int octaves, float omega, f3 p) {
  float s = 0.0f, l = 1.0f, o = 1.0f;
  for (int k = 0; k < octaves; ++k) {
    s = s + o * perlin3d(p.x * l, p.y * l, p.z * l);
    l = 1.99f * l;
    o = omega * o;
  }
  return s;
}

// Worley's cell noise (Texture.hs:253-307): the lcg and hash values stay in
// [0, 2^32), so `rem` is a mask
__device__ __forceinline__ long long cell_lcg(long long x) { return (1103515245ll * x + 12345ll) & 0xffffffffll; }
__device__ __forceinline__ long long cell_hash(long long x, long long y, long long z) {
  const long long h = (x * 73856093ll) ^ (y * 19349663ll) ^ (z * 83492791ll);
  const unsigned long long a = h < 0 ? 0ull - (unsigned long long)h : (unsigned long long)h;
  return (long long)(a % 4294967296ull);
}
__device__ __forceinline__ int cell_count(long long v) {  // prob
  return v < 393325350ll ? 1 : v < 1022645910ll ? 2 : v < 1861739990ll ? 3 : v < 2700834071ll ? 4
       : v < 3372109335ll ? 5 : v < 3819626178ll ? 6 : v < 4075350088ll ? 7 : v < 4203212043ll ? 8 : 9;
}
__device__ __forceinline__ float cell_dist(int kind, f3 a, f3 b) {
  const float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
  if (kind == 1) return dx * dx + dy * dy + dz * dz;
  if (kind == 2) return fabsf(dx) + fabsf(dy) + fabsf(dz);
  if (kind == 3) {
    float m = fabsf(dx);
    m = m <= fabsf(dy) ? fabsf(dy) : m;
    return m <= fabsf(dz) ? fabsf(dz) : m;
  }
  return __builtin_sqrtf(dx * dx + dy * dy + dz * dz);
}
__device__ __forceinline__ float cell_noise(int kind, f3 p) {
  const long long ox = floor_ll(p.x), oy = floor_ll(p.y), oz = floor_ll(p.z);
  bool first = true;
  float best = 0.0f;
  for (long long x = ox - 1; x <= ox + 1; ++x)
    for (long long y = oy - 1; y <= oy + 1; ++y)
      for (long long z = oz - 1; z <= oz + 1; ++z) {
        long long u = cell_lcg(cell_hash(x, y, z));
        const int n = cell_count(u);
        for (int k = 0; k < n; ++k) {
          const long long u1 = cell_lcg(u), u2 = cell_lcg(u1), u3 = cell_lcg(u2);
          const f3 q = mk3((float)x + (float)u1 / 4294967296.0f, (float)y + (float)u2 / 4294967296.0f,
                           (float)z + (float)u3 / 4294967296.0f);
          u = u3;
          const float d = cell_dist(kind, p, q);
          if (first || !(best <= d)) best = d;
          first = false;
        }
      }
  return best;
}

// quasiCrystal (Texture.hs:309-340), angles as GHC 8.0's enumFromThen makes them
__device__ __forceinline__ float quasi_crystal(int octaves, float x, float y) {
  const float step = 3.14159265358979323846f / (float)octaves;
  float s = 0.0f, t1 = 0.0f, t2 = 0.0f;
  for (int k = 0; k < octaves; ++k) {
    const float th = k == 0 ? 0.0f : (k == 1 ? step : (t1 + t1) - t2);
    t2 = t1;
    t1 = th;
    s = s + (cosf(cosf(th) * x + sinf(th) * y) + 1.0f) / 2.0f;
  }
  const float tr = truncf(s);
  long long k = floor_ll(tr);
  float v = s - tr;
  if (v < 0.0f) {
    k = k - 1;
    v = 1.0f + v;
  }
  return (k & 1) != 0 ? 1.0f - v : v;
}

// The 2d mappings (Texture.hs:162-179) and the 3d identity mapping (:150-156).
__device__ __forceinline__ void map2d(const bgpu_tex_node& n, const TexPt& t, float* u, float* v) {
  if (n.map == BGPU_MAP_PLANAR) {
    *u = (t.p.x * n.m[0] + t.p.y * n.m[1] + t.p.z * n.m[2]) + n.m[6];
    *v = (t.p.x * n.m[3] + t.p.y * n.m[4] + t.p.z * n.m[5]) + n.m[7];
  } else {
    *u = n.m[0] * t.u + n.m[2];
    *v = n.m[1] * t.v + n.m[3];
  }
}
__device__ __forceinline__ f3 map3d(const bgpu_tex_node& n, const TexPt& t) {  // transPoint
  const float* m = n.m;
  const float xp = m[0] * t.p.x + m[1] * t.p.y + m[2] * t.p.z + m[3];
  const float yp = m[4] * t.p.x + m[5] * t.p.y + m[6] * t.p.z + m[7];
  const float zp = m[8] * t.p.x + m[9] * t.p.y + m[10] * t.p.z + m[11];
  const float wp = m[12] * t.p.x + m[13] * t.p.y + m[14] * t.p.z + m[15];
  return wp == 1.0f ? mk3(xp, yp, zp) : mk3(xp / wp, yp / wp, zp / wp);
}
// getPixel (Texture.hs:89-104): the texel's index in its image
__device__ __forceinline__ long long image_texel(int4 im, float u, float v) {
  long long px = floor_ll(u * (float)im.x) % im.x;
  px = px < 0 ? px + im.x : px;
  long long py = floor_ll((-v) * (float)im.y) % im.y;
  py = py < 0 ? py + im.y : py;
  return py * im.x + px;
}

// The scene's texture tables.
struct Tex {
  const bgpu_tex_node* nodes;
  const int4* images;  // width, height, channels, offset of the first texel
  const float* texels;
};

// Pass A over [lo, root]: scalar values in sv, and per spectrum node the
// band pass's input (sv: blend or gradient weight; iv: selection, texel
// offset or gradient segment).
__device__ inline void pass_a(const Tex& T, int lo, int root, const TexPt& t, float* sv, int* iv) {
  for (int k = lo; k <= root; ++k) {
    const bgpu_tex_node& n = T.nodes[k];
    float s = 0.0f;
    int i = 0;
    switch (n.op) {
      case BGPU_TN_SCONST: s = n.q[0]; break;
      case BGPU_TN_SIMAGE: {
        float u, v;
        map2d(n, t, &u, &v);
        const int4 im = T.images[n.a];
        s = T.texels[im.w + image_texel(im, u, v)];
        break;
      }
      case BGPU_TN_CELL: s = cell_noise(n.a, map3d(n, t)); break;
      case BGPU_TN_FBM: s = fbm(n.a, n.q[0], map3d(n, t)); break;
      case BGPU_TN_PERLIN: {
        const f3 p = map3d(n, t);
        s = perlin3d(p.x, p.y, p.z);
        break;
      }
      case BGPU_TN_CRYSTAL: {
        float u, v;
        map2d(n, t, &u, &v);
        s = quasi_crystal(n.a, u, v);
        break;
      }
      case BGPU_TN_SCALE: s = n.q[0] + n.q[1] * sv[n.a - lo]; break;
      case BGPU_TN_IMAGE: {
        float u, v;
        map2d(n, t, &u, &v);
        const int4 im = T.images[n.a];
        i = (int)(im.w + 16 * image_texel(im, u, v));
        break;
      }
      case BGPU_TN_BLEND: s = sv[n.c - lo]; break;
      case BGPU_TN_CHECKER: {
        const long long c = floor_ll(t.p.x * n.q[0]) + floor_ll(t.p.y * n.q[1]) + floor_ll(t.p.z * n.q[2]);
        i = ((c % 2) + 2) % 2 == 0 ? 0 : 1;
        break;
      }
      case BGPU_TN_GRADIENT: {  // i: -1 first colour, -2 last, -3 none (NaN), else the segment's end
        const float f = sv[n.a - lo];
        float glo = T.nodes[n.b].q[0], ghi = glo;
        for (int j = 1; j < n.c; ++j) {
          const float p = T.nodes[n.b + j].q[0];
          glo = p < glo ? p : glo;
          ghi = p > ghi ? p : ghi;
        }
        if (f <= glo) {
          i = -1;
        } else if (f >= ghi) {
          i = -2;
        } else {
          i = -3;
          for (int j = 0; j < n.c; ++j)
            if (T.nodes[n.b + j].q[0] > f) {
              i = j;
              break;
            }
          if (i <= 0) {
            i = -3;
          } else {
            const float p0 = T.nodes[n.b + i - 1].q[0], p1 = T.nodes[n.b + i].q[0];
            s = (f - p0) / (p1 - p0);
          }
        }
        break;
      }
      case BGPU_TN_GRAPHPAPER: {
        float x, z;
        map2d(n, t, &x, &z);
        const float xf = fabsf(x - truncf(x)), zf = fabsf(z - truncf(z));
        const float l = n.q[0] / 2.0f, h = 1.0f - l;
        i = (xf < l || zf < l || xf > h || zf > h) ? 1 : 0;
        break;
      }
      default: break;
    }
    sv[k - lo] = s;
    iv[k - lo] = i;
  }
}

// Pass B for band b: the value of every spectrum node of [lo, root].
__device__ inline float pass_b(const Tex& T, int lo, int root, int b, const float* sv, const int* iv, float* w) {
  for (int k = lo; k <= root; ++k) {
    const bgpu_tex_node& n = T.nodes[k];
    float r = 0.0f;
    switch (n.op) {
      case BGPU_TN_CONST: r = n.spec[b]; break;
      case BGPU_TN_IMAGE: r = T.texels[iv[k - lo] + b]; break;
      case BGPU_TN_BLEND: {  // spectrumBlend (Texture.hs:133-145)
        const float x = sv[k - lo];
        const float v1 = w[n.a - lo], v2 = w[n.b - lo];
        r = x <= 0.0f ? v1 : (x >= 1.0f ? v2 : v1 * (1.0f - x) + v2 * x);
        break;
      }
      case BGPU_TN_CHECKER: r = iv[k - lo] == 0 ? w[n.a - lo] : w[n.b - lo]; break;
      case BGPU_TN_GRAPHPAPER: r = iv[k - lo] == 1 ? w[n.b - lo] : w[n.a - lo]; break;
      case BGPU_TN_GRADIENT: {  // gradient (Texture.hs:236-247)
        const int i = iv[k - lo];
        if (i == -1) {
          r = T.nodes[n.b].spec[b];
        } else if (i == -2) {
          r = T.nodes[n.b + n.c - 1].spec[b];
        } else if (i > 0) {
          const float wt = sv[k - lo];
          r = T.nodes[n.b + i - 1].spec[b] * (1.0f - wt) + T.nodes[n.b + i].spec[b] * wt;
        }
        break;
      }
      default: break;
    }
    w[k - lo] = r;
  }
  return w[root - lo];
}

__device__ inline float eval_scalar(const Tex& T, int lo, int root, const TexPt& t) {
  float sv[kMaxTexRange];
  int iv[kMaxTexRange];
  pass_a(T, lo, root, t, sv, iv);
  return sv[root - lo];
}
__device__ inline void eval_spectrum(const Tex& T, int lo, int root, const TexPt& t, float* out) {
  float sv[kMaxTexRange], w[kMaxTexRange];
  int iv[kMaxTexRange];
  pass_a(T, lo, root, t, sv, iv);
  for (int b = 0; b < 16; ++b) out[b] = pass_b(T, lo, root, b, sv, iv, w);
}

}  // namespace tx
}  // namespace bgpu
