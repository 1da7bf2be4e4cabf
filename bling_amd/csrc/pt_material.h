// BxDF records of a material from its type, spectra and scalars
// (Material.hs:32-127, Microfacet.hs:127-134 fixExponent), shared by the host
// (constant materials at scene creation) and the device (materials whose
// inputs are textures, per hit).
#pragma once
#include "pt.h"

namespace bgpu {

// The inputs of one material: bgpu_material's type, three spectra and four
// scalars (the textured ones already evaluated).
struct MatIn {
  int32_t type;
  float scalar[4];
  float spec[3][16];
};

namespace matb {
__host__ __device__ inline float hmax(float x, float y) { return (x <= y) ? y : x; }
__host__ __device__ inline float hmin(float x, float y) { return (x <= y) ? x : y; }
__host__ __device__ inline float clamp(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }
__host__ __device__ inline void zero(DBxdf& b) {
  b.kind = b.type = b.fr = b.flip = 0;
  b.A = b.B = b.e = b.ei = b.et = 0.0f;
  b.pad2[0] = b.pad2[1] = b.pad2[2] = 0.0f;
  for (int k = 0; k < 16; ++k) b.r[k] = b.eta[k] = b.k[k] = 0.0f;
}
// mkBlinn (1 / rough) with fixExponent
__host__ __device__ inline float blinn_e(float rough) {
  float e = 1.0f / rough;
  return (e > 10000.0f || e != e) ? 10000.0f : e;
}
// frApproxEta r = (1 + r') / (1 - r'), r' = sqrt (sClamp 0 0.999 r), and
// frApproxK r = sqrt (refl / (1 - refl)) * 2, refl = sClamp 0 0.999 r, per
// band (Fresnel.hs:72-78)
__host__ __device__ inline void approx_eta_k(float r, float* eta, float* k) {
  const float refl = hmax(0.0f, hmin(0.999f, r));
  const float rs = sqrtf(refl);
  *eta = (1.0f + rs) / (1.0f - rs);
  *k = sqrtf(refl / (1.0f - refl)) * 2.0f;
}
// Oren-Nayar's A and B of sigma (Diffuse.hs:29-40); kind Lambert at sigma 0
__host__ __device__ inline void diffuse_lobe(DBxdf& b, float sig) {
  if (sig == 0.0f) {
    b.kind = K_LAMBERT;
  } else {
    b.kind = K_OREN;
    float s = clamp(sig, 0.0f, 1.0f);
    float s2 = s * s;
    b.A = 1.0f - (s2 / (2.0f * (s2 + 0.33f)));
    b.B = 0.45f * s2 / (s2 + 0.09f);
  }
}
}  // namespace matb

// The material's components; an unknown type has none (black).
__host__ __device__ inline void build_material(const MatIn& m, DMaterial& o) {
  using namespace matb;
  o.n = 0;
  o.tex_kind = 0;
  o.tex_map = 0;
  o.alt = 0;
  for (int k = 0; k < 12; ++k) o.tex[k] = 0.0f;
  zero(o.c[0]);
  zero(o.c[1]);
  switch (m.type) {
    case BGPU_MAT_MATTE: {
      DBxdf& b = o.c[o.n++];
      for (int k = 0; k < 16; ++k) b.r[k] = m.spec[0][k];
      diffuse_lobe(b, m.scalar[0]);
      b.type = BX_R | BX_D;
      break;
    }
    case BGPU_MAT_PLASTIC: {
      DBxdf& d = o.c[o.n++];
      d.kind = K_LAMBERT;
      d.type = BX_R | BX_D;
      for (int k = 0; k < 16; ++k) d.r[k] = m.spec[0][k];
      DBxdf& g = o.c[o.n++];
      g.kind = K_MF;
      g.type = BX_R | BX_G;
      for (int k = 0; k < 16; ++k) g.r[k] = m.spec[1][k];
      g.e = blinn_e(m.scalar[0]);
      g.fr = FR_DIEL;
      g.ei = 1.0f;
      g.et = 1.5f;
      break;
    }
    case BGPU_MAT_METAL: {
      DBxdf& g = o.c[o.n++];
      g.kind = K_MF;
      g.type = BX_R | BX_G;
      for (int k = 0; k < 16; ++k) {
        g.r[k] = 1.0f;
        g.eta[k] = m.spec[0][k];
        g.k[k] = m.spec[1][k];
      }
      g.e = blinn_e(m.scalar[0]);
      g.fr = FR_COND;
      break;
    }
    case BGPU_MAT_GLASS: {
      DBxdf& r = o.c[o.n++];
      r.kind = K_SPECR;
      r.type = BX_R | BX_S;
      r.fr = FR_DIEL;
      r.ei = 1.0f;
      r.et = m.scalar[0];
      DBxdf& t = o.c[o.n++];
      t.kind = K_SPECT;
      t.type = BX_T | BX_S;
      t.ei = 1.0f;
      t.et = m.scalar[0];
      for (int k = 0; k < 16; ++k) {  // sClamp' (Spectrum.hs:459-461)
        r.r[k] = hmax(0.0f, hmin(1.0f, m.spec[0][k]));
        t.r[k] = hmax(0.0f, hmin(1.0f, m.spec[1][k]));
      }
      break;
    }
    case BGPU_MAT_TRANSMATTE: {  // translucentMatte (Material.hs:43-53)
      DBxdf& r = o.c[o.n++];
      DBxdf& t = o.c[o.n++];
      for (int k = 0; k < 16; ++k) {
        r.r[k] = hmax(0.0f, hmin(1.0f, m.spec[0][k]));
        t.r[k] = hmax(0.0f, hmin(1.0f, m.spec[1][k])) * (1.0f - r.r[k]);
      }
      diffuse_lobe(r, m.scalar[0]);
      diffuse_lobe(t, m.scalar[0]);
      r.type = BX_R | BX_D;
      t.type = BX_T | BX_D;  // bxdfTypeFlip [Reflection, Transmission]
      t.flip = 1;
      break;
    }
    case BGPU_MAT_SUBSTRATE: {  // mkSubstrate (Material.hs:111-127)
      DBxdf& b = o.c[o.n++];
      b.kind = K_FB;
      b.type = BX_R | BX_G;
      for (int k = 0; k < 16; ++k) {  // sClamp 0 1
        b.r[k] = hmax(0.0f, hmin(1.0f, m.spec[0][k]));
        b.eta[k] = hmax(0.0f, hmin(1.0f, m.spec[1][k]));
        b.k[k] = hmax(0.0f, hmin(1.0f, m.spec[2][k]));
      }
      b.e = blinn_e(hmax(0.0f, m.scalar[0]));  // mkAnisotropic (1 / u) (1 / v), u = max 0 urough
      b.ei = blinn_e(hmax(0.0f, m.scalar[1]));
      b.et = m.scalar[2];
      break;
    }
    case BGPU_MAT_SHINYMETAL: {  // mkShinyMetal (Material.hs:98-109)
      DBxdf& g = o.c[o.n++];
      g.kind = K_MF;
      g.type = BX_R | BX_G;
      g.e = blinn_e(m.scalar[0]);
      g.fr = FR_COND;
      DBxdf& r = o.c[o.n++];
      r.kind = K_SPECR;
      r.type = BX_R | BX_S;
      r.fr = FR_COND;
      for (int k = 0; k < 16; ++k) {
        g.r[k] = r.r[k] = 1.0f;
        approx_eta_k(m.spec[1][k], &g.eta[k], &g.k[k]);
        approx_eta_k(m.spec[0][k], &r.eta[k], &r.k[k]);
      }
      break;
    }
    case BGPU_MAT_MIRROR: {
      DBxdf& r = o.c[o.n++];
      r.kind = K_SPECR;
      r.type = BX_R | BX_S;
      r.fr = FR_NOOP;
      for (int k = 0; k < 16; ++k) r.r[k] = hmax(0.0f, hmin(1.0f, m.spec[0][k]));
      break;
    }
    default:
      break;
  }
}

}  // namespace bgpu
