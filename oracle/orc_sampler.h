// ORACLE -- test infrastructure only (see orc_math.h).
// Counter-based restatement of the stratified sampler of
// src/lib/Graphics/Bling/Sampling.hs:85-221 (runSample Stratified,
// stratified1D/2D, fill, rnd', rnd2D').  The reference draws its jitter and
// shuffles from an entropy-seeded MWC stream; here every value is a pure
// function of (job seed, pass, pixel, sample, dimension):
//   mix64        -- the splitmix64 finalizer (S. Vigna, public domain)
//   permute      -- a keyed bijection of [0, n) by cycle-walking a bijective
//                   integer mix on the enclosing power of two, standing in
//                   for `shuffle` (Random.hs:79-89)
// The GPU sampler implements the same specification.
#pragma once
#include <cstdint>

namespace orc {

inline uint64_t mix64(uint64_t x) {
  x ^= x >> 30;
  x *= 0xbf58476d1ce4e5b9ULL;
  x ^= x >> 27;
  x *= 0x94d049bb133111ebULL;
  x ^= x >> 31;
  return x;
}

inline uint64_t pixelKey(uint64_t seed, uint32_t pass, int ix, int iy) {
  uint64_t c = ((uint64_t)pass << 42) | ((uint64_t)(uint32_t)(iy + 1024) << 21) |
               (uint64_t)(uint32_t)(ix + 1024);
  return mix64(mix64(seed) ^ mix64(c + 0x9e3779b97f4a7c15ULL));
}

inline uint64_t rngBits(uint64_t key, uint32_t dim, uint32_t idx) {
  return mix64(key + 0x9e3779b97f4a7c15ULL * (((uint64_t)dim << 24) + (uint64_t)idx + 1ULL));
}

// [0, 1) with 24 bits
inline float uni(uint64_t bits) { return (float)(uint32_t)(bits >> 40) * 0x1p-24f; }

inline uint32_t permute(uint32_t i, uint32_t n, uint32_t p) {
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

// dimension ids
enum : uint32_t {
  D_CAM_X = 1, D_CAM_Y = 2, D_LENS_X = 3, D_LENS_Y = 4, D_LENS_P = 5,
  D_T1_J = 0x1000, D_T1_P = 0x2000, D_T2_JX = 0x3000, D_T2_JY = 0x4000, D_T2_P = 0x5000,
  D_R1 = 0x6000, D_R2X = 0x7000, D_R2Y = 0x8000
};

static const float kAlmostOne = 0x1.fffffep-1f;  // Sampling.hs:154-155

struct PixelSampler {
  uint64_t key;
  uint32_t nu, nv, N;
  uint32_t n1d, n2d;  // table sizes 4*sd, 3*sd (Integrator/Path.hs:30-39)

  // stratified2D entry i (Sampling.hs:164-171): (u, v) = quotRem i nu
  void strat2(uint32_t i, float ju, float jv, float* u, float* v) const {
    float du = 1.0f / (float)nu, dv = 1.0f / (float)nv;
    float a = ((float)(i / nu) + ju) * du, b = ((float)(i % nu) + jv) * dv;
    *u = a < kAlmostOne ? a : kAlmostOne;
    *v = b < kAlmostOne ? b : kAlmostOne;
  }
  // camera sample n (pixel offsets, not shuffled; Sampling.hs:117,128-132)
  void camera(uint32_t n, float* ox, float* oy) const {
    strat2(n, uni(rngBits(key, D_CAM_X, n)), uni(rngBits(key, D_CAM_Y, n)), ox, oy);
  }
  void lens(uint32_t n, float* u, float* v) const {  // shuffled (Sampling.hs:118-121)
    uint32_t i = permute(n, N, (uint32_t)rngBits(key, D_LENS_P, 0));
    strat2(i, uni(rngBits(key, D_LENS_X, i)), uni(rngBits(key, D_LENS_Y, i)), u, v);
  }
  // rnd' k (Sampling.hs:203-211)
  float get1(uint32_t n, uint32_t k) const {
    if (k < n1d) {
      uint32_t i = permute(n, N, (uint32_t)rngBits(key, D_T1_P + k, 0));
      float du = 1.0f / (float)N;
      float a = ((float)i + uni(rngBits(key, D_T1_J + k, i))) * du;  // stratified1D
      return a < kAlmostOne ? a : kAlmostOne;
    }
    return uni(rngBits(key, D_R1 + k, n));
  }
  // rnd2D' k (Sampling.hs:213-221)
  void get2(uint32_t n, uint32_t k, float* u, float* v) const {
    if (k < n2d) {
      uint32_t i = permute(n, N, (uint32_t)rngBits(key, D_T2_P + k, 0));
      strat2(i, uni(rngBits(key, D_T2_JX + k, i)), uni(rngBits(key, D_T2_JY + k, i)), u, v);
      return;
    }
    *u = uni(rngBits(key, D_R2X + k, n));
    *v = uni(rngBits(key, D_R2Y + k, n));
  }
};

}  // namespace orc
