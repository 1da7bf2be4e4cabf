// Perez sun/sky light map (SunSky.hs), evaluated per direction.
//
// A sun/sky map is a continuous function of the map coordinates (u, v): the
// direction sphToDir (cartToSph (u, v)) in the light's frame gets the sky's
// colour, Perez' luminance and chromaticity distributions converted
// XYZ -> sRGB -> rgbToSpectrumIllum, plus the attenuated solar spectrum
// where the direction falls inside the sun's disk (SunSky.hs:17-24).  Only
// its sampling distribution is discrete (640 x 480 grid points).
//
// The scene keeps one block of kSkyFloats floats per sun/sky light in its map
// buffer: the solar spectrum, the seven illuminant RGB bases and the model's
// parameters.  sky_color() gives the colour at (u, v) as the coefficients of
// rgbToSpectrum's combination of bases (white*cw + (b1*c1 + b2*c2),
// Spectrum.hs:146-159) and whether the sun is seen, so a caller evaluates a
// band with three table reads instead of holding 16 values.
#pragma once

#include <cmath>

enum {
  SKY_SUN = 0,                // sunSpectrum' (16 bands; black with the sun below the horizon)
  SKY_BASES = 16,             // rgbIllum bases: white, cyan, magenta, yellow, red, green, blue
  SKY_PAR = SKY_BASES + 112,  // the model's parameters (SKYP_*)
  kSkyFloats = SKY_PAR + 36,
};
enum { SKY_W = 0, SKY_C = 16, SKY_M = 32, SKY_Y = 48, SKY_R = 64, SKY_G = 80, SKY_B = 96 };
// sunDir and sunTheta of initSky, the Perez coefficients of x, y and Y, the
// zenith values, the sun disk's cosine bound (sunThetaMax2) and the sun
// direction sunSpectrum tests against (normalized from the unnormalized
// input, so it may differ from sunDir in the last bit), then the daylight
// basis functions' XYZ s0XYZ, s1XYZ, s2XYZ (Spectrum.hs:234-241)
enum { SKYP_SUNDIR = 0, SKYP_SUNTHETA = 3, SKYP_PX = 4, SKYP_PY = 9, SKYP_PYY = 14, SKYP_ZX = 19, SKYP_ZY = 20,
       SKYP_ZYY = 21, SKYP_STM = 22, SKYP_DISK = 23, SKYP_S = 26 };
constexpr int kSkyMapW = 640, kSkyMapH = 480;  // mkDiscreteTextureMap2d (640, 480) (SunSky.hs:17)

struct SkyColor {
  float cw, c1, c2;  // rgbToSpectrum's coefficients of the white basis and of the bases at i1, i2
  int i1, i2;        // float offsets of those bases within the block's basis table
  bool sky;          // false: the sky term is black (direction below the horizon)
  bool sun;          // the direction is inside the sun's disk
};

// perez (SunSky.hs:99-104)
__host__ __device__ inline float sky_perez(const float* p, float sunT, float t, float g, float lvz) {
  const float csg = cosf(g), cst = cosf(sunT);
  const float num = (1.0f + p[0] * expf(p[1] / cosf(t))) * (1.0f + p[2] * expf(p[3] * g)) + (p[4] * csg) * csg;
  const float den = (1.0f + p[0] * expf(p[1])) * (1.0f + p[2] * expf(p[3] * sunT)) + (p[4] * cst) * cst;
  return (lvz * num) / den;
}

// rgbToSpectrum's choice of bases (Spectrum.hs:146-159)
__host__ __device__ inline void sky_rgb_bases(float r, float g, float b, SkyColor& c) {
  if (r <= g && r <= b) {
    c.cw = r;
    if (g <= b) { c.i1 = SKY_C; c.c1 = g - r; c.i2 = SKY_B; c.c2 = b - g; }
    else        { c.i1 = SKY_C; c.c1 = b - r; c.i2 = SKY_G; c.c2 = g - b; }
  } else if (g <= r && g <= b) {
    c.cw = g;
    if (r <= b) { c.i1 = SKY_M; c.c1 = r - g; c.i2 = SKY_B; c.c2 = b - r; }
    else        { c.i1 = SKY_M; c.c1 = b - g; c.i2 = SKY_R; c.c2 = r - b; }
  } else {
    c.cw = b;
    if (r <= b) { c.i1 = SKY_Y; c.c1 = r - b; c.i2 = SKY_G; c.c2 = g - r; }
    else        { c.i1 = SKY_Y; c.c1 = g - b; c.i2 = SKY_R; c.c2 = r - g; }
  }
}

// The sun/sky map at (u, v) (SunSky.hs:17-20), `par` = block + SKY_PAR.
__host__ __device__ inline SkyColor sky_color(const float* par, float u, float v) {
  const float phi = (u * 2.0f) * 3.14159265f, th = v * 3.14159265f;  // cartToSph (Types.hs:31-33)
  const float st = sinf(th);
  const float dx = st * cosf(phi), dy = st * sinf(phi), dz0 = cosf(th);  // sphToDir (Math.hs:141-148)
  SkyColor c;
  c.cw = c.c1 = c.c2 = 0.0f;
  c.i1 = c.i2 = 0;
  const float sx = par[SKYP_SUNDIR], sy = par[SKYP_SUNDIR + 1], sz = par[SKYP_SUNDIR + 2];
  const float dz = -dz0;  // skySpectrum (SunSky.hs:80-96)
  c.sky = !(dz < 1e-4f);
  if (c.sky) {
    const float theta = acosf(dz);
    const float cg = (dx * sx + dy * sy) + dz0 * sz;
    const float gamma = acosf(cg < -1.0f ? -1.0f : (cg > 1.0f ? 1.0f : cg));
    const float sunT = par[SKYP_SUNTHETA];
    const float x = sky_perez(par + SKYP_PX, sunT, theta, gamma, par[SKYP_ZX]);
    const float y = sky_perez(par + SKYP_PY, sunT, theta, gamma, par[SKYP_ZY]);
    const float yy = sky_perez(par + SKYP_PYY, sunT, theta, gamma, par[SKYP_ZYY]) * 1e-4f;
    // chromaticityToXYZ x y (Spectrum.hs:229-251)
    const float *s0 = par + SKYP_S, *s1 = s0 + 3, *s2 = s0 + 6;
    const float dd = (0.0241f + 0.2562f * x) - 0.7341f * y;
    const float m1 = ((-1.3515f - 1.7703f * x) + 5.9114f * y) / dd;
    const float m2 = ((0.03f - 31.4424f * x) + 30.0717f * y) / dd;
    const float cx = (s0[0] + m1 * s1[0]) + m2 * s2[0];
    const float cy = (s0[1] + m1 * s1[1]) + m2 * s2[1];
    const float cz = (s0[2] + m1 * s1[2]) + m2 * s2[2];
    const float X = (cx * yy) / cy, Z = (cz * yy) / cy;
    // xyzToRgb (Spectrum.hs:162-168)
    const float r = (3.240479f * X - 1.537150f * yy) - 0.498535f * Z;
    const float g = (-0.969256f * X + 1.875991f * yy) + 0.041556f * Z;
    const float b = (0.055648f * X - 0.204043f * yy) + 1.057311f * Z;
    sky_rgb_bases(r, g, b, c);
  }
  // sunSpectrum (SunSky.hs:106-113): (sunD * (1, 1, -1)) . dir > the disk's cosine bound
  const float* sd = par + SKYP_DISK;
  const float d = (sd[0] * dx + sd[1] * dy) + (-sd[2]) * dz0;
  c.sun = d > par[SKYP_STM];
  return c;
}

// One band of the map's spectrum: the sky's combination of bases plus the
// sun's spectrum, as skySpectrum + sunSpectrum adds them (a black term adds
// an exact zero and is skipped).
__host__ __device__ inline float sky_band(const float* blk, const SkyColor& c, int b) {
  const float* bs = blk + SKY_BASES;
  const float sky = c.sky ? bs[SKY_W + b] * c.cw + (bs[c.i1 + b] * c.c1 + bs[c.i2 + b] * c.c2) : 0.0f;
  if (!c.sun) return sky;
  return c.sky ? sky + blk[SKY_SUN + b] : blk[SKY_SUN + b];
}
