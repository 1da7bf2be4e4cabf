// Host preparation of the path tracer's per-scene data:
//   materials  -> BxDF records (Material.hs:32-96, Microfacet.hs:127-134 fixExponent)
//   lights     -> explicit lights, then geometric (emissive shape) lights in
//                 shape order, the list of mkScene (Scene.hs:37-43)
//   env maps   -> the Dist2D of mkInfiniteAreaLight (Light.hs:76-77,
//                 Montecarlo.hs:34-104), built in binary32 in the same order;
//                 a sun/sky map's parameter block (SunSky.hs)
//   film       -> 16x16 filter table (Image.hs:48-61) and sample extent (Image.hs:162-168)
#include <algorithm>
#include <cmath>
#include <cstring>

#include "bgpu_internal.h"
#include "pt_material.h"
#include "pt_sky.h"
#include "pt_texture.h"
#include "spectral_tables.h"

namespace bgpu {

namespace {

inline float hmaxf(float x, float y) { return (x <= y) ? y : x; }

float sY(const float* s) {  // Spectrum.hs:371-373
  float a = 0.0f;
  for (int i = 0; i < 16; ++i) a = a + s[i] * BGPU_CIE_Y[i];
  return a / BGPU_CIE_Y_SUM;
}

const float kPiF = 3.14159265358979323846f;

float eval_filter(const bgpu_film& f, float x, float y) {  // Filter.hs:68-96
  switch (f.filter) {
    case BGPU_FILTER_GAUSS: {
      float w = f.fparams[0], h = f.fparams[1], a = f.fparams[2];
      float ex = std::exp(-a * w * w), ey = std::exp(-a * h * h);
      return hmaxf(0.0f, std::exp(-a * x * x) - ex) * hmaxf(0.0f, std::exp(-a * y * y) - ey);
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
        float xp = std::fabs(v) * kPiF;
        return (std::sin(xp) / xp) * (std::sin(xp * tau) / (xp * tau));
      };
      return s1(x * ix) * s1(y * iy);
    }
    case BGPU_FILTER_TRIANGLE: {
      float w = f.fparams[0], h = f.fparams[1];
      return (hmaxf(0.0f, w - std::fabs(x * 2.0f)) * hmaxf(0.0f, h - std::fabs(y * 2.0f))) / (w * h);
    }
    default:
      return (std::fabs(x) < 0.5f && std::fabs(y) < 0.5f) ? 1.0f : 0.0f;
  }
}

// Dist1D [n, func[n], cdf[n+1], fi] (Montecarlo.hs:40-48)
void push_dist1d(std::vector<float>& out, const std::vector<float>& f) {
  int n = (int)f.size();
  float nb;
  std::memcpy(&nb, &n, 4);
  out.push_back(nb);
  for (float v : f) out.push_back(v);
  std::vector<float> i(n + 1);
  i[0] = 0.0f;
  for (int k = 0; k < n; ++k) i[k + 1] = i[k] + f[k] / (float)n;
  float fi = i[n];
  for (int k = 0; k <= n; ++k) out.push_back(fi != 0.0f ? i[k] / fi : (float)k / (float)n);
  out.push_back(fi);
}

float dist1d_fi(const std::vector<float>& f) {
  int n = (int)f.size();
  float acc = 0.0f;
  for (int k = 0; k < n; ++k) acc = acc + f[k] / (float)n;
  return acc;
}

// A texel map's value at (u, v): rgbfToTexMap (IO/Bitmap.hs:21-28)
const float* tex_px(const bgpu_light& l, float u, float v) {
  int px = (int)std::floor((1.0f - u) * (float)l.width);
  int py = (int)std::floor((1.0f - v) * (float)l.height);
  px = px < 0 ? 0 : (px > l.width - 1 ? l.width - 1 : px);
  py = py < 0 ? 0 : (py > l.height - 1 ? l.height - 1 : py);
  return l.radiance + 16 * ((size_t)py * l.width + px);
}

// ---------------------------------------------------------------- sun/sky
struct V3 {
  float x, y, z;
};
V3 v_norm(V3 v) {  // normalize (Math.hs:349-353)
  const float sq = (v.x * v.x + v.y * v.y) + v.z * v.z;
  if (sq == 0.0f) return {0.0f, 1.0f, 0.0f};
  const float il = 1.0f / std::sqrt(sq);
  return {v.x * il, v.y * il, v.z * il};
}
V3 v_cross(V3 u, V3 v) { return {u.y * v.z - u.z * v.y, -(u.x * v.z - u.z * v.x), u.x * v.y - u.y * v.x}; }
float v_dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }

// evalSpd of a regular SPD (Spectrum.hs:271-280)
float spd_regular(const float* a, int n, float l0, float l1, float l) {
  if (l <= l0) return a[0];
  if (l >= l1) return a[n - 1];
  const float d1 = 1.0f / ((l1 - l0) / (float)(n - 1));
  const float x = (l - l0) * d1;
  const int b0 = (int)std::floor(x);
  const int b1 = std::min(b0 + 1, n - 1);
  const float dx = x - (float)b0;
  return (1.0f - dx) * a[b0] + dx * a[b1];
}
// evalSpd of an irregular SPD, its binary search included (Spectrum.hs:258-269)
float spd_irregular(const float* ls, const float* vs, int n, float l) {
  if (l <= ls[0]) return vs[0];
  if (l >= ls[n - 1]) return vs[n - 1];
  int lo = 0, hi = n - 1;
  for (;;) {
    const int mid = (lo + hi) / 2;
    if (lo == mid) break;
    if (ls[mid] == l) {
      lo = mid;
      break;
    }
    if (ls[mid] < l) lo = mid;
    else hi = mid;
  }
  const float t = (l - ls[lo]) / (ls[lo + 1] - ls[lo]);
  return (1.0f - t) * vs[lo] + t * vs[lo + 1];  // lerp (Math.hs:108-110)
}

// The attenuated solar spectrum's SPD at lambda l (sunSpectrum', SunSky.hs:119-155)
float sun_spd(float l, float m, float turb) {
  const float tR = std::exp(-((m * 0.008735f) * std::pow(l / 1000.0f, -4.08f)));
  const float beta = 0.04608365822050f * turb - 0.04586025928522f;
  const float tA = std::exp(-((m * beta) * std::pow(l / 1000.0f, -1.3f)));
  const float tO = std::exp(-((m * spd_irregular(BGPU_SKY_KO_L, BGPU_SKY_KO_A, 64, l)) * 0.35f));
  const float kg = spd_irregular(BGPU_SKY_KG_L, BGPU_SKY_KG_A, 4, l);
  const float tG = std::exp(-(((1.41f * kg) * m) / std::pow(1.0f + (118.93f * kg) * m, 0.45f)));
  const float kwa = spd_irregular(BGPU_SKY_KWA_L, BGPU_SKY_KWA_A, 13, l);
  const float tWA = std::exp(-((((0.2385f * kwa) * 2.0f) * m) / std::pow(1.0f + ((20.07f * kwa) * 2.0f) * m, 0.45f)));
  return ((((spd_regular(BGPU_SKY_SOL, 38, 380.0f, 750.0f, l) * tR) * tA) * tO) * tG) * tWA;
}

// The block sky_color reads (pt_sky.h) for mkSunSkyLight east sunDir turbidity:
// initSky (SunSky.hs:49-74), sunSpectrum' (:115-155), sunThetaMax2 (:42-47).
void sky_init(const float* east, const float* sun, float t, float* blk) {
  std::memset(blk, 0, sizeof(float) * kSkyFloats);
  const float* bases[7] = {BGPU_RGB_ILLUM_WHITE, BGPU_RGB_ILLUM_CYAN, BGPU_RGB_ILLUM_MAGENTA, BGPU_RGB_ILLUM_YELLOW,
                           BGPU_RGB_ILLUM_RED,   BGPU_RGB_ILLUM_GREEN, BGPU_RGB_ILLUM_BLUE};
  for (int k = 0; k < 7; ++k) std::memcpy(blk + SKY_BASES + 16 * k, bases[k], 64);
  // basis = coordinateSystem' (normalize up) (normalize east), up = (0, 1, 0) (Math.hs:427-432)
  const V3 w = v_norm(v_norm({0.0f, 1.0f, 0.0f}));
  const V3 u = v_norm(v_cross(v_norm({east[0], east[1], east[2]}), w));
  const V3 v = v_cross(w, u);
  auto to_local = [&](V3 d) { return V3{v_dot(d, u), v_dot(d, v), v_dot(d, w)}; };  // worldToLocal
  const V3 sdw = {sun[0], sun[1], sun[2]};
  const V3 sd = v_norm(to_local(v_norm(sdw)));
  const V3 disk = v_norm(to_local(sdw));
  const float z = sd.z;
  const float st = std::acos(z < -1.0f ? -1.0f : (z > 1.0f ? 1.0f : z));
  const float st2 = st * st, st3 = (st * st) * st, t2 = t * t;
  const float chi = (4.0f / 9.0f - t / 120.0f) * (kPiF - 2.0f * st);
  float* par = blk + SKY_PAR;
  par[SKYP_SUNDIR] = sd.x, par[SKYP_SUNDIR + 1] = sd.y, par[SKYP_SUNDIR + 2] = sd.z;
  par[SKYP_SUNTHETA] = st;
  const float py[5] = {0.17872f * t - 1.46303f, -0.35540f * t + 0.42749f, -0.02266f * t + 5.32505f,
                       0.12064f * t - 2.57705f, -0.06696f * t + 0.37027f};
  const float px[5] = {-0.01925f * t - 0.25922f, -0.06651f * t + 0.00081f, -0.00041f * t + 0.21247f,
                       -0.06409f * t - 0.89887f, -0.00325f * t + 0.04517f};
  const float pyc[5] = {-0.01669f * t - 0.26078f, -0.09495f * t + 0.00921f, -0.00792f * t + 0.21023f,
                        -0.04405f * t - 1.65369f, -0.01092f * t + 0.05291f};
  std::memcpy(par + SKYP_PX, px, sizeof(px));
  std::memcpy(par + SKYP_PY, pyc, sizeof(pyc));
  std::memcpy(par + SKYP_PYY, py, sizeof(py));
  par[SKYP_ZYY] = ((((4.04530f * t - 4.97100f) * std::tan(chi)) - 0.2155f * t) + 2.4192f) * 1000.0f;
  par[SKYP_ZX] = ((((0.00165f * st3 - 0.00374f * st2) + 0.00208f * st) * t2) +
                  ((((-0.02902f * st3 + 0.06377f * st2) - 0.03202f * st) + 0.00394f) * t)) +
                 (((0.11693f * st3 - 0.21196f * st2) + 0.06052f * st) + 0.25885f);
  par[SKYP_ZY] = ((((0.00275f * st3 - 0.00610f * st2) + 0.00316f * st) * t2) +
                  ((((-0.04212f * st3 + 0.08970f * st2) - 0.04153f * st) + 0.00515f) * t)) +
                 (((0.15346f * st3 - 0.26756f * st2) + 0.06669f * st) + 0.26688f);
  const float sint2 = 6.955e5f / 1.496e8f;
  par[SKYP_STM] = std::sqrt(hmaxf(0.0f, 1.0f - sint2));
  par[SKYP_DISK] = disk.x, par[SKYP_DISK + 1] = disk.y, par[SKYP_DISK + 2] = disk.z;
  std::memcpy(par + SKYP_S, BGPU_S0_XYZ, 12);
  std::memcpy(par + SKYP_S + 3, BGPU_S1_XYZ, 12);
  std::memcpy(par + SKYP_S + 6, BGPU_S2_XYZ, 12);
  if (!(sd.z < 0.0f)) {  // below the horizon the sun is black
    const float m = 1.0f / (std::cos(st) + 0.000940f * std::pow(1.6386f - st, -1.253f));
    for (int i = 0; i < 16; ++i) {  // fromSpd of an SPD function: the band ends' mean (Spectrum.hs:316,329-335)
      const float a = (float)i / 16.0f, b = (float)(i + 1) / 16.0f;
      const float l0 = (1.0f - a) * 400.0f + a * 700.0f, l1 = (1.0f - b) * 400.0f + b * 700.0f;
      blk[SKY_SUN + i] = (sun_spd(l0, m, t) + sun_spd(l1, m, t)) * 0.5f;
    }
  }
}

}  // namespace

int build_pt_data(Ctx* ctx, const bgpu_scene_desc* d, Scene* sc) {
  hipStream_t st = ctx->stream;
  PtScene& pt = sc->pt;
  DevPtScene& H = pt.host;
  std::memset(&H, 0, sizeof(H));
  const uint32_t nt = d->num_triangles, nv = d->num_vertices, ns = d->num_shapes;

  // geometry for shading
  if (nt) {
    if (int rc = sc->tri_idx.upload(ctx, d->indices, 12ull * nt, st)) return rc;
    if (int rc = sc->verts.upload(ctx, d->vertices, 12ull * nv, st)) return rc;
    if (d->normals)
      if (int rc = sc->normals.upload(ctx, d->normals, 12ull * nv, st)) return rc;
    if (d->uvs)
      if (int rc = sc->uvs.upload(ctx, d->uvs, 8ull * nv, st)) return rc;
    std::vector<int32_t> info(2ull * nt);
    uint32_t nmesh = d->num_meshes;
    for (uint32_t t = 0; t < nt; ++t) {
      int m = d->tri_mesh ? d->tri_mesh[t] : 0;
      bgpu_mesh mesh = (nmesh && d->meshes) ? d->meshes[m] : bgpu_mesh{0, 0};
      info[2 * t] = mesh.material;
      info[2 * t + 1] = mesh.flags;
    }
    if (int rc = sc->tri_mat.upload(ctx, info.data(), info.size() * 4, st)) return rc;
  }
  std::vector<DShapeX> sx(std::max<uint32_t>(ns, 1));
  for (uint32_t s = 0; s < ns; ++s) std::memcpy(sx[s].o2w, d->shapes[s].o2w, sizeof(sx[s].o2w));
  if (int rc = sc->shape_x.upload(ctx, sx.data(), sx.size() * sizeof(DShapeX), st)) return rc;

  // materials (Material.hs:32-96)
  std::vector<DMaterial> mats(std::max<uint32_t>(d->num_materials, 1));
  // one material from its type and scalars, with spectra spec[0], spec[1]
  // (the constant ones, or a selector texture's tex2 alternates)
  auto build = [&](const bgpu_material& m, const float (*spec)[16]) {
    MatIn in;
    in.type = m.type;
    for (int k = 0; k < 4; ++k) in.scalar[k] = m.scalar[k];
    for (int k = 0; k < 16; ++k) {
      in.spec[0][k] = spec[0][k];
      in.spec[1][k] = spec[1][k];
      in.spec[2][k] = m.spec[2][k];
    }
    DMaterial o;
    std::memset(&o, 0, sizeof(o));
    build_material(in, o);
    return o;
  };
  // the texture node list: every child precedes its node and has the kind
  // its operation takes; lo[k] is the lowest node of k's subtree, and a
  // texture is evaluated over [lo, root], at most tx::kMaxTexRange nodes
  const uint32_t nn = d->num_tex_nodes;
  if (nn && !d->tex_nodes) return fail(BGPU_E_INVALID, "tex_nodes is NULL");
  if (d->num_images && !d->images) return fail(BGPU_E_INVALID, "images is NULL");
  std::vector<int32_t> lo(nn);
  std::vector<int4> images(d->num_images);
  std::vector<float> texels;
  for (uint32_t k = 0; k < d->num_images; ++k) {
    const bgpu_image& im = d->images[k];
    if (im.width <= 0 || im.height <= 0 || (im.channels != 1 && im.channels != 16) || !im.texels ||
        (uint64_t)im.width * (uint64_t)im.height > (1ull << 26))
      return fail(BGPU_E_INVALID, "image " + std::to_string(k) + ": bad size, channels or texels");
    if (texels.size() + (size_t)im.channels * im.width * im.height > 0x7fffffffull)
      return fail(BGPU_E_INVALID, "images: more than 2^31 texels in all");
    images[k] = int4{im.width, im.height, im.channels, (int32_t)texels.size()};
    texels.insert(texels.end(), im.texels, im.texels + (size_t)im.channels * im.width * im.height);
  }
  auto is_scalar = [](int op) { return op >= BGPU_TN_SCONST && op <= BGPU_TN_SCALE; };
  auto is_spectrum = [](int op) { return op >= BGPU_TN_CONST && op <= BGPU_TN_GRAPHPAPER; };
  for (uint32_t k = 0; k < nn; ++k) {
    const bgpu_tex_node& t = d->tex_nodes[k];
    const int i = (int)k;
    int low = i;
    auto child = [&](int c, bool scalar) {
      if (c < 0 || c >= i) return false;
      if (!(scalar ? is_scalar(d->tex_nodes[c].op) : is_spectrum(d->tex_nodes[c].op))) return false;
      low = std::min(low, lo[c]);
      return true;
    };
    auto map2 = [&]() { return t.map == BGPU_MAP_UV || t.map == BGPU_MAP_PLANAR; };
    auto img = [&](int ch) { return t.a >= 0 && (uint32_t)t.a < d->num_images && d->images[t.a].channels == ch; };
    bool ok = true;
    switch (t.op) {
      case BGPU_TN_SCONST: case BGPU_TN_CONST: case BGPU_TN_GSTEP: break;
      case BGPU_TN_SIMAGE: ok = img(1) && map2(); break;
      case BGPU_TN_IMAGE: ok = img(16) && map2(); break;
      case BGPU_TN_CELL: ok = t.a >= 0 && t.a <= 3 && t.map == BGPU_MAP_IDENTITY; break;
      case BGPU_TN_FBM: ok = t.a >= 0 && t.a <= 64 && t.map == BGPU_MAP_IDENTITY; break;
      case BGPU_TN_PERLIN: ok = t.map == BGPU_MAP_IDENTITY; break;
      case BGPU_TN_CRYSTAL: ok = t.a >= 1 && t.a <= 1024 && map2(); break;
      case BGPU_TN_SCALE: ok = child(t.a, true); break;
      case BGPU_TN_BLEND: ok = child(t.a, false) && child(t.b, false) && child(t.c, true); break;
      case BGPU_TN_CHECKER: ok = child(t.a, false) && child(t.b, false); break;
      case BGPU_TN_GRAPHPAPER: ok = child(t.a, false) && child(t.b, false) && map2(); break;
      case BGPU_TN_GRADIENT:  // steps: GSTEP nodes b .. b + c - 1, ascending by position
        ok = child(t.a, true) && t.c >= 1 && t.b >= 0 && t.b + t.c <= i;
        for (int j = 0; ok && j < t.c; ++j) {
          ok = d->tex_nodes[t.b + j].op == BGPU_TN_GSTEP &&
               (j == 0 || d->tex_nodes[t.b + j - 1].q[0] <= d->tex_nodes[t.b + j].q[0]);
          low = std::min(low, t.b + j);
        }
        break;
      default: ok = false;
    }
    if (!ok) return fail(BGPU_E_INVALID, "texture node " + std::to_string(k) + ": bad operation or operands");
    lo[k] = low;
  }
  std::vector<DMatSrc> srcs;
  for (uint32_t i = 0; i < d->num_materials; ++i) {
    const bgpu_material& m = d->materials[i];
    mats[i] = build(m, m.spec);
    const bgpu_texture& tx = m.tex;
    if (tx.kind == BGPU_TEX_NONE) continue;
    if (tx.kind == BGPU_TEX_NODES) {  // inputs evaluated per hit by the texture kernel
      if (tx.slots <= 0 || (tx.slots & ~0x1f7) != 0)
        return fail(BGPU_E_INVALID, "material " + std::to_string(i) + ": bad textured slots");
      DMatSrc src;
      std::memset(&src, 0, sizeof(src));
      src.type = m.type;
      src.slots = tx.slots;
      std::memcpy(src.scalar, m.scalar, sizeof(src.scalar));
      std::memcpy(src.spec, m.spec, sizeof(src.spec));
      for (int k = 0; k < 9; ++k) {  // spectra 0..2, scalars 4..7, the bump displacement 8
        if (!((tx.slots >> k) & 1)) continue;
        const int r = tx.roots[k];
        if (r < 0 || (uint32_t)r >= nn ||
            !(k < 3 ? is_spectrum(d->tex_nodes[r].op) : is_scalar(d->tex_nodes[r].op)) ||
            r - lo[r] + 1 > tx::kMaxTexRange)
          return fail(BGPU_E_INVALID, "material " + std::to_string(i) + ": bad texture root, or a texture of more "
                                      "than " + std::to_string(tx::kMaxTexRange) + " nodes");
        src.roots[k] = r;
        src.lo[k] = lo[r];
      }
      mats[i].tex_kind = BGPU_TEX_NODES;
      mats[i].alt = (int32_t)srcs.size();
      srcs.push_back(src);
      pt.node_tex = true;
      continue;
    }
    if ((tx.kind != BGPU_TEX_CHECKER && tx.kind != BGPU_TEX_GRAPHPAPER) ||
        (tx.mapping != BGPU_MAP_UV && tx.mapping != BGPU_MAP_PLANAR) || tx.slots < 1 || tx.slots > 3)
      return fail(BGPU_E_INVALID, "material texture: unknown kind, mapping or slots");
    // the tex2 variant: the textured slots take the alternate spectra
    float alt[2][16];
    for (int k = 0; k < 2; ++k) std::memcpy(alt[k], ((tx.slots >> k) & 1) ? tx.alt[k] : m.spec[k], 64);
    mats[i].tex_kind = tx.kind;
    mats[i].tex_map = tx.mapping;
    std::memcpy(mats[i].tex, tx.params, sizeof(mats[i].tex));
    mats[i].alt = (int32_t)mats.size();
    mats.push_back(build(m, alt));
  }
  // eval_bsdf / sample_bsdf take a FresnelBlend lobe (K_FB) to be its
  // material's only component: its value replaces the per-component sum
  for (const DMaterial& dm : mats)
    for (int c = 0; c < dm.n; ++c)
      if (dm.c[c].kind == K_FB && dm.n != 1)
        return fail(BGPU_E_INVALID, "material: a FresnelBlend lobe must be the material's only BxDF");
  if (int rc = sc->materials.upload(ctx, mats.data(), mats.size() * sizeof(DMaterial), st)) return rc;
  // distinct BxDF makeups among the materials primitives use: a scene with
  // more than one benefits from shading queues grouped by material kind
  {
    std::vector<char> used(mats.size(), 0);
    for (uint32_t t = 0; t < d->num_triangles; ++t) {
      const int32_t mi = d->tri_mesh ? d->tri_mesh[t] : 0;
      const int32_t m = (d->meshes && mi >= 0 && (uint32_t)mi < std::max<uint32_t>(d->num_meshes, 1))
                            ? d->meshes[mi].material
                            : 0;
      if (m >= 0 && (size_t)m < used.size()) used[m] = 1;
    }
    for (uint32_t s = 0; s < ns; ++s)
      if (d->shapes[s].material >= 0 && (size_t)d->shapes[s].material < used.size()) used[d->shapes[s].material] = 1;
    for (size_t i = 0; i < d->num_materials; ++i)  // a texture may select the alternate entry
      if (used[i] && (mats[i].tex_kind == BGPU_TEX_CHECKER || mats[i].tex_kind == BGPU_TEX_GRAPHPAPER))
        used[mats[i].alt] = 1;
    std::vector<int> sigs;
    bool diffuse = true;
    for (size_t i = 0; i < mats.size(); ++i) {
      if (!used[i]) continue;
      const DMaterial& m = mats[i];
      const int sig = m.n | (m.n > 0 ? m.c[0].kind << 2 : 0) | (m.n > 1 ? m.c[1].kind << 5 : 0);
      if (std::find(sigs.begin(), sigs.end(), sig) == sigs.end()) sigs.push_back(sig);
      for (int c = 0; c < m.n; ++c) diffuse = diffuse && (m.c[c].kind == K_LAMBERT || m.c[c].kind == K_OREN);
    }
    sc->pt.shade_kinds = (int)sigs.size();
    sc->pt.diffuse_only = diffuse;
  }

  // lights: explicit first, then emissive shapes (Scene.hs:38-42)
  std::vector<DLight> lights;
  std::vector<float> maps, dist;
  for (uint32_t i = 0; i < d->num_lights; ++i) {
    const bgpu_light& l = d->lights[i];
    if (l.type == BGPU_LIGHT_POINT || l.type == BGPU_LIGHT_DIRECTIONAL) {  // mkPointLight, mkDirectional
      if (!l.radiance) return fail(BGPU_E_INVALID, "point or directional light without an intensity");
      DLight o;
      std::memset(&o, 0, sizeof(o));
      o.type = l.type == BGPU_LIGHT_POINT ? LT_POINT : LT_DIR;
      o.shape = -1;
      o.id = -1;
      std::memcpy(o.L, l.radiance, 64);
      float x = l.vec[0], y = l.vec[1], z = l.vec[2];
      if (o.type == LT_DIR) {  // mkDirectional s n = Directional s (normalize n) (Light.hs:53-54)
        const float sq = (x * x + y * y) + z * z;
        if (sq != 0.0f) {
          const float il = 1.0f / std::sqrt(sq);
          x *= il;
          y *= il;
          z *= il;
        } else {
          x = 0.0f, y = 1.0f, z = 0.0f;
        }
      }
      o.vec[0] = x;
      o.vec[1] = y;
      o.vec[2] = z;
      lights.push_back(o);
      continue;
    }
    const bool sky = l.type == BGPU_LIGHT_INFINITE && l.map == BGPU_LIGHTMAP_SUNSKY;
    if (l.type != BGPU_LIGHT_INFINITE || (l.map != BGPU_LIGHTMAP_TEXELS && !sky) ||
        (!sky && (l.width <= 0 || l.height <= 0 || !l.radiance)))
      return fail(BGPU_E_INVALID, "unsupported or empty explicit light");
    if (sky && !(l.sky[6] == l.sky[6]))
      return fail(BGPU_E_INVALID, "sun/sky map: turbidity is not a number");
    DLight o;
    std::memset(&o, 0, sizeof(o));
    o.type = LT_INF;
    o.shape = -1;
    o.id = -1;
    o.sky = sky ? 1 : 0;
    o.w = sky ? kSkyMapW : l.width;
    o.h = sky ? kSkyMapH : l.height;
    o.map_off = (int32_t)maps.size();
    if (sky) {
      maps.resize(maps.size() + kSkyFloats);
      sky_init(l.sky, l.sky + 3, l.sky[6], maps.data() + o.map_off);
    } else {
      maps.insert(maps.end(), l.radiance, l.radiance + 16ull * l.width * l.height);
    }
    std::memcpy(o.w2l, l.w2l, 64);
    std::memcpy(o.l2w, l.l2w, 64);
    // mkDist2D (texSize rmap) (sY . eval), eval (x, y) at Cartesian (x/sx, y/sy)
    o.dist_off = (int32_t)dist.size();
    int nu = o.w, nvv = o.h;
    float b;
    std::memcpy(&b, &nu, 4);
    dist.push_back(b);
    std::memcpy(&b, &nvv, 4);
    dist.push_back(b);
    float fsx = (float)nu, fsy = (float)nvv;
    std::vector<std::vector<float>> rows(nvv);
    std::vector<float> fis(nvv);
    const float* blk = maps.data() + o.map_off;
    for (int y = 0; y < nvv; ++y) {
      rows[y].resize(nu);
      for (int x = 0; x < nu; ++x) {
        const float u = (float)x / fsx, v = (float)y / fsy;
        if (sky) {
          const SkyColor c = sky_color(blk + SKY_PAR, u, v);
          float s16[16];
          for (int k = 0; k < 16; ++k) s16[k] = sky_band(blk, c, k);
          rows[y][x] = sY(s16);
        } else {
          rows[y][x] = sY(tex_px(l, u, v));
        }
      }
      fis[y] = dist1d_fi(rows[y]);
    }
    push_dist1d(dist, fis);
    for (int y = 0; y < nvv; ++y) push_dist1d(dist, rows[y]);
    lights.push_back(o);
  }
  pt.shape_light_host.assign(ns, -1);
  for (uint32_t s = 0; s < ns; ++s) {
    if (!d->shapes[s].emissive) continue;
    DLight o;
    std::memset(&o, 0, sizeof(o));
    o.type = LT_AREA;
    o.shape = (int32_t)s;
    o.id = d->shapes[s].light_id;
    std::memcpy(o.L, d->shapes[s].emission, 64);
    o.sh.type = d->shapes[s].type;  // the shape's DShape record, as the scene's shape table holds it
    o.sh.material = d->shapes[s].material;
    o.sh.light = (int32_t)lights.size();
    std::memcpy(o.sh.p, d->shapes[s].params, sizeof(o.sh.p));
    std::memcpy(o.sh.w2o, d->shapes[s].w2o, sizeof(o.sh.w2o));
    std::memcpy(o.l2w, d->shapes[s].o2w, sizeof(o.l2w));
    pt.shape_light_host[s] = (int32_t)lights.size();
    lights.push_back(o);
  }
  pt.light_kinds = 0;
  for (const DLight& l : lights)
    pt.light_kinds |= l.type == LT_INF     ? (1 << kLightInf) | (l.sky ? 1 << kLightSky : 0)
                      : l.type == LT_POINT ? (1 << kLightPoint)
                      : l.type == LT_DIR   ? (1 << kLightDir)
                                           : (1 << d->shapes[l.shape].type);
  if (lights.empty()) lights.push_back(DLight{});
  if (maps.empty()) maps.push_back(0.0f);
  if (dist.empty()) dist.push_back(0.0f);
  if (int rc = sc->lights.upload(ctx, lights.data(), lights.size() * sizeof(DLight), st)) return rc;
  if (int rc = sc->light_maps.upload(ctx, maps.data(), maps.size() * 4, st)) return rc;
  if (int rc = sc->light_dist.upload(ctx, dist.data(), dist.size() * 4, st)) return rc;

  H.tri_idx = sc->tri_idx.as<int32_t>();
  H.verts = sc->verts.as<float>();
  H.normals = d->normals ? sc->normals.as<float>() : nullptr;
  H.uvs = d->uvs ? sc->uvs.as<float>() : nullptr;
  H.tri_info = sc->tri_mat.as<int2>();
  H.shapex = sc->shape_x.as<DShapeX>();
  H.mats = sc->materials.as<DMaterial>();
  H.lights = sc->lights.as<DLight>();
  H.maps = sc->light_maps.as<float>();
  H.dist = sc->light_dist.as<float>();
  pt.tex_lo_host = lo;
  pt.tex_op_host.resize(nn);
  for (uint32_t k = 0; k < nn; ++k) pt.tex_op_host[k] = d->tex_nodes[k].op;
  if (nn) {
    if (int rc = sc->tex_nodes.upload(ctx, d->tex_nodes, sizeof(bgpu_tex_node) * nn, st)) return rc;
    if (int rc = sc->tex_lo.upload(ctx, lo.data(), 4 * lo.size(), st)) return rc;
    if (!images.empty())
      if (int rc = sc->images.upload(ctx, images.data(), sizeof(int4) * images.size(), st)) return rc;
    if (!texels.empty())
      if (int rc = sc->texels.upload(ctx, texels.data(), 4 * texels.size(), st)) return rc;
    if (!srcs.empty())
      if (int rc = sc->mat_src.upload(ctx, srcs.data(), sizeof(DMatSrc) * srcs.size(), st)) return rc;
    H.tex_nodes = sc->tex_nodes.as<bgpu_tex_node>();
    H.tex_lo = sc->tex_lo.as<int32_t>();
    H.images = images.empty() ? nullptr : sc->images.as<int4>();
    H.texels = texels.empty() ? nullptr : sc->texels.as<float>();
    H.mat_src = srcs.empty() ? nullptr : sc->mat_src.as<DMatSrc>();
  }
  H.ntri = (int32_t)nt;
  H.nlights = (int32_t)(d->num_lights + std::count_if(pt.shape_light_host.begin(),
                                                      pt.shape_light_host.end(),
                                                      [](int32_t v) { return v >= 0; }));
  H.nmats = (int32_t)d->num_materials;
  return BGPU_OK;
}

int make_view(const bgpu_camera& cam, const bgpu_film& f, DevView* v) {
  if (cam.type != BGPU_CAMERA_PERSPECTIVE && cam.type != BGPU_CAMERA_ENVIRONMENT)
    return fail(BGPU_E_INVALID, "unsupported camera type");
  if (cam.type == BGPU_CAMERA_ENVIRONMENT && !(cam.xres > 0.0f && cam.yres > 0.0f))
    return fail(BGPU_E_INVALID, "environment camera: raster size must be positive");
  if (f.width <= 0 || f.height <= 0 || f.width > 65536 || f.height > 65536)
    return fail(BGPU_E_INVALID, "film size must be in [1, 65536]");
  if (f.filter < BGPU_FILTER_BOX || f.filter > BGPU_FILTER_TRIANGLE)
    return fail(BGPU_E_INVALID, "unknown filter type");
  std::memset(v, 0, sizeof(*v));
  v->cam_type = cam.type;
  std::memcpy(v->c2w, cam.c2w, 64);
  std::memcpy(v->r2c, cam.r2c, 64);
  v->lens_radius = cam.lens_radius;
  v->focal_distance = cam.focal_distance;
  v->env_sx = cam.xres;
  v->env_sy = cam.yres;
  v->film_w = f.width;
  v->film_h = f.height;
  if (f.filter == BGPU_FILTER_BOX) {
    v->fw = 0.5f;
    v->fh = 0.5f;
  } else {
    v->fw = f.fparams[0];
    v->fh = f.fparams[1];
  }
  if (!(v->fw > 0.0f && v->fw <= 64.0f && v->fh > 0.0f && v->fh <= 64.0f))
    return fail(BGPU_E_INVALID, "filter extent must be in (0, 64]");
  for (int y = 0; y < 16; ++y) {
    float fy = ((float)y + 0.5f) * v->fh / 16.0f;
    for (int x = 0; x < 16; ++x) {
      float fx = ((float)x + 0.5f) * v->fw / 16.0f;
      v->filter[y * 16 + x] = eval_filter(f, fx, fy);
    }
  }
  int x0 = (int)std::floor(0.5f - v->fw), x1 = (int)std::floor(0.5f + (float)f.width + v->fw);
  int y0 = (int)std::floor(0.5f - v->fh), y1 = (int)std::floor(0.5f + (float)f.height + v->fh);
  v->ext_x0 = x0;
  v->ext_y0 = y0;
  v->ext_w = x1 - x0 + 1;
  v->ext_h = y1 - y0 + 1;
  if ((long long)v->ext_w * v->ext_h > (1ll << 30)) return fail(BGPU_E_INVALID, "film too large");
  return BGPU_OK;
}

}  // namespace bgpu
