// Host preparation of the path tracer's per-scene data:
//   materials  -> BxDF records (Material.hs:32-96, Microfacet.hs:127-134 fixExponent)
//   lights     -> explicit lights, then geometric (emissive shape) lights in
//                 shape order, the list of mkScene (Scene.hs:37-43)
//   env maps   -> the Dist2D of mkInfiniteAreaLight (Light.hs:76-77,
//                 Montecarlo.hs:34-104), built in binary32 in the same order
//   film       -> 16x16 filter table (Image.hs:48-61) and sample extent (Image.hs:162-168)
#include <algorithm>
#include <cmath>
#include <cstring>

#include "bgpu_internal.h"
#include "spectral_tables.h"

namespace bgpu {

namespace {

inline float hmaxf(float x, float y) { return (x <= y) ? y : x; }
inline float hminf(float x, float y) { return (x <= y) ? x : y; }
inline float clampf(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }

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

const float* tex_px(const bgpu_light& l, float u, float v) {  // Texture.hs:94-99
  int px = (int)std::floor(u * (float)l.width) % l.width;
  px = px < 0 ? px + l.width : px;
  int py = (int)std::floor((-v) * (float)l.height) % l.height;
  py = py < 0 ? py + l.height : py;
  return l.radiance + 16 * ((size_t)py * l.width + px);
}

DBxdf bx_zero() {
  DBxdf b;
  std::memset(&b, 0, sizeof(b));
  return b;
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
  auto blinn_e = [](float rough) {
    float e = 1.0f / rough;
    return (e > 10000.0f || std::isnan(e)) ? 10000.0f : e;
  };
  // one material from its type and scalars, with spectra spec[0], spec[1]
  auto build = [&](const bgpu_material& m, const float (*spec)[16]) {
    DMaterial o;
    std::memset(&o, 0, sizeof(o));
    switch (m.type) {
      case BGPU_MAT_MATTE: {
        DBxdf b = bx_zero();
        std::memcpy(b.r, spec[0], 64);
        float sig = m.scalar[0];
        if (sig == 0.0f) {
          b.kind = K_LAMBERT;
        } else {
          b.kind = K_OREN;
          float s = clampf(sig, 0.0f, 1.0f);
          float s2 = s * s;
          b.A = 1.0f - (s2 / (2.0f * (s2 + 0.33f)));
          b.B = 0.45f * s2 / (s2 + 0.09f);
        }
        b.type = BX_R | BX_D;
        o.c[o.n++] = b;
        break;
      }
      case BGPU_MAT_PLASTIC: {
        DBxdf dd = bx_zero();
        dd.kind = K_LAMBERT;
        dd.type = BX_R | BX_D;
        std::memcpy(dd.r, spec[0], 64);
        DBxdf g = bx_zero();
        g.kind = K_MF;
        g.type = BX_R | BX_G;
        std::memcpy(g.r, spec[1], 64);
        g.e = blinn_e(m.scalar[0]);
        g.fr = FR_DIEL;
        g.ei = 1.0f;
        g.et = 1.5f;
        o.c[o.n++] = dd;
        o.c[o.n++] = g;
        break;
      }
      case BGPU_MAT_METAL: {
        DBxdf g = bx_zero();
        g.kind = K_MF;
        g.type = BX_R | BX_G;
        for (int k = 0; k < 16; ++k) g.r[k] = 1.0f;
        g.e = blinn_e(m.scalar[0]);
        g.fr = FR_COND;
        std::memcpy(g.eta, spec[0], 64);
        std::memcpy(g.k, spec[1], 64);
        o.c[o.n++] = g;
        break;
      }
      case BGPU_MAT_GLASS: {
        DBxdf r = bx_zero();
        r.kind = K_SPECR;
        r.type = BX_R | BX_S;
        r.fr = FR_DIEL;
        r.ei = 1.0f;
        r.et = m.scalar[0];
        DBxdf t = bx_zero();
        t.kind = K_SPECT;
        t.type = BX_T | BX_S;
        t.ei = 1.0f;
        t.et = m.scalar[0];
        for (int k = 0; k < 16; ++k) {  // sClamp' (Spectrum.hs:459-461)
          r.r[k] = hmaxf(0.0f, hminf(1.0f, spec[0][k]));
          t.r[k] = hmaxf(0.0f, hminf(1.0f, spec[1][k]));
        }
        o.c[o.n++] = r;
        o.c[o.n++] = t;
        break;
      }
      case BGPU_MAT_TRANSMATTE: {  // translucentMatte (Material.hs:43-53)
        const float sig = m.scalar[0];
        auto lobe = [&](const float* s, bool flip) {
          DBxdf b = bx_zero();
          std::memcpy(b.r, s, 64);
          if (sig == 0.0f) {
            b.kind = K_LAMBERT;
          } else {
            b.kind = K_OREN;
            float c = clampf(sig, 0.0f, 1.0f);
            float c2 = c * c;
            b.A = 1.0f - (c2 / (2.0f * (c2 + 0.33f)));
            b.B = 0.45f * c2 / (c2 + 0.09f);
          }
          b.type = flip ? (BX_T | BX_D) : (BX_R | BX_D);  // bxdfTypeFlip [Reflection, Transmission]
          b.flip = flip ? 1 : 0;
          return b;
        };
        float r[16], t[16];
        for (int k = 0; k < 16; ++k) {
          r[k] = hmaxf(0.0f, hminf(1.0f, spec[0][k]));
          t[k] = hmaxf(0.0f, hminf(1.0f, spec[1][k])) * (1.0f - r[k]);
        }
        o.c[o.n++] = lobe(r, false);
        o.c[o.n++] = lobe(t, true);
        break;
      }
      case BGPU_MAT_SUBSTRATE: {  // mkSubstrate (Material.hs:111-127)
        DBxdf b = bx_zero();
        b.kind = K_FB;
        b.type = BX_R | BX_G;
        for (int k = 0; k < 16; ++k) {  // sClamp 0 1
          b.r[k] = hmaxf(0.0f, hminf(1.0f, spec[0][k]));
          b.eta[k] = hmaxf(0.0f, hminf(1.0f, spec[1][k]));
          b.k[k] = hmaxf(0.0f, hminf(1.0f, m.spec[2][k]));
        }
        b.e = blinn_e(hmaxf(0.0f, m.scalar[0]));  // mkAnisotropic (1 / u) (1 / v), u = max 0 urough
        b.ei = blinn_e(hmaxf(0.0f, m.scalar[1]));
        b.et = m.scalar[2];
        o.c[o.n++] = b;
        break;
      }
      case BGPU_MAT_MIRROR: {
        DBxdf r = bx_zero();
        r.kind = K_SPECR;
        r.type = BX_R | BX_S;
        r.fr = FR_NOOP;
        for (int k = 0; k < 16; ++k) r.r[k] = hmaxf(0.0f, hminf(1.0f, spec[0][k]));
        o.c[o.n++] = r;
        break;
      }
      default:
        break;
    }
    return o;
  };
  for (uint32_t i = 0; i < d->num_materials; ++i) {
    const bgpu_material& m = d->materials[i];
    mats[i] = build(m, m.spec);
    const bgpu_texture& tx = m.tex;
    if (tx.kind == BGPU_TEX_NONE) continue;
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
      if (used[i] && mats[i].tex_kind != BGPU_TEX_NONE) used[mats[i].alt] = 1;
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
    if (l.type != BGPU_LIGHT_INFINITE || l.width <= 0 || l.height <= 0 || !l.radiance)
      return fail(BGPU_E_INVALID, "unsupported or empty explicit light");
    DLight o;
    std::memset(&o, 0, sizeof(o));
    o.type = LT_INF;
    o.shape = -1;
    o.id = -1;
    o.w = l.width;
    o.h = l.height;
    o.map_off = (int32_t)maps.size();
    maps.insert(maps.end(), l.radiance, l.radiance + 16ull * l.width * l.height);
    std::memcpy(o.w2l, l.w2l, 64);
    std::memcpy(o.l2w, l.l2w, 64);
    // mkDist2D (texSize rmap) (sY . eval), eval (x, y) at Cartesian (x/sx, y/sy)
    o.dist_off = (int32_t)dist.size();
    int nu = l.width, nvv = l.height;
    float b;
    std::memcpy(&b, &nu, 4);
    dist.push_back(b);
    std::memcpy(&b, &nvv, 4);
    dist.push_back(b);
    float fsx = (float)nu, fsy = (float)nvv;
    std::vector<std::vector<float>> rows(nvv);
    std::vector<float> fis(nvv);
    for (int y = 0; y < nvv; ++y) {
      rows[y].resize(nu);
      for (int x = 0; x < nu; ++x) rows[y][x] = sY(tex_px(l, (float)x / fsx, (float)y / fsy));
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
    pt.shape_light_host[s] = (int32_t)lights.size();
    lights.push_back(o);
  }
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
  H.ntri = (int32_t)nt;
  H.nlights = (int32_t)(d->num_lights + std::count_if(pt.shape_light_host.begin(),
                                                      pt.shape_light_host.end(),
                                                      [](int32_t v) { return v >= 0; }));
  H.nmats = (int32_t)d->num_materials;
  return BGPU_OK;
}

int make_view(const bgpu_camera& cam, const bgpu_film& f, DevView* v) {
  if (cam.type != BGPU_CAMERA_PERSPECTIVE) return fail(BGPU_E_INVALID, "unsupported camera type");
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
