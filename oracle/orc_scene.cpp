// ORACLE -- test infrastructure only (see orc_math.h).
#include "orc_scene.h"

#include <algorithm>
#include <cstring>
#include <numeric>

namespace orc {

static M4 toM4(const float* m) {
  M4 r;
  std::memcpy(r.m, m, sizeof(r.m));
  return r;
}

AABB Scene::primBounds(int id) const {
  if (id < ntri) {
    V3 a, b, c;
    tri(id, &a, &b, &c);
    return triangleBounds(a, b, c);
  }
  const ShapePrim& sp = shapes[id - ntri];
  return transBox(sp.o2w, objectBounds(sp.shape));  // Shape.hs:287-292
}

bool primIntersect(const Scene& s, int id, const Ray& r, Hit* h) {
  if (id < s.ntri) {
    V3 a, b, c;
    s.tri(id, &a, &b, &c);
    TriHit th;
    if (!triangleIntersect(a, b, c, r, &th)) return false;
    h->t = th.t;
    h->b1 = th.b1;
    h->b2 = th.b2;
    h->prim = id;
    h->eps = 1e-3f * th.t;  // TriangleMesh.hs:206
    h->axis = 0;
    return true;
  }
  const ShapePrim& sp = s.shapes[id - s.ntri];
  Ray ro = transRay(sp.w2o, r);  // Geometry.hs:33-36
  ShapeHit sh;
  if (!shapeIntersect(sp.shape, ro, &sh)) return false;
  h->t = sh.t;
  h->b1 = 0.0f;
  h->b2 = 0.0f;
  h->prim = id;
  h->eps = sh.eps;
  h->axis = sh.axis;
  return true;
}

bool primIntersects(const Scene& s, int id, const Ray& r) {
  if (id < s.ntri) {
    V3 a, b, c;
    s.tri(id, &a, &b, &c);
    return triangleIntersects(a, b, c, r);
  }
  const ShapePrim& sp = s.shapes[id - s.ntri];
  return shapeIntersects(sp.shape, transRay(sp.w2o, r));  // Geometry.hs:26-27
}

// (t, id) lexicographic minimum; counts exact ties at the running minimum.
static inline void consider(const Hit& c, Hit* best, int* ties) {
  if (best->prim < 0 || c.t < best->t) {
    *best = c;
    *ties = 0;
  } else if (c.t == best->t) {
    ++*ties;
    if (c.prim < best->prim) *best = c;
  }
}

// Pruning test in double precision.  Boxes holding analytic shapes are grown
// by 3e-3 * (distance from the ray origin + box diagonal): the binary32
// quadratic of Shape.hs:173-187 evaluates c = |o|^2 - r^2 with cancellation,
// so it reports hits up to ~3*sqrt(2^-24)*|o| outside the true surface.
static bool nodeOverlap(const Scene::Node& n, const Ray& r, double tlo, double thi) {
  double o[3] = {r.o.x, r.o.y, r.o.z}, d[3] = {r.d.x, r.d.y, r.d.z};
  double lo[3] = {n.lo[0], n.lo[1], n.lo[2]}, hi[3] = {n.hi[0], n.hi[1], n.hi[2]};
  if (n.has_shape) {
    double dist = 0.0, diag = 0.0;
    for (int k = 0; k < 3; ++k) {
      dist += std::max(0.0, std::max(lo[k] - o[k], o[k] - hi[k]));
      diag += hi[k] - lo[k];
    }
    double pad = 3e-3 * (dist + diag);
    for (int k = 0; k < 3; ++k) {
      lo[k] -= pad;
      hi[k] += pad;
    }
  }
  double a = tlo, b = thi;
  for (int k = 0; k < 3; ++k) {
    if (d[k] == 0.0) {
      if (o[k] < lo[k] || o[k] > hi[k]) return false;
      continue;
    }
    double t0 = (lo[k] - o[k]) / d[k], t1 = (hi[k] - o[k]) / d[k];
    if (t0 > t1) std::swap(t0, t1);
    a = std::max(a, t0);
    b = std::min(b, t1);
    if (a > b) return false;
  }
  return true;
}

static double tMargin(const Scene& s, const Ray& r) {
  double diam = 0.0;
  for (int k = 0; k < 3; ++k) {
    double e = (double)comp(s.bounds.hi, k) - (double)comp(s.bounds.lo, k);
    if (e > diam) diam = e;
  }
  double dl = std::sqrt((double)r.d.x * r.d.x + (double)r.d.y * r.d.y + (double)r.d.z * r.d.z);
  if (!(dl > 0.0)) return 1e30;
  return 1e-3 * (diam + 1.0) / dl;
}

bool sceneIntersect(const Scene& s, const Ray& r, Hit* h, bool brute, int* tiesOut) {
  h->prim = -1;
  int ties = 0;
  if (!intersectAABB(s.bounds, r, nullptr, nullptr)) {  // KdTree.hs:240
    if (tiesOut) *tiesOut = 0;
    return false;
  }
  if (brute || s.nodes.empty()) {
    for (int id = 0; id < s.nprims(); ++id) {
      Hit c;
      if (primIntersect(s, id, r, &c)) consider(c, h, &ties);
    }
  } else {
    double m = tMargin(s, r);
    int stack[256];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
      const Scene::Node& n = s.nodes[stack[--sp]];
      double thi = (h->prim >= 0 ? (double)h->t : (double)r.tmax) + m;
      if (!nodeOverlap(n, r, (double)r.tmin - m, thi)) continue;
      if (n.right < 0) {
        for (int i = 0; i < n.count; ++i) {
          Hit c;
          if (primIntersect(s, s.order[n.left + i], r, &c)) consider(c, h, &ties);
        }
      } else {
        stack[sp++] = n.left;
        stack[sp++] = n.right;
      }
    }
  }
  if (tiesOut) *tiesOut = ties;
  return h->prim >= 0;
}

bool sceneOccluded(const Scene& s, const Ray& r, bool brute) {
  if (!intersectAABB(s.bounds, r, nullptr, nullptr)) return false;  // KdTree.hs:244
  if (brute || s.nodes.empty()) {
    for (int id = 0; id < s.nprims(); ++id)
      if (primIntersects(s, id, r)) return true;
    return false;
  }
  double m = tMargin(s, r);
  int stack[256];
  int sp = 0;
  stack[sp++] = 0;
  while (sp) {
    const Scene::Node& n = s.nodes[stack[--sp]];
    if (!nodeOverlap(n, r, (double)r.tmin - m, (double)r.tmax + m)) continue;
    if (n.right < 0) {
      for (int i = 0; i < n.count; ++i)
        if (primIntersects(s, s.order[n.left + i], r)) return true;
    } else {
      stack[sp++] = n.left;
      stack[sp++] = n.right;
    }
  }
  return false;
}

// Median-split BVH over centroids, boxes in double and padded.
static void buildNodes(Scene& s) {
  int n = s.nprims();
  s.nodes.clear();
  s.order.resize(n);
  std::iota(s.order.begin(), s.order.end(), 0);
  if (n == 0) return;
  std::vector<double> blo(3 * (size_t)n), bhi(3 * (size_t)n), cen(3 * (size_t)n);
  double diam = 0.0;
  for (int k = 0; k < 3; ++k)
    diam = std::max(diam, (double)comp(s.bounds.hi, k) - (double)comp(s.bounds.lo, k));
  if (!(diam < 1e30)) diam = 1e30;
  double pad = 1e-4 * (diam + 1.0);
  for (int i = 0; i < n; ++i) {
    AABB b = s.primBounds(i);
    for (int k = 0; k < 3; ++k) {
      blo[3 * i + k] = (double)comp(b.lo, k) - pad;
      bhi[3 * i + k] = (double)comp(b.hi, k) + pad;
      cen[3 * i + k] = 0.5 * (blo[3 * i + k] + bhi[3 * i + k]);
    }
  }
  struct Task {
    int node, begin, end;
  };
  std::vector<Task> work;
  s.nodes.push_back(Scene::Node{});
  work.push_back(Task{0, 0, n});
  while (!work.empty()) {
    Task t = work.back();
    work.pop_back();
    Scene::Node nd;
    for (int k = 0; k < 3; ++k) {
      nd.lo[k] = 1e300;
      nd.hi[k] = -1e300;
    }
    double clo[3] = {1e300, 1e300, 1e300}, chi[3] = {-1e300, -1e300, -1e300};
    nd.has_shape = 0;
    for (int i = t.begin; i < t.end; ++i) {
      int p = s.order[i];
      if (p >= s.ntri) nd.has_shape = 1;
      for (int k = 0; k < 3; ++k) {
        nd.lo[k] = std::min(nd.lo[k], blo[3 * p + k]);
        nd.hi[k] = std::max(nd.hi[k], bhi[3 * p + k]);
        clo[k] = std::min(clo[k], cen[3 * p + k]);
        chi[k] = std::max(chi[k], cen[3 * p + k]);
      }
    }
    int cnt = t.end - t.begin;
    if (cnt <= 4) {
      nd.left = t.begin;
      nd.right = -1;
      nd.count = cnt;
      s.nodes[t.node] = nd;
      continue;
    }
    int ax = 0;
    for (int k = 1; k < 3; ++k)
      if (chi[k] - clo[k] > chi[ax] - clo[ax]) ax = k;
    int mid = (t.begin + t.end) / 2;
    std::nth_element(s.order.begin() + t.begin, s.order.begin() + mid, s.order.begin() + t.end,
                     [&](int a, int b) { return cen[3 * a + ax] < cen[3 * b + ax]; });
    int l = (int)s.nodes.size();
    s.nodes.push_back(Scene::Node{});
    int r = (int)s.nodes.size();
    s.nodes.push_back(Scene::Node{});
    nd.left = l;
    nd.right = r;
    nd.count = 0;
    s.nodes[t.node] = nd;
    work.push_back(Task{l, t.begin, mid});
    work.push_back(Task{r, mid, t.end});
  }
}

Scene* buildScene(const bgpu_scene_desc* d, const char** err) {
  *err = nullptr;
  Scene* s = new Scene();
  s->ntri = (int)d->num_triangles;
  s->verts.resize(d->num_vertices);
  for (uint32_t i = 0; i < d->num_vertices; ++i)
    s->verts[i] = mk(d->vertices[3 * i], d->vertices[3 * i + 1], d->vertices[3 * i + 2]);
  if (d->normals) {
    s->normals.resize(d->num_vertices);
    for (uint32_t i = 0; i < d->num_vertices; ++i)
      s->normals[i] = mk(d->normals[3 * i], d->normals[3 * i + 1], d->normals[3 * i + 2]);
  }
  if (d->uvs) s->uvs.assign(d->uvs, d->uvs + 2 * (size_t)d->num_vertices);
  s->idx.assign(d->indices, d->indices + 3 * (size_t)d->num_triangles);
  for (int v : s->idx) {  // mkTriangleMesh checks (TriangleMesh.hs:46-55)
    if (v < 0) { *err = "mkTriangleMesh: contains negative indices"; delete s; return nullptr; }
    if ((uint32_t)v >= d->num_vertices) {
      *err = "mkTriangleMesh: contains out of bounds indices";
      delete s;
      return nullptr;
    }
  }
  s->tri_mesh.assign(d->num_triangles, 0);
  if (d->tri_mesh) s->tri_mesh.assign(d->tri_mesh, d->tri_mesh + d->num_triangles);
  s->meshes.assign(d->meshes, d->meshes + d->num_meshes);
  if (s->meshes.empty()) s->meshes.push_back(bgpu_mesh{0, 0});
  for (int m : s->tri_mesh)
    if (m < 0 || m >= (int)s->meshes.size()) { *err = "bad tri_mesh index"; delete s; return nullptr; }
  for (uint32_t i = 0; i < d->num_shapes; ++i) {
    const bgpu_shape& in = d->shapes[i];
    ShapePrim sp;
    sp.shape.type = in.type;
    std::memcpy(sp.shape.p, in.params, sizeof(sp.shape.p));
    sp.o2w = toM4(in.o2w);
    sp.w2o = toM4(in.w2o);
    sp.material = in.material;
    sp.emissive = in.emissive;
    sp.light_id = in.light_id;
    std::memcpy(sp.emission, in.emission, sizeof(sp.emission));
    if (in.type < 0 || in.type > 4) { *err = "bad shape type"; delete s; return nullptr; }
    s->shapes.push_back(sp);
  }
  s->materials.assign(d->materials, d->materials + d->num_materials);
  for (uint32_t i = 0; i < d->num_lights; ++i) {
    bgpu_light l = d->lights[i];
    std::vector<float> map;
    if (l.radiance) map.assign(l.radiance, l.radiance + 16 * (size_t)l.width * l.height);
    s->light_maps.push_back(map);
    s->lights.push_back(l);
  }
  s->camera = d->camera;
  s->film = d->film;
  // texture nodes: every child precedes its node, references are in range
  if (d->num_tex_nodes && !d->tex_nodes) { *err = "tex_nodes is NULL"; delete s; return nullptr; }
  s->textures.nodes.assign(d->tex_nodes, d->tex_nodes + d->num_tex_nodes);
  for (uint32_t i = 0; i < d->num_images; ++i) {
    const bgpu_image& im = d->images[i];
    if (im.width <= 0 || im.height <= 0 || (im.channels != 1 && im.channels != 16) || !im.texels) {
      *err = "bad image"; delete s; return nullptr;
    }
    tex::Textures::Image o;
    o.w = im.width;
    o.h = im.height;
    o.ch = im.channels;
    o.texels.assign(im.texels, im.texels + (size_t)im.channels * im.width * im.height);
    s->textures.images.push_back(std::move(o));
  }
  if (const char* e = tex::validate(s->textures)) { *err = e; delete s; return nullptr; }
  for (const bgpu_material& m : s->materials)
    if (m.tex.kind == BGPU_TEX_NODES)
      for (int k = 0; k < 9; ++k) {  // spectra 0..2, scalars 4..7, the bump displacement 8
        if (!((m.tex.slots >> k) & 1)) continue;
        const int r = m.tex.roots[k];
        const bool ok = k != 3 && r >= 0 && r < (int)d->num_tex_nodes &&
                        (k < 3 ? tex::isSpectrumOp(d->tex_nodes[r].op) : tex::isScalarOp(d->tex_nodes[r].op));
        if (!ok || (m.tex.slots >> 9) != 0) { *err = "bad texture root"; delete s; return nullptr; }
      }
  // mkKdTree: bounds = foldl' mappend mempty (map worldBounds ps)  (KdTree.hs:111)
  s->bounds = emptyAABB();
  for (int id = 0; id < s->nprims(); ++id) s->bounds = extend(s->bounds, s->primBounds(id));
  buildNodes(*s);
  return s;
}

}  // namespace orc

// ---- C exports for the tests (ctypes) ------------------------------------
extern "C" {

void* orc_scene_create(const bgpu_scene_desc* d, char* errbuf, int errlen) {
  const char* err = nullptr;
  orc::Scene* s = orc::buildScene(d, &err);
  if (!s && errbuf && errlen > 0) {
    std::strncpy(errbuf, err ? err : "error", errlen - 1);
    errbuf[errlen - 1] = 0;
  }
  return s;
}

void orc_scene_destroy(void* s) { delete (orc::Scene*)s; }

void orc_scene_bounds(void* sp, float* out6) {
  orc::Scene* s = (orc::Scene*)sp;
  out6[0] = s->bounds.lo.x; out6[1] = s->bounds.lo.y; out6[2] = s->bounds.lo.z;
  out6[3] = s->bounds.hi.x; out6[4] = s->bounds.hi.y; out6[5] = s->bounds.hi.z;
}

static orc::Ray getRay(const bgpu_rays* r, size_t i) {
  return orc::Ray{orc::mk(r->ox[i], r->oy[i], r->oz[i]), orc::mk(r->dx[i], r->dy[i], r->dz[i]),
                  r->tmin[i], r->tmax[i]};
}

int orc_intersect_closest(void* sp, const bgpu_rays* rays, size_t n, bgpu_hits* out, int brute,
                          int nthreads, int64_t* nties) {
  orc::Scene* s = (orc::Scene*)sp;
  int64_t ties = 0;
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads) reduction(+ : ties)
  for (long i = 0; i < (long)n; ++i) {
    orc::Hit h;
    int t = 0;
    orc::Ray r = getRay(rays, (size_t)i);
    if (orc::sceneIntersect(*s, r, &h, brute != 0, &t)) {
      out->t[i] = h.t;
      out->b1[i] = h.b1;
      out->b2[i] = h.b2;
      out->prim[i] = h.prim;
      if (t) ties++;
    } else {
      out->t[i] = 0.0f;
      out->b1[i] = 0.0f;
      out->b2[i] = 0.0f;
      out->prim[i] = -1;
    }
  }
  if (nties) *nties = ties;
  return 0;
}

int orc_intersect_any(void* sp, const bgpu_rays* rays, size_t n, uint8_t* occ, int brute,
                      int nthreads) {
  orc::Scene* s = (orc::Scene*)sp;
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads)
  for (long i = 0; i < (long)n; ++i)
    occ[i] = orc::sceneOccluded(*s, getRay(rays, (size_t)i), brute != 0) ? 1 : 0;
  return 0;
}

}  // extern "C"
