// Scene flattening: the description (world-space vertices after the host's
// transPoint, shapes with their matrices, materials, lights, camera, film) is
// validated like mkTriangleMesh (Primitive/TriangleMesh.hs:46-55), bounded
// like mkKdTree (Primitive/KdTree.hs:107-112), organised into a BVH and copied
// once into device-resident arrays.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <thread>

#include "bgpu_internal.h"

namespace bgpu {

Ctx::~Ctx() {
  if (sched) destroy_scheduler(sched);
  sched = nullptr;
  if (plain) destroy_engine(plain);
  plain = nullptr;
  stage_rays.release();
  stage_hits.release();
  count_buf.release();
  work_buf.release();
  if (ev0) (void)hipEventDestroy(ev0);
  if (ev1) (void)hipEventDestroy(ev1);
  if (stream) (void)hipStreamDestroy(stream);
}

namespace {

inline float hmax_h(float x, float y) { return (x <= y) ? y : x; }
inline float hmin_h(float x, float y) { return (x <= y) ? x : y; }

inline void xp_host(const float* m, const float* p, float* o) {  // Transform.hs:247-258
  float xp = m[0] * p[0] + m[1] * p[1] + m[2] * p[2] + m[3];
  float yp = m[4] * p[0] + m[5] * p[1] + m[6] * p[2] + m[7];
  float zp = m[8] * p[0] + m[9] * p[1] + m[10] * p[2] + m[11];
  float wp = m[12] * p[0] + m[13] * p[1] + m[14] * p[2] + m[15];
  if (wp == 1.0f) {
    o[0] = xp; o[1] = yp; o[2] = zp;
  } else {
    o[0] = xp / wp; o[1] = yp / wp; o[2] = zp / wp;
  }
}

Box empty_box() {
  Box b;
  for (int k = 0; k < 3; ++k) {
    b.lo[k] = INFINITY;
    b.hi[k] = -INFINITY;
  }
  return b;
}
void extend_p(Box& b, const float* p) {  // extendAABBP (AABB.hs:43-46)
  for (int k = 0; k < 3; ++k) {
    b.lo[k] = hmin_h(b.lo[k], p[k]);
    b.hi[k] = hmax_h(b.hi[k], p[k]);
  }
}
void extend_b(Box& a, const Box& b) {  // extendAABB (AABB.hs:34-41)
  for (int k = 0; k < 3; ++k) {
    a.lo[k] = hmin_h(a.lo[k], b.lo[k]);
    a.hi[k] = hmax_h(a.hi[k], b.hi[k]);
  }
}

// transBox o2w (objectBounds s)  (Shape.hs:286-311, Transform.hs:281-296)
Box shape_world_bounds(const bgpu_shape& s) {
  float lo[3], hi[3];
  const float* p = s.params;
  switch (s.type) {
    case BGPU_SHAPE_BOX:
      lo[0] = p[0]; lo[1] = p[1]; lo[2] = p[2]; hi[0] = p[3]; hi[1] = p[4]; hi[2] = p[5];
      break;
    case BGPU_SHAPE_CYLINDER:
      lo[0] = -p[0]; lo[1] = -p[0]; lo[2] = p[1]; hi[0] = p[0]; hi[1] = p[0]; hi[2] = p[2];
      break;
    case BGPU_SHAPE_DISK:
      lo[0] = -p[1]; lo[1] = -p[1]; lo[2] = p[0]; hi[0] = p[1]; hi[1] = p[1]; hi[2] = p[0];
      break;
    case BGPU_SHAPE_QUAD:
      lo[0] = -p[0]; lo[1] = -p[1]; lo[2] = 0.0f; hi[0] = p[0]; hi[1] = p[1]; hi[2] = 0.0f;
      break;
    default:
      lo[0] = lo[1] = lo[2] = -p[0];
      hi[0] = hi[1] = hi[2] = p[0];
      break;
  }
  float c[8][3] = {{lo[0], lo[1], lo[2]}, {lo[0], lo[1], hi[2]}, {lo[0], hi[1], lo[2]},
                   {lo[0], hi[1], hi[2]}, {hi[0], lo[1], lo[2]}, {hi[0], lo[1], hi[2]},
                   {hi[0], hi[1], lo[2]}, {hi[0], hi[1], hi[2]}};
  Box b = empty_box();
  for (int i = 0; i < 8; ++i) {
    float o[3];
    xp_host(s.o2w, c[i], o);
    extend_p(b, o);
  }
  return b;
}

inline uint32_t fbits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return u;
}
// Upper bound of the largest singular value of the linear part of a
// row-major 4x4 matrix: sigma^2 <= the largest absolute row sum of L^T L.
double sigma_bound(const float* m) {
  double best = 0.0;
  for (int j = 0; j < 3; ++j) {
    double row = 0.0;
    for (int k = 0; k < 3; ++k) {
      double g = 0.0;
      for (int i = 0; i < 3; ++i) g += (double)m[4 * i + j] * (double)m[4 * i + k];
      row += std::fabs(g);
    }
    best = std::max(best, row);
  }
  return std::sqrt(best) * (1.0 + 1e-9);
}
// Runs f(begin, end) over [0, n) in contiguous chunks on up to 16 threads.
template <class F>
void parallel_for(size_t n, F&& f, size_t grain = 32768) {
  const size_t t = std::min<size_t>(16, std::max(1u, std::thread::hardware_concurrency()));
  const size_t k = std::min(t, n / grain);
  if (k <= 1) {
    f((size_t)0, n);
    return;
  }
  std::vector<std::thread> th;
  for (size_t i = 0; i < k; ++i) th.emplace_back([&f, n, k, i] { f(n * i / k, n * (i + 1) / k); });
  for (auto& x : th) x.join();
}
inline float bitsf(int32_t i) {
  float f;
  std::memcpy(&f, &i, 4);
  return f;
}

}  // namespace

int build_pt_data(Ctx* ctx, const bgpu_scene_desc* d, Scene* sc);

int flatten_and_upload(Ctx* ctx, const bgpu_scene_desc* d, Scene* sc) {
  // -- validation (TriangleMesh.hs:46-55 raise `error`; here: BGPU_E_INVALID)
  if (d->num_triangles && (!d->indices || !d->vertices))
    return fail(BGPU_E_INVALID, "triangles given without vertices/indices");
  if (d->num_shapes && !d->shapes) return fail(BGPU_E_INVALID, "shapes pointer is NULL");
  if ((uint64_t)d->num_triangles + d->num_shapes > (1u << 27) - 1)
    return fail(BGPU_E_INVALID, "too many primitives (limit 2^27-1)");
  for (uint64_t i = 0; i < 3ull * d->num_triangles; ++i) {
    int v = d->indices[i];
    if (v < 0) return fail(BGPU_E_INVALID, "mkTriangleMesh: contains negative indices");
    if ((uint32_t)v >= d->num_vertices)
      return fail(BGPU_E_INVALID, "mkTriangleMesh: contains out of bounds indices");
  }
  uint32_t nmesh = d->num_meshes ? d->num_meshes : 1;
  if (d->tri_mesh)
    for (uint32_t i = 0; i < d->num_triangles; ++i)
      if (d->tri_mesh[i] < 0 || (uint32_t)d->tri_mesh[i] >= nmesh)
        return fail(BGPU_E_INVALID, "tri_mesh index out of range");
  for (uint32_t i = 0; i < d->num_shapes; ++i) {
    const bgpu_shape& s = d->shapes[i];
    if (s.type < BGPU_SHAPE_BOX || s.type > BGPU_SHAPE_SPHERE)
      return fail(BGPU_E_INVALID, "unknown shape type");
  }

  sc->ntri = d->num_triangles;
  sc->nshape = d->num_shapes;
  const uint32_t np = d->num_triangles + d->num_shapes;

  // shading data first: the device build reads its vertices and indices
  if (int rc = build_pt_data(ctx, d, sc)) return rc;

  // -- the triangle tree on the GPU when the scene is large enough; it
  // reports a fallback for inputs the host builder handles differently
  DeviceBvh dbvh;
  bool dev_tri = false;
  if (ctx->bvh_device && d->num_triangles > (uint32_t)kSmall) {
    bool fallback = true;
    const int rc = dbvh.build(ctx, sc->verts.as<float>(), sc->tri_idx.as<int32_t>(), (int)d->num_triangles, &fallback);
    if (rc == BGPU_E_BUDGET || rc == BGPU_E_NOMEM) {
      dbvh.release();  // its scratch does not fit: the host builds with less device memory
      fallback = true;
    } else if (rc) {
      return rc;
    }
    dev_tri = !fallback;
  }

  // -- primitive bounds and the root box fold (KdTree.hs:111)
  std::vector<Box> pb(dev_tri ? d->num_shapes : np);
  const uint32_t soff = dev_tri ? 0 : d->num_triangles;  // shapes' first index in pb
  Box root = empty_box();  // the fold in primitive order
  if (!dev_tri) {
    parallel_for(d->num_triangles, [&](size_t tb, size_t te) {
      for (size_t t = tb; t < te; ++t) {
        Box b = empty_box();  // triangleBounds (TriangleMesh.hs:136-138)
        for (int k = 0; k < 3; ++k) extend_p(b, d->vertices + 3 * (size_t)d->indices[3 * t + k]);
        pb[t] = b;
      }
    });
    for (uint32_t i = 0; i < d->num_triangles; ++i) extend_b(root, pb[i]);
  } else {
    // the same fold without storing the boxes: per-chunk folds combined in
    // chunk order (extendAABB keeps the earlier value on ties)
    std::vector<Box> part(64, empty_box());
    const size_t nt = d->num_triangles;
    parallel_for(64, [&](size_t cb, size_t ce) {
      for (size_t c = cb; c < ce; ++c) {
        Box acc = empty_box();
        for (size_t t = nt * c / 64; t < nt * (c + 1) / 64; ++t) {
          Box b = empty_box();
          for (int k = 0; k < 3; ++k) extend_p(b, d->vertices + 3 * (size_t)d->indices[3 * t + k]);
          extend_b(acc, b);
        }
        part[c] = acc;
      }
    }, 1);
    for (const Box& b : part) extend_b(root, b);
  }
  for (uint32_t s = 0; s < d->num_shapes; ++s) pb[soff + s] = shape_world_bounds(d->shapes[s]);
  for (uint32_t s = 0; s < d->num_shapes; ++s) extend_b(root, pb[soff + s]);

  // -- BVHs: one over the triangles, one over the shapes, built binary and
  // collapsed to the context's width into one node array
  std::vector<BNode> tnodes, snodes;
  std::vector<int32_t> order;
  BuildParams bp;
  float aniso = 1.0f;
  std::vector<float> hfac;  // per shape: sigma(o2w) sigma(w2o)^2 / (2 r), +inf if not a sphere
  if (dev_tri) {
    sc->build = BuildStats{};
    sc->build.device = 1;
    sc->build.leaves = dbvh.leaves;
    sc->build.max_depth = dbvh.max_depth;
    sc->build.build_ms = dbvh.build_ms;
  } else if (d->num_triangles) {
    if (d->num_shapes == 0) {
      build_bvh2(pb, bp, &tnodes, &order, &sc->build);
    } else {
      std::vector<Box> tb(pb.begin(), pb.begin() + d->num_triangles);
      build_bvh2(tb, bp, &tnodes, &order, &sc->build);
    }
  }
  if (d->num_shapes) {
    std::vector<Box> sb(pb.begin() + soff, pb.end());
    std::vector<int32_t> so;
    BuildStats sst;
    build_bvh2(sb, bp, &snodes, &so, &sst);
    // the shape leaves follow the triangle leaves in the leaf records
    const int32_t loff = (int32_t)d->num_triangles;
    for (BNode& n : snodes)
      for (int s = 0; s < 2; ++s)
        if (n.link[s] < 0) {
          int enc = ~n.link[s];
          n.link[s] = ~((((enc >> kLeafCountBits) + loff) << kLeafCountBits) | (enc & kMaxLeaf));
        }
    for (int32_t id : so) order.push_back(id + (int32_t)d->num_triangles);
    sc->build.leaves += sst.leaves;
    sc->build.max_depth = std::max(sc->build.max_depth, sst.max_depth);
    // transform factors of the false-hit margin (isect.h)
    hfac.resize(d->num_shapes);
    for (uint32_t s = 0; s < d->num_shapes; ++s) {
      const bgpu_shape& sh = d->shapes[s];
      const double s1 = sigma_bound(sh.o2w), s2 = sigma_bound(sh.w2o);
      const double a = std::isfinite(s1 * s2) ? std::min(s1 * s2, 1e6) : 1e6;
      aniso = std::max(aniso, (float)(a * (1.0 + 1e-6)));
      const double h = s1 * s2 * s2 / (2.0 * (double)sh.params[0]);
      hfac[s] = (sh.type == BGPU_SHAPE_SPHERE && sh.params[0] > 0.0f && std::isfinite(h) && h < 1e30)
                  ? (float)(h * (1.0 + 1e-6))
                  : INFINITY;
    }
  }
  // the widest width whose worst-case traversal stack fits the device stack
  // `nodes` holds the host-collapsed trees; with the device triangle tree
  // it holds the shape tree only, placed after the dev_count device nodes
  std::vector<WNode> nodes;
  int32_t tri_root = -1, shape_root = -1;
  int dev_count = 0;
  int width = ctx->bvh_width;
  for (;; width /= 2) {
    nodes.clear();
    nodes.reserve(tnodes.size() + snodes.size() + 2);
    int bound = 0;
    if (dev_tri) {
      int tb = 0;
      if (int rc = dbvh.collapse(ctx, width, &dev_count, &tb)) return rc;
      tri_root = 0;
      bound = tb;
    } else {
      tri_root = tnodes.empty() ? -1 : collapse_bvh(tnodes, 0, width, &nodes);
      if (tri_root >= 0) bound = std::max(bound, stack_bound(nodes, tri_root));
    }
    shape_root = snodes.empty() ? -1 : collapse_bvh(snodes, 0, width, &nodes);
    if (shape_root >= 0) bound = std::max(bound, stack_bound(nodes, shape_root));
    sc->build.stack_bound = bound;
    if (bound <= kStackCapacity || width == 2) break;
  }
  if (sc->build.stack_bound > kStackCapacity)
    return fail(BGPU_E_INVALID, "BVH too deep for the traversal stack");
  sc->build.width = width;
  sc->build.nodes = dev_count + (int)nodes.size();
  const int obase = dev_tri ? (int)d->num_triangles : 0;  // leaf position of order[0]

  // node layout, 8 * width floats: lo.x[w] hi.x[w] lo.y[w] hi.y[w] lo.z[w]
  // hi.z[w] link[w] h[w]; unused slots carry the link kNoChild and repeat
  // child 0's box and h (so that the union of a node's child boxes is its box)
  const int W = width;
  std::vector<float> hchild(nodes.size() * W, 0.0f);  // shape tree: max h below each child
  std::function<float(int32_t)> hmax = [&](int32_t link) -> float {
    float m = 0.0f;
    if (link >= 0) {
      const WNode& n = nodes[link];
      for (int c = 0; c < n.n; ++c) {
        const float v = hmax(n.link[c]);
        hchild[(size_t)link * W + c] = v;
        m = std::max(m, v);
      }
      return m;
    }
    const int enc = ~link;
    for (int k = enc >> kLeafCountBits; k < (enc >> kLeafCountBits) + (enc & kMaxLeaf); ++k)
      m = std::max(m, hfac[order[k - obase] - (int32_t)d->num_triangles]);
    return m;
  };
  if (shape_root >= 0) (void)hmax(shape_root);
  std::vector<float> hn((size_t)8 * W * std::max<size_t>(nodes.size(), 1), 0.0f);
  parallel_for(nodes.size(), [&](size_t ib, size_t ie) {
    for (size_t i = ib; i < ie; ++i) {
      const WNode& n = nodes[i];
      float* o = &hn[(size_t)8 * W * i];
      for (int c = 0; c < W; ++c) {
        const bool used = c < n.n;
        const Box& b = n.child[used ? c : 0];
        for (int k = 0; k < 3; ++k) {
          o[(2 * k) * W + c] = b.lo[k];
          o[(2 * k + 1) * W + c] = b.hi[k];
        }
        const int32_t lk = used ? n.link[c] : kNoChild;
        o[6 * W + c] = bitsf(lk >= 0 && lk != kNoChild ? lk + dev_count : lk);
        o[7 * W + c] = hchild[i * W + (used ? c : 0)];
      }
    }
  });
  // leaf records of order[k] at position obase + k (the device writes the
  // triangles' records itself)
  std::vector<float> hl(12 * std::max<size_t>(order.size(), 1), 0.0f);
  parallel_for(order.size(), [&](size_t kb, size_t ke) {
    for (size_t k = kb; k < ke; ++k) {
      int id = order[k];
      float* o = &hl[12 * k];
      if ((uint32_t)id < d->num_triangles) {
        const float* p1 = d->vertices + 3 * (size_t)d->indices[3 * (size_t)id];
        const float* p2 = d->vertices + 3 * (size_t)d->indices[3 * (size_t)id + 1];
        const float* p3 = d->vertices + 3 * (size_t)d->indices[3 * (size_t)id + 2];
        o[0] = p1[0]; o[1] = p1[1]; o[2] = p1[2]; o[3] = bitsf(id);
        // e1 = p2 - p1, e2 = p3 - p1: the same IEEE subtractions as TriangleMesh.hs:169
        o[4] = p2[0] - p1[0]; o[5] = p2[1] - p1[1]; o[6] = p2[2] - p1[2]; o[7] = bitsf(0);
        o[8] = p3[0] - p1[0]; o[9] = p3[1] - p1[1]; o[10] = p3[2] - p1[2]; o[11] = 0.0f;
      } else {
        int si = id - (int)d->num_triangles;
        o[3] = bitsf(id);
        o[7] = bitsf(~si);
      }
    }
  });
  std::vector<DShape> hs(std::max<uint32_t>(d->num_shapes, 1));
  for (uint32_t s = 0; s < d->num_shapes; ++s) {
    const bgpu_shape& in = d->shapes[s];
    DShape& o = hs[s];
    std::memset(&o, 0, sizeof(o));
    o.type = in.type;
    o.material = in.material;
    o.light = -1;
    std::memcpy(o.p, in.params, sizeof(o.p));
    std::memcpy(o.w2o, in.w2o, sizeof(o.w2o));
  }

  hipStream_t st = ctx->stream;
  if (dev_tri) {
    const size_t nb = (size_t)8 * W * 4;  // bytes per node
    if (int rc = sc->nodes.alloc(ctx, nb * (dev_count + nodes.size()))) return rc;
    BGPU_HIP(hipMemcpyAsync(sc->nodes.ptr, dbvh.wide_nodes(), nb * dev_count, hipMemcpyDeviceToDevice, st));
    if (!nodes.empty())
      BGPU_HIP(hipMemcpyAsync(static_cast<char*>(sc->nodes.ptr) + nb * dev_count, hn.data(), nb * nodes.size(),
                              hipMemcpyHostToDevice, st));
    if (int rc = sc->leaf.alloc(ctx, 48ull * (d->num_triangles + d->num_shapes))) return rc;
    if (int rc = dbvh.leaves_to(ctx, sc->verts.as<float>(), sc->tri_idx.as<int32_t>(), sc->leaf.as<float4>()))
      return rc;
    if (!order.empty())
      BGPU_HIP(hipMemcpyAsync(static_cast<char*>(sc->leaf.ptr) + 48ull * d->num_triangles, hl.data(),
                              48ull * order.size(), hipMemcpyHostToDevice, st));
  } else {
    if (int rc = sc->nodes.upload(ctx, hn.data(), hn.size() * 4, st)) return rc;
    if (int rc = sc->leaf.upload(ctx, hl.data(), hl.size() * 4, st)) return rc;
  }
  // shape light indices are assigned by build_pt_data (light list order)
  for (uint32_t s = 0; s < d->num_shapes; ++s) hs[s].light = sc->pt.shape_light_host.empty() ? -1 : sc->pt.shape_light_host[s];
  if (int rc = sc->shapes.upload(ctx, hs.data(), hs.size() * sizeof(DShape), st)) return rc;
  BGPU_HIP(hipStreamSynchronize(st));

  DevScene& ds = sc->dev;
  ds.nodes = sc->nodes.as<float4>();
  ds.leaf = sc->leaf.as<float4>();
  ds.shapes = sc->shapes.as<DShape>();
  for (int k = 0; k < 3; ++k) {
    ds.root_lo[k] = root.lo[k];
    ds.root_hi[k] = root.hi[k];
  }
  ds.num_nodes = (int32_t)(dev_count + nodes.size());
  ds.width = width;
  ds.tri_root = tri_root;
  ds.shape_root = shape_root < 0 ? -1 : shape_root + dev_count;
  ds.shape_pad_k = kSphereK0 * aniso;
  ds.shape_floor_k = 64.0f * 0x1p-24f * aniso;
  sc->film = d->film;
  sc->camera = d->camera;
  return BGPU_OK;
}

}  // namespace bgpu
