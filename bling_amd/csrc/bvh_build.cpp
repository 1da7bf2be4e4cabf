// Host binned-SAH BVH builder.  See bvh_build.h.
#include "bvh_build.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <future>
#include <memory>
#include <thread>

namespace bgpu {

namespace {

inline Box empty_box() {
  Box b;
  for (int k = 0; k < 3; ++k) {
    b.lo[k] = INFINITY;
    b.hi[k] = -INFINITY;
  }
  return b;
}
inline void grow(Box& a, const Box& b) {
  for (int k = 0; k < 3; ++k) {
    a.lo[k] = std::min(a.lo[k], b.lo[k]);
    a.hi[k] = std::max(a.hi[k], b.hi[k]);
  }
}
inline float area(const Box& b) {
  float d[3];
  for (int k = 0; k < 3; ++k) d[k] = std::max(0.0f, b.hi[k] - b.lo[k]);
  return 2.0f * (d[0] * d[1] + d[0] * d[2] + d[1] * d[2]);
}

struct TNode {
  Box box[2];
  int32_t leaf[2] = {0, 0};  // encoded leaf if child pointer null
  std::unique_ptr<TNode> kid[2];
};

// A primitive range's bounds and centroid bounds.
struct Extent {
  Box box;
  float cmin[3], cmax[3];
};
inline Extent empty_extent() {
  Extent x;
  x.box = empty_box();
  for (int k = 0; k < 3; ++k) {
    x.cmin[k] = INFINITY;
    x.cmax[k] = -INFINITY;
  }
  return x;
}
inline void grow(Extent& a, const Extent& b) {
  grow(a.box, b.box);
  for (int k = 0; k < 3; ++k) {
    a.cmin[k] = std::min(a.cmin[k], b.cmin[k]);
    a.cmax[k] = std::max(a.cmax[k], b.cmax[k]);
  }
}

constexpr int kMaxBins = 64;
// Per-axis bins of one pass: bounds, centroid bounds and counts.
struct Bins {
  Extent x[3][kMaxBins];
  int c[3][kMaxBins];
  void clear(int nb) {
    for (int a = 0; a < 3; ++a)
      for (int i = 0; i < nb; ++i) {
        x[a][i] = empty_extent();
        c[a][i] = 0;
      }
  }
};

// Runs f(begin, end) over [b, e) in up to `threads` contiguous chunks.
template <class F>
void parallel_chunks(int b, int e, int threads, F&& f) {
  const int n = e - b;
  const int t = std::max(1, std::min(threads, n / 65536));
  if (t == 1) {
    f(0, b, e);
    return;
  }
  std::vector<std::thread> th;
  for (int i = 0; i < t; ++i) {
    const int cb = b + (int)((long long)n * i / t), ce = b + (int)((long long)n * (i + 1) / t);
    th.emplace_back([&f, i, cb, ce] { f(i, cb, ce); });
  }
  for (auto& x : th) x.join();
}

// A primitive's bounds, centroid and id, stored in build order: ranges are
// partitioned in place, so every pass over a range reads it sequentially.
struct Prim {
  Box box;
  float c[3];
  int32_t id;
};

struct Builder {
  std::vector<Prim> prims;
  BuildParams p;
  int par_depth = 0;
  int threads = 1;

  Builder(const std::vector<Box>& b, const BuildParams& bp) : p(bp) {
    prims.resize(b.size());
    threads = p.threads > 0 ? p.threads : (int)std::max(1u, std::thread::hardware_concurrency());
    threads = std::min(threads, 16);
    parallel_chunks(0, (int)b.size(), threads, [&](int, int cb, int ce) {
      for (int i = cb; i < ce; ++i) {
        Prim& q = prims[i];
        q.box = b[i];
        for (int k = 0; k < 3; ++k) {
          float lo = b[i].lo[k], hi = b[i].hi[k];
          float c = 0.5f * (lo + hi);
          q.c[k] = std::isfinite(c) ? c : 0.0f;
        }
        q.id = (int32_t)i;
      }
    });
  }

  // Bounds and centroid bounds of [b, e) (one pass, chunked when large).
  Extent range_extent(int b, int e, int nthreads) const {
    std::vector<Extent> part(std::max(1, nthreads), empty_extent());
    parallel_chunks(b, e, nthreads, [&](int t, int cb, int ce) {
      Extent x = empty_extent();
      for (int i = cb; i < ce; ++i) {
        const Prim& q = prims[i];
        grow(x.box, q.box);
        for (int k = 0; k < 3; ++k) {
          x.cmin[k] = std::min(x.cmin[k], q.c[k]);
          x.cmax[k] = std::max(x.cmax[k], q.c[k]);
        }
      }
      part[t] = x;
    });
    Extent r = empty_extent();
    for (const Extent& x : part) grow(r, x);
    return r;
  }

  // Splits [b, e) with extent `ex`: returns the split position (begin < mid
  // < end) and the children's extents, or sets *make_leaf.  Binning of all
  // three axes is one pass over the range (chunked over threads when large);
  // the children's extents are unions of the chosen bins.
  int split(int b, int e, int depth, const Extent& ex, int nthreads, bool* make_leaf, Extent* le,
            Extent* re) {
    int n = e - b;
    *make_leaf = false;
    const float* cmin = ex.cmin;
    const float* cmax = ex.cmax;
    bool balanced = depth > 40;  // keeps the total depth under kMaxDepth
    float best_cost = INFINITY;
    int best_axis = -1, best_bin = -1;
    const int NB = std::max(2, std::min(p.bins, kMaxBins));
    float parea = area(ex.box);
    std::unique_ptr<Bins> bins;
    if (!balanced && parea > 0.0f) {
      float sc[3];
      bool use[3];
      for (int ax = 0; ax < 3; ++ax) {
        float ext = cmax[ax] - cmin[ax];
        use[ax] = ext > 0.0f;
        sc[ax] = use[ax] ? NB / ext : 0.0f;
      }
      const int nt = std::max(1, std::min(nthreads, n / 65536));
      std::vector<std::unique_ptr<Bins>> part(nt);
      for (auto& x : part) {
        x.reset(new Bins());
        x->clear(NB);
      }
      parallel_chunks(b, e, nt, [&](int t, int cb, int ce) {
        Bins& B = *part[t];
        for (int i = cb; i < ce; ++i) {
          const Prim& q = prims[i];
          for (int ax = 0; ax < 3; ++ax) {
            if (!use[ax]) continue;
            int bi = std::min(NB - 1, std::max(0, (int)((q.c[ax] - cmin[ax]) * sc[ax])));
            Extent& x = B.x[ax][bi];
            grow(x.box, q.box);
            for (int k = 0; k < 3; ++k) {
              x.cmin[k] = std::min(x.cmin[k], q.c[k]);
              x.cmax[k] = std::max(x.cmax[k], q.c[k]);
            }
            B.c[ax][bi]++;
          }
        }
      });
      bins = std::move(part[0]);
      for (int t = 1; t < nt; ++t)
        for (int ax = 0; ax < 3; ++ax)
          for (int i = 0; i < NB; ++i) {
            grow(bins->x[ax][i], part[t]->x[ax][i]);
            bins->c[ax][i] += part[t]->c[ax][i];
          }
      for (int ax = 0; ax < 3; ++ax) {
        if (!use[ax]) continue;
        float ra[kMaxBins];
        int rc[kMaxBins];
        Box acc = empty_box();
        int cnt = 0;
        for (int i = NB - 1; i > 0; --i) {
          grow(acc, bins->x[ax][i].box);
          cnt += bins->c[ax][i];
          ra[i] = area(acc);
          rc[i] = cnt;
        }
        acc = empty_box();
        cnt = 0;
        for (int i = 0; i < NB - 1; ++i) {
          grow(acc, bins->x[ax][i].box);
          cnt += bins->c[ax][i];
          if (cnt == 0 || rc[i + 1] == 0) continue;
          float c = p.cost_node + (area(acc) * cnt + ra[i + 1] * rc[i + 1]) / parea;
          if (c < best_cost) {
            best_cost = c;
            best_axis = ax;
            best_bin = i;
          }
        }
      }
    }
    if (!balanced && n <= p.max_leaf && (best_axis < 0 || (float)n <= best_cost)) {
      *make_leaf = true;
      return b;
    }
    if (!balanced && best_axis < 0 && n <= kMaxLeaf) {
      *make_leaf = true;
      return b;
    }
    if (balanced && n <= kMaxLeaf) {
      *make_leaf = true;
      return b;
    }
    if (best_axis >= 0) {
      float ext = cmax[best_axis] - cmin[best_axis];
      float sc = NB / ext;
      int ax = best_axis;
      auto it = std::partition(prims.begin() + b, prims.begin() + e, [&](const Prim& q) {
        int bi = std::min(NB - 1, std::max(0, (int)((q.c[ax] - cmin[ax]) * sc)));
        return bi <= best_bin;
      });
      int mid = (int)(it - prims.begin());
      if (mid > b && mid < e) {
        *le = empty_extent();
        *re = empty_extent();
        for (int i = 0; i < NB; ++i) grow(i <= best_bin ? *le : *re, bins->x[ax][i]);
        return mid;
      }
    }
    // balanced / degenerate: median split on the widest centroid axis
    int ax = 0;
    for (int k = 1; k < 3; ++k)
      if (cmax[k] - cmin[k] > cmax[ax] - cmin[ax]) ax = k;
    int mid = b + n / 2;
    std::nth_element(prims.begin() + b, prims.begin() + mid, prims.begin() + e,
                     [&](const Prim& x, const Prim& y) { return x.c[ax] < y.c[ax]; });
    *le = range_extent(b, mid, nthreads);
    *re = range_extent(mid, e, nthreads);
    return mid;
  }

  // Threads a subtree at `depth` may use: the builder's threads shared by
  // the 2^depth subtrees that run at once.
  int threads_at(int depth) const { return std::max(1, threads >> std::min(depth, 30)); }

  // Builds the subtree for [b, e) with extent ex into `slot` of `parent`.
  void child(TNode* parent, int slot, int b, int e, int depth, const Extent& ex) {
    parent->box[slot] = pad_box(ex.box);
    bool leaf = (e - b) <= 1;
    Extent le, re;
    int mid = b;
    if (!leaf) mid = split(b, e, depth, ex, threads_at(depth), &leaf, &le, &re);
    if (leaf) {
      parent->leaf[slot] = ~((b << kLeafCountBits) | (e - b));
      return;
    }
    parent->kid[slot].reset(new TNode());
    fill(parent->kid[slot].get(), b, mid, e, depth + 1, le, re);
  }

  void fill(TNode* node, int b, int mid, int e, int depth, const Extent& le, const Extent& re) {
    if (depth < par_depth && (e - b) > 65536) {
      auto f = std::async(std::launch::async, [&] { child(node, 0, b, mid, depth, le); });
      child(node, 1, mid, e, depth, re);
      f.get();
    } else {
      child(node, 0, b, mid, depth, le);
      child(node, 1, mid, e, depth, re);
    }
  }
};

void flatten(TNode* t, std::vector<BNode>* out, std::vector<int32_t>* order,
             const std::vector<int32_t>& ids, int depth, BuildStats* st) {
  int self = (int)out->size();
  out->push_back(BNode{});
  st->max_depth = std::max(st->max_depth, depth);
  for (int s = 0; s < 2; ++s) {
    (*out)[self].child[s] = t->box[s];
    if (t->kid[s]) {
      int idx = (int)out->size();
      flatten(t->kid[s].get(), out, order, ids, depth + 1, st);
      (*out)[self].link[s] = idx;
    } else {
      int enc = ~t->leaf[s];
      int first = enc >> kLeafCountBits, cnt = enc & kMaxLeaf;
      int nf = (int)order->size();
      for (int i = 0; i < cnt; ++i) order->push_back(ids[first + i]);
      (*out)[self].link[s] = ~((nf << kLeafCountBits) | cnt);
      st->leaves++;
      st->max_leaf_size = std::max(st->max_leaf_size, cnt);
    }
  }
}

}  // namespace

Box pad_box(const Box& b) {
  Box r = b;
  float ext = 0.0f;
  for (int k = 0; k < 3; ++k) ext = std::max(ext, b.hi[k] - b.lo[k]);
  if (!std::isfinite(ext)) ext = 0.0f;
  for (int k = 0; k < 3; ++k) {
    float m = std::max(std::fabs(b.lo[k]), std::fabs(b.hi[k]));
    if (!std::isfinite(m)) continue;
    float e = ext * 1e-5f + m * 4e-6f + 1e-30f;
    r.lo[k] = b.lo[k] - e;
    r.hi[k] = b.hi[k] + e;
  }
  return r;
}

void build_bvh2(const std::vector<Box>& prim_bounds, const BuildParams& p,
                std::vector<BNode>* nodes, std::vector<int32_t>* order, BuildStats* stats) {
  auto t0 = std::chrono::steady_clock::now();
  BuildStats st;
  nodes->clear();
  order->clear();
  Builder bd(prim_bounds, p);
  int hw = p.threads > 0 ? p.threads : (int)std::max(1u, std::thread::hardware_concurrency());
  bd.par_depth = 0;
  while ((1 << bd.par_depth) < hw * 2 && bd.par_depth < 8) bd.par_depth++;
  int n = (int)prim_bounds.size();
  TNode root;
  if (n == 0) {
    root.box[0] = root.box[1] = empty_box();
    root.leaf[0] = root.leaf[1] = ~0;
  } else if (n == 1) {
    root.box[0] = pad_box(prim_bounds[0]);
    root.leaf[0] = ~((0 << kLeafCountBits) | 1);
    root.box[1] = empty_box();
    root.leaf[1] = ~0;
  } else {
    bool leaf = false;
    const Extent all = bd.range_extent(0, n, bd.threads);
    Extent le, re;
    int mid = bd.split(0, n, 0, all, bd.threads, &leaf, &le, &re);
    if (leaf) {  // the root is always an inner node
      mid = std::max(1, n / 2);
      le = bd.range_extent(0, mid, 1);
      re = bd.range_extent(mid, n, 1);
    }
    bd.fill(&root, 0, mid, n, 1, le, re);
  }
  std::vector<int32_t> ids(n);
  parallel_chunks(0, n, bd.threads, [&](int, int cb, int ce) {
    for (int i = cb; i < ce; ++i) ids[i] = bd.prims[i].id;
  });
  flatten(&root, nodes, order, ids, 0, &st);
  st.nodes = (int)nodes->size();
  // SAH cost estimate (node visits + prim tests, per unit root area)
  double ra = 0.0;
  {
    Box u = empty_box();
    grow(u, (*nodes)[0].child[0]);
    grow(u, (*nodes)[0].child[1]);
    ra = area(u);
  }
  double sah = 0.0;
  if (ra > 0.0)
    for (const BNode& nd : *nodes)
      for (int s = 0; s < 2; ++s) {
        double a = area(nd.child[s]) / ra;
        if (nd.link[s] >= 0) sah += a * p.cost_node;
        else sah += a * ((~nd.link[s]) & kMaxLeaf);
      }
  st.sah = sah;
  st.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (stats) *stats = st;
}

int collapse_bvh(const std::vector<BNode>& bin, int root, int width, std::vector<WNode>* out) {
  // BVH2 inner index -> output index; children are assigned in DFS order
  width = std::max(2, std::min(width, kMaxWidth));
  out->reserve(out->size() + bin.size());  // at least one BVH2 node per wide node
  const int base = (int)out->size();
  std::vector<std::pair<int, int>> todo{{root, base}};  // (BVH2 node, output node)
  out->push_back(WNode{});
  while (!todo.empty()) {
    const int s = todo.back().first, o = todo.back().second;
    todo.pop_back();
    WNode w;
    w.child[0] = bin[s].child[0];
    w.child[1] = bin[s].child[1];
    w.link[0] = bin[s].link[0];
    w.link[1] = bin[s].link[1];
    w.n = 2;
    // open the largest inner child until the node is `width` wide
    while (w.n < width) {
      int pick = -1;
      float pa = -1.0f;
      for (int i = 0; i < w.n; ++i)
        if (w.link[i] >= 0 && area(w.child[i]) > pa) {
          pa = area(w.child[i]);
          pick = i;
        }
      if (pick < 0) break;
      const int g = w.link[pick];
      w.child[pick] = bin[g].child[0];
      w.link[pick] = bin[g].link[0];
      w.child[w.n] = bin[g].child[1];
      w.link[w.n] = bin[g].link[1];
      w.n++;
    }
    for (int i = 0; i < w.n; ++i) {
      if (w.link[i] >= 0) {
        const int ni = (int)out->size();
        out->push_back(WNode{});
        todo.push_back({w.link[i], ni});
        w.link[i] = ni;
      }
    }
    (*out)[o] = w;
  }
  return base;
}

int stack_bound(const std::vector<WNode>& w, int root) {
  // post-order over the tree: bound(n) = (children - 1) + max bound(inner child)
  std::vector<int> bound(w.size(), -1);
  std::vector<int> st{root};
  while (!st.empty()) {
    const int n = st.back();
    bool ready = true;
    for (int i = 0; i < w[n].n; ++i) {
      const int32_t l = w[n].link[i];
      if (l >= 0 && bound[l] < 0) {
        st.push_back(l);
        ready = false;
      }
    }
    if (!ready) continue;
    st.pop_back();
    int m = 0;
    for (int i = 0; i < w[n].n; ++i)
      if (w[n].link[i] >= 0) m = std::max(m, bound[w[n].link[i]]);
    bound[n] = w[n].n - 1 + m;
  }
  return bound[root];
}

}  // namespace bgpu
