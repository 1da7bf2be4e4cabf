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

struct Builder {
  const std::vector<Box>& pb;
  std::vector<float> cen;  // 3 per prim
  std::vector<int32_t> ids;
  BuildParams p;
  int par_depth = 0;

  Builder(const std::vector<Box>& b, const BuildParams& bp) : pb(b), p(bp) {
    cen.resize(3 * b.size());
    for (size_t i = 0; i < b.size(); ++i)
      for (int k = 0; k < 3; ++k) {
        float lo = b[i].lo[k], hi = b[i].hi[k];
        float c = 0.5f * (lo + hi);
        cen[3 * i + k] = std::isfinite(c) ? c : 0.0f;
      }
    ids.resize(b.size());
    for (size_t i = 0; i < b.size(); ++i) ids[i] = (int32_t)i;
  }

  Box range_box(int b, int e) const {
    Box r = empty_box();
    for (int i = b; i < e; ++i) grow(r, pb[ids[i]]);
    return r;
  }

  // Returns the split position (begin < mid < end); fills child boxes.
  int split(int b, int e, int depth, bool* make_leaf) {
    int n = e - b;
    *make_leaf = false;
    float cmin[3] = {INFINITY, INFINITY, INFINITY}, cmax[3] = {-INFINITY, -INFINITY, -INFINITY};
    Box nb = empty_box();
    for (int i = b; i < e; ++i) {
      int id = ids[i];
      grow(nb, pb[id]);
      for (int k = 0; k < 3; ++k) {
        cmin[k] = std::min(cmin[k], cen[3 * id + k]);
        cmax[k] = std::max(cmax[k], cen[3 * id + k]);
      }
    }
    bool balanced = depth > 40;  // keeps the total depth under kMaxDepth
    float best_cost = INFINITY;
    int best_axis = -1, best_bin = -1;
    const int NB = std::max(2, std::min(p.bins, 64));
    float parea = area(nb);
    if (!balanced && parea > 0.0f) {
      for (int ax = 0; ax < 3; ++ax) {
        float ext = cmax[ax] - cmin[ax];
        if (!(ext > 0.0f)) continue;
        Box bb[64];
        int bc[64];
        for (int i = 0; i < NB; ++i) {
          bb[i] = empty_box();
          bc[i] = 0;
        }
        float sc = NB / ext;
        for (int i = b; i < e; ++i) {
          int id = ids[i];
          int bi = std::min(NB - 1, std::max(0, (int)((cen[3 * id + ax] - cmin[ax]) * sc)));
          grow(bb[bi], pb[id]);
          bc[bi]++;
        }
        float ra[64];
        int rc[64];
        Box acc = empty_box();
        int cnt = 0;
        for (int i = NB - 1; i > 0; --i) {
          grow(acc, bb[i]);
          cnt += bc[i];
          ra[i] = area(acc);
          rc[i] = cnt;
        }
        acc = empty_box();
        cnt = 0;
        for (int i = 0; i < NB - 1; ++i) {
          grow(acc, bb[i]);
          cnt += bc[i];
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
      auto it = std::partition(ids.begin() + b, ids.begin() + e, [&](int id) {
        int bi = std::min(NB - 1, std::max(0, (int)((cen[3 * id + ax] - cmin[ax]) * sc)));
        return bi <= best_bin;
      });
      int mid = (int)(it - ids.begin());
      if (mid > b && mid < e) return mid;
    }
    // balanced / degenerate: median split on the widest centroid axis
    int ax = 0;
    for (int k = 1; k < 3; ++k)
      if (cmax[k] - cmin[k] > cmax[ax] - cmin[ax]) ax = k;
    int mid = b + n / 2;
    std::nth_element(ids.begin() + b, ids.begin() + mid, ids.begin() + e,
                     [&](int x, int y) { return cen[3 * x + ax] < cen[3 * y + ax]; });
    return mid;
  }

  // Builds the subtree for [b, e); returns a node (never a bare leaf unless
  // called for a child slot, see child()).
  void child(TNode* parent, int slot, int b, int e, int depth) {
    parent->box[slot] = pad_box(range_box(b, e));
    bool leaf = false;
    int mid = (e - b) <= 1 ? b : split(b, e, depth, &leaf);
    if ((e - b) <= 1) leaf = true;
    if (leaf) {
      parent->leaf[slot] = ~((b << kLeafCountBits) | (e - b));
      return;
    }
    parent->kid[slot].reset(new TNode());
    fill(parent->kid[slot].get(), b, mid, e, depth + 1);
  }

  void fill(TNode* node, int b, int mid, int e, int depth) {
    if (depth < par_depth && (e - b) > 65536) {
      auto f = std::async(std::launch::async, [&] { child(node, 0, b, mid, depth); });
      child(node, 1, mid, e, depth);
      f.get();
    } else {
      child(node, 0, b, mid, depth);
      child(node, 1, mid, e, depth);
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
    int mid = bd.split(0, n, 0, &leaf);
    if (leaf) mid = std::max(1, n / 2);  // the root is always an inner node
    bd.fill(&root, 0, mid, n, 1);
  }
  flatten(&root, nodes, order, bd.ids, 0, &st);
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
  const int base = (int)out->size();
  std::vector<std::pair<int, int>> todo{{root, base}};  // (BVH2 node, output node)
  out->push_back(WNode{});
  while (!todo.empty()) {
    const int s = todo.back().first, o = todo.back().second;
    todo.pop_back();
    std::vector<Box> cb{bin[s].child[0], bin[s].child[1]};
    std::vector<int32_t> cl{bin[s].link[0], bin[s].link[1]};
    // open the largest inner child until the node is `width` wide
    while ((int)cl.size() < width) {
      int pick = -1;
      float pa = -1.0f;
      for (size_t i = 0; i < cl.size(); ++i)
        if (cl[i] >= 0 && area(cb[i]) > pa) {
          pa = area(cb[i]);
          pick = (int)i;
        }
      if (pick < 0) break;
      const int g = cl[pick];
      cb[pick] = bin[g].child[0];
      cl[pick] = bin[g].link[0];
      cb.push_back(bin[g].child[1]);
      cl.push_back(bin[g].link[1]);
    }
    for (size_t i = 0; i < cl.size(); ++i) {
      if (cl[i] >= 0) {
        const int ni = (int)out->size();
        out->push_back(WNode{});
        todo.push_back({cl[i], ni});
        cl[i] = ni;
      }
    }
    (*out)[o].child = cb;
    (*out)[o].link = cl;
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
    for (int32_t l : w[n].link)
      if (l >= 0 && bound[l] < 0) {
        st.push_back(l);
        ready = false;
      }
    if (!ready) continue;
    st.pop_back();
    int m = 0;
    for (int32_t l : w[n].link)
      if (l >= 0) m = std::max(m, bound[l]);
    bound[n] = (int)w[n].link.size() - 1 + m;
  }
  return bound[root];
}

}  // namespace bgpu
