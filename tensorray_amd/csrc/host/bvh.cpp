// Binned-SAH BVH build, collapsed to a 4-wide BVH laid out for the GPU
// traversal kernel (128-byte nodes, triangles reordered into leaf order).
//
// Replaces the OptiX GAS build/compaction/update of the reference
// (LibTensorRay/Renderer/Scene.cpp:171-366, 368-469).
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <thread>
#include <vector>

#include "scene_host.h"

namespace trc {

namespace {

struct Box {
  float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
  float hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  void grow(const float* p) {
    for (int k = 0; k < 3; ++k) { lo[k] = std::min(lo[k], p[k]); hi[k] = std::max(hi[k], p[k]); }
  }
  void grow(const Box& b) {
    for (int k = 0; k < 3; ++k) { lo[k] = std::min(lo[k], b.lo[k]); hi[k] = std::max(hi[k], b.hi[k]); }
  }
  bool valid() const { return lo[0] <= hi[0]; }
  double area() const {
    if (!valid()) return 0.0;
    double dx = (double)hi[0] - lo[0], dy = (double)hi[1] - lo[1], dz = (double)hi[2] - lo[2];
    return 2.0 * (dx * dy + dy * dz + dz * dx);
  }
};

struct BinNode {
  Box box;
  int32_t left = -1, right = -1;  // children (inner) or -1
  uint32_t first = 0, count = 0;  // leaf range in the order array
  bool leaf() const { return left < 0; }
};

constexpr int kBins = 32;
constexpr double kCostTraverse = kSahTraverse;
constexpr double kCostIntersect = kSahIntersect;
constexpr int kMaxLeaf = 4;

struct Builder {
  const std::vector<float>& pos;
  const std::vector<uint32_t>& idx;
  std::vector<Box> tbox;
  std::vector<float> cent;  // [T][3]
  std::vector<uint32_t> order;
  std::vector<BinNode> nodes;

  Builder(const std::vector<float>& p, const std::vector<uint32_t>& i) : pos(p), idx(i) {}

  void prepare_range(size_t lo, size_t hi) {
    for (size_t t = lo; t < hi; ++t) {
      Box b;
      for (int j = 0; j < 3; ++j) b.grow(&pos[3 * (size_t)idx[3 * t + j]]);
      tbox[t] = b;
      for (int k = 0; k < 3; ++k) cent[3 * t + k] = 0.5f * (b.lo[k] + b.hi[k]);
      order[t] = (uint32_t)t;
    }
  }

  // prepare() on `threads` threads over contiguous ranges (same result)
  void prepare_par(unsigned threads) {
    const size_t T = idx.size() / 3;
    tbox.resize(T);
    cent.resize(3 * T);
    order.resize(T);
    if (threads <= 1 || T < ((size_t)1 << 16)) {
      prepare_range(0, T);
      return;
    }
    std::vector<std::thread> th;
    for (unsigned i = 0; i < threads; ++i) th.emplace_back([&, i] { prepare_range(T * i / threads, T * (i + 1) / threads); });
    for (auto& x : th) x.join();
  }

  static int32_t make_leaf(std::vector<BinNode>& out, uint32_t first, uint32_t count, const Box& box) {
    BinNode n;
    n.box = box;
    n.first = first;
    n.count = count;
    out.push_back(n);
    return (int32_t)out.size() - 1;
  }

  // The subtree over order[first, first + count) appended to `out` in
  // preorder (node, left subtree, right subtree); its root is the first node
  // appended.  With par_depth > 0 the two subtrees of a large node are built
  // on two threads into their own arrays and appended left then right, which
  // gives the same array as the sequential build.
  int32_t build_into(std::vector<BinNode>& out, uint32_t first, uint32_t count, int par_depth) {
    Box box, cbox;
    for (uint32_t i = first; i < first + count; ++i) {
      box.grow(tbox[order[i]]);
      cbox.grow(&cent[3 * (size_t)order[i]]);
    }
    if (count <= 1) return make_leaf(out, first, count, box);
    double best_cost = DBL_MAX;
    int best_axis = -1, best_split = -1;
    for (int axis = 0; axis < 3; ++axis) {
      float ext = cbox.hi[axis] - cbox.lo[axis];
      if (!(ext > 0.0f)) continue;
      Box bb[kBins];
      uint32_t bc[kBins] = {0};
      double scale = (double)kBins / ext;
      for (uint32_t i = first; i < first + count; ++i) {
        uint32_t t = order[i];
        int b = (int)((cent[3 * (size_t)t + axis] - cbox.lo[axis]) * scale);
        b = std::min(std::max(b, 0), kBins - 1);
        bb[b].grow(tbox[t]);
        bc[b]++;
      }
      double right_area[kBins];
      uint32_t right_cnt[kBins];
      Box acc;
      uint32_t ac = 0;
      for (int b = kBins - 1; b > 0; --b) {
        acc.grow(bb[b]);
        ac += bc[b];
        right_area[b] = acc.area();
        right_cnt[b] = ac;
      }
      Box lacc;
      uint32_t lc = 0;
      for (int b = 0; b < kBins - 1; ++b) {
        lacc.grow(bb[b]);
        lc += bc[b];
        if (lc == 0 || right_cnt[b + 1] == 0) continue;
        double cost = lacc.area() * lc + right_area[b + 1] * right_cnt[b + 1];
        if (cost < best_cost) { best_cost = cost; best_axis = axis; best_split = b; }
      }
    }
    double parent_area = std::max(box.area(), 1e-30);
    double leaf_cost = kCostIntersect * count;
    double split_cost = best_axis >= 0 ? kCostTraverse + kCostIntersect * best_cost / parent_area : DBL_MAX;
    if (count <= (uint32_t)kMaxLeaf && leaf_cost <= split_cost) return make_leaf(out, first, count, box);
    uint32_t mid;
    if (best_axis < 0) {
      // all centroids coincide: split the range in half
      mid = first + count / 2;
    } else {
      float ext = cbox.hi[best_axis] - cbox.lo[best_axis];
      double scale = (double)kBins / ext;
      auto it = std::partition(order.begin() + first, order.begin() + first + count, [&](uint32_t t) {
        int b = (int)((cent[3 * (size_t)t + best_axis] - cbox.lo[best_axis]) * scale);
        b = std::min(std::max(b, 0), kBins - 1);
        return b <= best_split;
      });
      mid = (uint32_t)(it - order.begin());
      if (mid == first || mid == first + count) mid = first + count / 2;
    }
    int32_t self = (int32_t)out.size();
    out.push_back(BinNode{});
    out[self].box = box;
    if (par_depth > 0 && count >= kParallelMin) {
      std::vector<BinNode> lv, rv;
      std::thread th([&] { build_into(lv, first, mid - first, par_depth - 1); });
      build_into(rv, mid, first + count - mid, par_depth - 1);
      th.join();
      const int32_t l = (int32_t)out.size();
      for (BinNode n : lv) {
        if (!n.leaf()) { n.left += l; n.right += l; }
        out.push_back(n);
      }
      const int32_t r = (int32_t)out.size();
      for (BinNode n : rv) {
        if (!n.leaf()) { n.left += r; n.right += r; }
        out.push_back(n);
      }
      out[self].left = l;
      out[self].right = r;
      return self;
    }
    int32_t l = build_into(out, first, mid - first, 0);
    int32_t r = build_into(out, mid, first + count - mid, 0);
    out[self].left = l;
    out[self].right = r;
    return self;
  }

  // Subtrees smaller than this are built on the calling thread.
  static constexpr uint32_t kParallelMin = 1u << 15;
};

// Outward rounding of a box so that the slab test stays conservative even
// with the rounding of the traversal arithmetic.
inline float pad_lo(float x) { return x - (std::fabs(x) * 9.5367431640625e-7f + 1e-7f); }
inline float pad_hi(float x) { return x + (std::fabs(x) * 9.5367431640625e-7f + 1e-7f); }

void set_child_box(Bvh4Node& n, int slot, const Box& b) {
  n.lox[slot] = pad_lo(b.lo[0]); n.loy[slot] = pad_lo(b.lo[1]); n.loz[slot] = pad_lo(b.lo[2]);
  n.hix[slot] = pad_hi(b.hi[0]); n.hiy[slot] = pad_hi(b.hi[1]); n.hiz[slot] = pad_hi(b.hi[2]);
}

void set_empty(Bvh4Node& n, int slot) {
  // an inverted box never intersects
  n.lox[slot] = n.loy[slot] = n.loz[slot] = FLT_MAX;
  n.hix[slot] = n.hiy[slot] = n.hiz[slot] = -FLT_MAX;
  n.child[slot] = kEmpty;
}

struct Collapser {
  const std::vector<BinNode>& bn;
  std::vector<Bvh4Node>& out;
  std::vector<TriRecord>& tris;
  const std::vector<uint32_t>& order;
  const std::vector<float>& pos;
  const std::vector<uint32_t>& idx;
  BvhBuildStats& st;
  std::vector<int> depth_of;  // depth of every 4-wide node (root 1)

  // Emits one leaf's triangles, splitting ranges larger than kMaxLeafTris.
  uint32_t emit_leaf(const BinNode& n) {
    uint32_t first = (uint32_t)tris.size();
    for (uint32_t i = n.first; i < n.first + n.count; ++i) {
      uint32_t t = order[i];
      TriRecord r;
      for (int k = 0; k < 3; ++k) {
        r.v0[k] = pos[3 * (size_t)idx[3 * t] + k];
        r.v1[k] = pos[3 * (size_t)idx[3 * t + 1] + k];
        r.v2[k] = pos[3 * (size_t)idx[3 * t + 2] + k];
      }
      std::memcpy(&r.v0[3], &t, 4);
      r.v1[3] = 0.0f;
      r.v2[3] = 0.0f;
      tris.push_back(r);
    }
    st.leaves++;
    return make_leaf(first, n.count);
  }

  // Converts binary node `b` (an inner node) into a 4-wide node at index
  // `slot_index`; returns the worst-case stack need below it.
  int collapse(int32_t b, uint32_t out_index, int depth) {
    st.max_depth = std::max(st.max_depth, depth);
    if (depth_of.size() <= out_index) depth_of.resize(out_index + 1, 0);
    depth_of[out_index] = depth;
    std::vector<int32_t> kids = {bn[b].left, bn[b].right};
    while (kids.size() < 4) {
      int best = -1;
      double best_area = -1.0;
      for (size_t i = 0; i < kids.size(); ++i)
        if (!bn[kids[i]].leaf() && bn[kids[i]].box.area() > best_area) {
          best_area = bn[kids[i]].box.area();
          best = (int)i;
        }
      if (best < 0) break;
      int32_t k = kids[best];
      kids.erase(kids.begin() + best);
      kids.push_back(bn[k].left);
      kids.push_back(bn[k].right);
    }
    Bvh4Node node;
    for (int s = 0; s < 4; ++s) set_empty(node, s);
    int need = 0;
    int nk = (int)kids.size();
    std::vector<std::pair<int, uint32_t>> inner;  // (slot, binary id)
    for (int s = 0; s < nk; ++s) {
      const BinNode& c = bn[kids[s]];
      set_child_box(node, s, c.box);
      if (c.leaf()) {
        node.child[s] = emit_leaf(c);
      } else {
        inner.push_back({s, (uint32_t)kids[s]});
      }
    }
    out[out_index] = node;
    for (auto& pr : inner) {
      uint32_t ci = (uint32_t)out.size();
      out.push_back(Bvh4Node{});
      out[out_index].child[pr.first] = ci;
      // at most nk-1 siblings are pending on the stack while a child is visited
      int sub = collapse((int32_t)pr.second, ci, depth + 1);
      need = std::max(need, (nk - 1) + sub);
    }
    return need;
  }
};

}  // namespace

void bvh_build(const std::vector<float>& pos, const std::vector<uint32_t>& idx,
               std::vector<Bvh4Node>& nodes, std::vector<TriRecord>& tris,
               std::vector<uint32_t>& order, BvhBuildStats& stats, BvhRefitPlan* plan) {
  const size_t T = idx.size() / 3;
  if (T == 0) throw std::runtime_error("bvh_build: no triangles");
  if (T >= (size_t)kLeafFirstMask) throw std::runtime_error("bvh_build: too many triangles");
  Builder b(pos, idx);
  // threads: up to 16, as 2^par_depth subtree builders
  const unsigned hc = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  int par_depth = 0;
  while ((2u << par_depth) <= hc) ++par_depth;
  b.prepare_par(hc);
  b.nodes.reserve(2 * T);
  int32_t root = b.build_into(b.nodes, 0, (uint32_t)T, par_depth);
  stats = BvhBuildStats{};
  nodes.clear();
  tris.clear();
  tris.reserve(T);
  nodes.reserve(T / 2 + 1);
  nodes.push_back(Bvh4Node{});
  Collapser c{b.nodes, nodes, tris, b.order, pos, idx, stats};
  if (b.nodes[root].leaf()) {
    // single-leaf scene: a root node with one leaf child
    Bvh4Node n;
    for (int s = 0; s < 4; ++s) set_empty(n, s);
    set_child_box(n, 0, b.nodes[root].box);
    n.child[0] = c.emit_leaf(b.nodes[root]);
    nodes[0] = n;
    stats.stack_need = 1;
  } else {
    stats.stack_need = c.collapse(root, 0, 1) + 1;
  }
  // refit plan: nodes grouped by depth, deepest level first
  if (plan) {
    c.depth_of.resize(nodes.size(), 1);
    const int maxd = *std::max_element(c.depth_of.begin(), c.depth_of.end());
    plan->level_off.assign(1, 0u);
    plan->nodes.clear();
    plan->nodes.reserve(nodes.size());
    for (int d = maxd; d >= 1; --d) {
      for (size_t i = 0; i < nodes.size(); ++i)
        if (c.depth_of[i] == d) plan->nodes.push_back((uint32_t)i);
      plan->level_off.push_back((uint32_t)plan->nodes.size());
    }
  }
  // the triangle order is the leaf order of the records
  order.resize(T);
  for (size_t i = 0; i < T; ++i) std::memcpy(&order[i], &tris[i].v0[3], 4);
  stats.nodes = (int32_t)nodes.size();
  // SAH cost of the 4-wide tree (diagnostic)
  double root_area = b.nodes[root].box.area();
  double cost = 0.0;
  for (const auto& n : nodes)
    for (int s = 0; s < 4; ++s) {
      if (n.child[s] == kEmpty) continue;
      double dx = (double)n.hix[s] - n.lox[s], dy = (double)n.hiy[s] - n.loy[s], dz = (double)n.hiz[s] - n.loz[s];
      double a = 2.0 * (dx * dy + dy * dz + dz * dx) / std::max(root_area, 1e-30);
      cost += a * (child_is_leaf(n.child[s]) ? leaf_count(n.child[s]) * kCostIntersect : kCostTraverse);
    }
  stats.sah_cost = cost;
}

void bvh_refit(const std::vector<float>& pos, const std::vector<uint32_t>& idx,
               const std::vector<uint32_t>& order, std::vector<Bvh4Node>& nodes,
               std::vector<TriRecord>& tris) {
  // Triangle records follow the new positions; boxes are recomputed bottom-up
  // (children always have larger indices than their parent).
  for (size_t i = 0; i < order.size(); ++i) {
    uint32_t t = order[i];
    for (int k = 0; k < 3; ++k) {
      tris[i].v0[k] = pos[3 * (size_t)idx[3 * t] + k];
      tris[i].v1[k] = pos[3 * (size_t)idx[3 * t + 1] + k];
      tris[i].v2[k] = pos[3 * (size_t)idx[3 * t + 2] + k];
    }
  }
  std::vector<Box> nb(nodes.size());
  for (size_t ni = nodes.size(); ni-- > 0;) {
    Bvh4Node& n = nodes[ni];
    Box all;
    for (int s = 0; s < 4; ++s) {
      uint32_t c = n.child[s];
      if (c == kEmpty) continue;
      Box cb;
      if (child_is_leaf(c)) {
        for (uint32_t i = leaf_first(c); i < leaf_first(c) + leaf_count(c); ++i) {
          cb.grow(tris[i].v0); cb.grow(tris[i].v1); cb.grow(tris[i].v2);
        }
      } else {
        cb = nb[c];
      }
      set_child_box(n, s, cb);
      all.grow(cb);
    }
    nb[ni] = all;
  }
}

}  // namespace trc
