// Binned-SAH BVH2 builder for the GPU layout (ssolh_bvh_node: both child
// boxes in the parent). Used for each object's BLAS over its triangles and
// for the TLAS over instance boxes.
#include <algorithm>
#include <cfloat>
#include <cstring>

#include "ssol_internal.hpp"

namespace ssol {

namespace {

constexpr int NBINS = 16;

struct Box {
  float lo[3], hi[3];
  void reset() {
    for (int k = 0; k < 3; ++k) { lo[k] = FLT_MAX; hi[k] = -FLT_MAX; }
  }
  void grow(const Box& b) {
    for (int k = 0; k < 3; ++k) { lo[k] = std::min(lo[k], b.lo[k]); hi[k] = std::max(hi[k], b.hi[k]); }
  }
  void grow(const float* p) {
    for (int k = 0; k < 3; ++k) { lo[k] = std::min(lo[k], p[k]); hi[k] = std::max(hi[k], p[k]); }
  }
  double area() const {
    if (lo[0] > hi[0]) return 0;
    const double dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    return 2 * (dx * dy + dy * dz + dz * dx);
  }
};

struct Builder {
  const std::vector<aabb>& boxes;
  std::vector<uint32_t>& order;
  std::vector<ssolh_bvh_node>& nodes;
  std::vector<float> cent;  // [n][3]
  unsigned max_leaf;
  unsigned max_depth = 0;

  Builder(const std::vector<aabb>& b, std::vector<uint32_t>& o, std::vector<ssolh_bvh_node>& n,
          unsigned ml)
      : boxes(b), order(o), nodes(n), max_leaf(ml) {}

  Box bounds(size_t first, size_t count) const {
    Box b;
    b.reset();
    for (size_t i = first; i < first + count; ++i) {
      Box x;
      memcpy(x.lo, boxes[order[i]].lo, sizeof(x.lo));
      memcpy(x.hi, boxes[order[i]].hi, sizeof(x.hi));
      b.grow(x);
    }
    return b;
  }

  // split [first, first+count) and return the split point
  size_t split(size_t first, size_t count) {
    Box cb;
    cb.reset();
    for (size_t i = first; i < first + count; ++i) cb.grow(&cent[3 * order[i]]);
    double best = DBL_MAX;
    int best_axis = -1, best_bin = 0;
    for (int axis = 0; axis < 3; ++axis) {
      const float ext = cb.hi[axis] - cb.lo[axis];
      if (!(ext > 0)) continue;
      Box bb[NBINS];
      size_t bc[NBINS] = {0};
      for (auto& b : bb) b.reset();
      for (size_t i = first; i < first + count; ++i) {
        const uint32_t id = order[i];
        int bi = (int)((cent[3 * id + axis] - cb.lo[axis]) / ext * NBINS);
        bi = std::min(std::max(bi, 0), NBINS - 1);
        Box x;
        memcpy(x.lo, boxes[id].lo, sizeof(x.lo));
        memcpy(x.hi, boxes[id].hi, sizeof(x.hi));
        bb[bi].grow(x);
        ++bc[bi];
      }
      Box left[NBINS], right[NBINS];
      size_t nl[NBINS], nr[NBINS];
      Box acc;
      acc.reset();
      size_t n = 0;
      for (int i = 0; i < NBINS; ++i) {
        if (bc[i]) acc.grow(bb[i]);
        n += bc[i];
        left[i] = acc;
        nl[i] = n;
      }
      acc.reset();
      n = 0;
      for (int i = NBINS - 1; i >= 0; --i) {
        if (bc[i]) acc.grow(bb[i]);
        n += bc[i];
        right[i] = acc;
        nr[i] = n;
      }
      for (int i = 0; i < NBINS - 1; ++i) {
        if (!nl[i] || !nr[i + 1]) continue;
        const double cost = left[i].area() * (double)nl[i] + right[i + 1].area() * (double)nr[i + 1];
        if (cost < best) { best = cost; best_axis = axis; best_bin = i; }
      }
    }
    if (best_axis < 0) return first + count / 2;
    const float ext = cb.hi[best_axis] - cb.lo[best_axis];
    auto mid = std::partition(order.begin() + (long)first, order.begin() + (long)(first + count),
                              [&](uint32_t id) {
                                int bi = (int)((cent[3 * id + best_axis] - cb.lo[best_axis]) / ext * NBINS);
                                bi = std::min(std::max(bi, 0), NBINS - 1);
                                return bi <= best_bin;
                              });
    size_t m = (size_t)(mid - order.begin());
    if (m == first || m == first + count) m = first + count / 2;
    return m;
  }

  // Fill child slot c of node `ni` with the subtree over [first, first+count)
  void child(uint32_t ni, int c, size_t first, size_t count, unsigned depth) {
    const Box b = bounds(first, count);
    memcpy(nodes[ni].lo[c], b.lo, sizeof(b.lo));
    memcpy(nodes[ni].hi[c], b.hi, sizeof(b.hi));
    if (count <= max_leaf) {
      nodes[ni].child[c] = (uint32_t)first;
      nodes[ni].count[c] = (uint32_t)count;
      max_depth = std::max(max_depth, depth);
      return;
    }
    const uint32_t id = (uint32_t)nodes.size();
    nodes.push_back(ssolh_bvh_node{});
    nodes[ni].child[c] = id;
    nodes[ni].count[c] = 0;
    inner(id, first, count, depth + 1);
  }

  void inner(uint32_t ni, size_t first, size_t count, unsigned depth) {
    const size_t m = split(first, count);
    child(ni, 0, first, m - first, depth);
    child(ni, 1, m, first + count - m, depth);
  }
};

}  // namespace

uint32_t bvh_build(const std::vector<aabb>& boxes, unsigned max_leaf,
                   std::vector<ssolh_bvh_node>& nodes, std::vector<uint32_t>& order,
                   unsigned* depth) {
  order.resize(boxes.size());
  for (size_t i = 0; i < boxes.size(); ++i) order[i] = (uint32_t)i;
  Builder b(boxes, order, nodes, max_leaf);
  b.cent.resize(boxes.size() * 3);
  for (size_t i = 0; i < boxes.size(); ++i)
    for (int k = 0; k < 3; ++k) b.cent[3 * i + k] = 0.5f * (boxes[i].lo[k] + boxes[i].hi[k]);
  const uint32_t root = (uint32_t)nodes.size();
  nodes.push_back(ssolh_bvh_node{});
  if (boxes.size() <= max_leaf) {
    // single leaf in child 0, child 1 absent
    b.child(root, 0, 0, boxes.size(), 1);
    nodes[root].count[1] = SSOLH_BVH_EMPTY;
    nodes[root].child[1] = 0;
    for (int k = 0; k < 3; ++k) { nodes[root].lo[1][k] = 1; nodes[root].hi[1][k] = -1; }
    if (boxes.empty()) nodes[root].count[0] = SSOLH_BVH_EMPTY;
  } else {
    b.inner(root, 0, boxes.size(), 1);
  }
  if (depth) *depth = b.max_depth;
  return root;
}

}  // namespace ssol
