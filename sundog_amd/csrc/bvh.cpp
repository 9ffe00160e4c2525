// Deterministic binned-SAH build + collapse to 4-wide nodes.
#include "bvh.h"

#include <xmmintrin.h>

#include <algorithm>
#include <stdexcept>
#include <thread>

namespace sdk {

namespace {

constexpr int kBins = 32;

struct BNode {
  Aabb box;
  int left = -1, right = -1;
  int first = 0, count = 0;  // leaf when left < 0
};

// Binary binned-SAH builder. Each node bins its primitives on all three
// axes in one pass, evaluates the splits axis by axis and partitions its
// index range stably through a shared scratch array (concurrent subtrees own
// disjoint ranges of both). The right subtrees of the top kParDepth levels
// that hold at least kParMin primitives are built on their own threads into
// their own node lists and spliced in after the left one; the split of every
// node is the same computation either way, so the tree is the same.
constexpr int kParDepth = 4;
constexpr int kParMin = 512;

struct Builder {
  const std::vector<Aabb>& prims;
  std::vector<vec3> centers;
  std::vector<uint32_t> idx, scratch;
  int maxLeaf;

  Builder(const std::vector<Aabb>& p, int ml) : prims(p), maxLeaf(ml) {
    centers.resize(p.size());
    idx.resize(p.size());
    scratch.resize(p.size());
    for (size_t i = 0; i < p.size(); i++) {
      centers[i] = p[i].center();
      idx[i] = (uint32_t)i;
    }
  }

  // Builds the subtree over idx[first, first + count) into `out` in
  // preorder; child links index `out`.
  int build(std::vector<BNode>& out, int first, int count, int depth) {
    int id = (int)out.size();
    out.emplace_back();
    Aabb box, cbox;
    for (int i = first; i < first + count; i++) {
      box.grow(prims[idx[i]]);
      cbox.grow(centers[idx[i]]);
    }
    out[id].box = box;
    if (count <= maxLeaf) {
      out[id].first = first;
      out[id].count = count;
      return id;
    }
    float lo[3], scale[3];
    bool live[3];
    for (int axis = 0; axis < 3; axis++) {
      lo[axis] = comp(cbox.lo, axis);
      const float hi = comp(cbox.hi, axis);
      live[axis] = hi > lo[axis];
      scale[axis] = live[axis] ? kBins / (hi - lo[axis]) : 0.0f;
    }
    Aabb bb[3][kBins];
    int bc[3][kBins] = {};
    for (int i = first; i < first + count; i++) {
      const uint32_t p = idx[i];
      const vec3 c = centers[p];
      const float cc[3] = {c.x, c.y, c.z};
      for (int axis = 0; axis < 3; axis++) {
        if (!live[axis]) continue;
        const int b = std::min(kBins - 1, (int)((cc[axis] - lo[axis]) * scale[axis]));
        bb[axis][b].grow(prims[p]);
        bc[axis][b]++;
      }
    }
    // Binned SAH over the centroid bounds, axes in order, strict improvement.
    // Every split between two occupied bins has the same two sides, so only
    // the split just after each occupied bin is costed (the first of a run
    // of equal costs, which the strict comparison keeps); empty bins leave
    // the accumulated boxes unchanged.
    int bestAxis = -1, bestSplit = -1;
    float bestCost = 3e38f;
    for (int axis = 0; axis < 3; axis++) {
      if (!live[axis]) continue;
      int occ[kBins], m = 0;
      for (int b = 0; b < kBins; b++)
        if (bc[axis][b] > 0) occ[m++] = b;
      float rightArea[kBins];
      int rightCount[kBins];
      Aabb acc;
      int n = 0;
      for (int k = m - 1; k > 0; k--) {  // sides from occupied bin occ[k] up
        acc.grow(bb[axis][occ[k]]);
        n += bc[axis][occ[k]];
        rightArea[k] = acc.area();
        rightCount[k] = n;
      }
      Aabb lacc;
      int ln = 0;
      for (int k = 0; k + 1 < m; k++) {
        lacc.grow(bb[axis][occ[k]]);
        ln += bc[axis][occ[k]];
        const float cost = lacc.area() * ln + rightArea[k + 1] * rightCount[k + 1];
        if (cost < bestCost) {
          bestCost = cost;
          bestAxis = axis;
          bestSplit = occ[k];
        }
      }
    }
    int mid;
    if (bestAxis < 0) {
      mid = first + count / 2;  // coincident centroids: split by index
    } else {
      // stable partition: left-side indices compact in place, right-side
      // ones go through scratch
      int nl = 0, nr = 0;
      for (int i = first; i < first + count; i++) {
        const uint32_t p = idx[i];
        const int b = std::min(kBins - 1, (int)((comp(centers[p], bestAxis) - lo[bestAxis]) * scale[bestAxis]));
        if (b <= bestSplit) idx[first + nl++] = p;
        else scratch[first + nr++] = p;
      }
      std::copy(scratch.begin() + first, scratch.begin() + first + nr, idx.begin() + first + nl);
      mid = first + nl;
      if (mid == first || mid == first + count) mid = first + count / 2;
    }
    int l, r;
    const int nRight = first + count - mid;
    if (depth < kParDepth && nRight >= kParMin) {
      std::vector<BNode> right;
      right.reserve((size_t)nRight * 2);
      std::thread t([&] { build(right, mid, nRight, depth + 1); });
      l = build(out, first, mid - first, depth + 1);
      t.join();
      const int base = (int)out.size();
      for (BNode& nd : right) {
        if (nd.left >= 0) {
          nd.left += base;
          nd.right += base;
        }
        out.push_back(nd);
      }
      r = base;
    } else {
      l = build(out, first, mid - first, depth + 1);
      r = build(out, mid, nRight, depth + 1);
    }
    out[id].left = l;
    out[id].right = r;
    return id;
  }
};

constexpr uint32_t kLeafFlag = 0x80000000u;

uint32_t leafRef(int first, int count) {
  if (count < 1 || count > 16 || first < 0 || first >= (1 << 27))
    throw std::runtime_error("bvh: leaf does not fit the packed reference");
  return kLeafFlag | ((uint32_t)(count - 1) << 27) | (uint32_t)first;
}

struct Collapser {
  const std::vector<BNode>& bn;
  Bvh4& out;
  float rootArea;

  // Returns the stack bound of the subtree rooted at the 4-wide node `id`.
  int emit(int bid, int id, int depth) {
    out.depth = std::max(out.depth, depth);
    int kids[4] = {bn[bid].left, bn[bid].right, -1, -1};
    int k = 2;
    while (k < 4) {
      int best = -1;
      float bestA = -1.0f;
      for (int i = 0; i < k; i++) {
        const BNode& c = bn[kids[i]];
        if (c.left >= 0 && c.box.area() > bestA) {
          bestA = c.box.area();
          best = i;
        }
      }
      if (best < 0) break;
      int b = kids[best];
      kids[best] = bn[b].left;
      kids[k++] = bn[b].right;
    }
    BvhNode4 nd;
    for (int i = 0; i < 4; i++) {
      nd.lox[i] = nd.loy[i] = nd.loz[i] = 3e38f;
      nd.hix[i] = nd.hiy[i] = nd.hiz[i] = 3e38f;
      nd.child[i] = (int32_t)0xFFFFFFFFu;
      nd.count[i] = -1;
    }
    out.sahCost += rootArea > 0 ? bn[bid].box.area() / rootArea : 1.0;
    out.nodes[id] = nd;
    int pendingInner[4];
    int nInner = 0;
    for (int i = 0; i < k; i++) {
      const BNode& c = bn[kids[i]];
      BvhNode4& n = out.nodes[id];
      n.lox[i] = c.box.lo.x; n.loy[i] = c.box.lo.y; n.loz[i] = c.box.lo.z;
      n.hix[i] = c.box.hi.x; n.hiy[i] = c.box.hi.y; n.hiz[i] = c.box.hi.z;
      if (c.left < 0) {
        n.child[i] = (int32_t)leafRef(c.first, c.count);
        n.count[i] = c.count;
        out.sahCost += rootArea > 0 ? c.box.area() / rootArea * c.count : c.count;
      } else {
        int nid = (int)out.nodes.size();
        out.nodes.emplace_back();
        n.child[i] = nid;
        n.count[i] = 0;
        pendingInner[nInner++] = i;
      }
    }
    int sub = 0;
    for (int j = 0; j < nInner; j++) {
      int i = pendingInner[j];
      int nid = out.nodes[id].child[i];
      sub = std::max(sub, emit(kids[i], nid, depth + 1));
    }
    return (k - 1) + sub;
  }
};

}  // namespace

Bvh4 buildBvh4(const std::vector<Aabb>& prims, int maxLeaf) {
  if (prims.empty()) throw std::runtime_error("bvh: no primitives");
  if (maxLeaf < 1 || maxLeaf > 16) throw std::runtime_error("bvh: bad leaf size");
  Builder b(prims, maxLeaf);
  std::vector<BNode> nodes;
  nodes.reserve(prims.size() * 2);
  b.build(nodes, 0, (int)prims.size(), 0);
  Bvh4 out;
  out.bounds = nodes[0].box;
  out.primOrder = b.idx;
  if (nodes[0].left < 0) {
    // Single leaf: wrap it in one 4-wide node with one used slot.
    BvhNode4 nd;
    for (int i = 0; i < 4; i++) {
      nd.lox[i] = nd.loy[i] = nd.loz[i] = 3e38f;
      nd.hix[i] = nd.hiy[i] = nd.hiz[i] = 3e38f;
      nd.child[i] = (int32_t)0xFFFFFFFFu;
      nd.count[i] = -1;
    }
    const BNode& c = nodes[0];
    nd.lox[0] = c.box.lo.x; nd.loy[0] = c.box.lo.y; nd.loz[0] = c.box.lo.z;
    nd.hix[0] = c.box.hi.x; nd.hiy[0] = c.box.hi.y; nd.hiz[0] = c.box.hi.z;
    nd.child[0] = (int32_t)leafRef(c.first, c.count);
    nd.count[0] = c.count;
    out.nodes.push_back(nd);
    out.stackBound = 0;
    out.depth = 1;
    out.sahCost = 1.0 + c.count;
    return out;
  }
  // 4-wide nodes never outnumber binary inner nodes; reserving keeps the
  // node references taken inside emit() valid.
  out.nodes.reserve(nodes.size());
  out.nodes.emplace_back();
  Collapser col{nodes, out, nodes[0].box.area()};
  out.stackBound = col.emit(0, 0, 1);
  return out;
}

SoaPoints toSoa(const std::vector<vec3>& pts) {
  SoaPoints s;
  const size_t n = (pts.size() + 3) / 4 * 4;
  s.x.resize(n);
  s.y.resize(n);
  s.z.resize(n);
  for (size_t i = 0; i < n; i++) {
    const vec3& p = pts[std::min(i, pts.size() - 1)];
    s.x[i] = p.x;
    s.y[i] = p.y;
    s.z[i] = p.z;
  }
  return s;
}

// Four points per step with SSE: ((m0 x + m1 y) + m2 z) + m3 per row, the
// order of Affine::point; min and max are exact, so lane order does not
// matter.
Aabb transformedBounds(const Affine& a, const SoaPoints& pts) {
  __m128 m[12];
  for (int k = 0; k < 12; k++) m[k] = _mm_set1_ps(a.m[k]);
  __m128 lo[3], hi[3];
  for (int r = 0; r < 3; r++) {
    lo[r] = _mm_set1_ps(3e38f);
    hi[r] = _mm_set1_ps(-3e38f);
  }
  const size_t n = pts.x.size();
  for (size_t i = 0; i < n; i += 4) {
    const __m128 x = _mm_loadu_ps(&pts.x[i]), y = _mm_loadu_ps(&pts.y[i]), z = _mm_loadu_ps(&pts.z[i]);
    for (int r = 0; r < 3; r++) {
      const __m128 t = _mm_add_ps(
          _mm_add_ps(_mm_add_ps(_mm_mul_ps(m[4 * r], x), _mm_mul_ps(m[4 * r + 1], y)), _mm_mul_ps(m[4 * r + 2], z)),
          m[4 * r + 3]);
      lo[r] = _mm_min_ps(lo[r], t);
      hi[r] = _mm_max_ps(hi[r], t);
    }
  }
  float l[3][4], h[3][4];
  for (int r = 0; r < 3; r++) {
    _mm_storeu_ps(l[r], lo[r]);
    _mm_storeu_ps(h[r], hi[r]);
  }
  Aabb b;
  for (int k = 0; k < 4; k++) {
    b.grow(vec3{l[0][k], l[1][k], l[2][k]});
    b.grow(vec3{h[0][k], h[1][k], h[2][k]});
  }
  return b;
}

}  // namespace sdk
