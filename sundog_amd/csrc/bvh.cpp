// Deterministic binned-SAH build + collapse to 4-wide nodes.
#include "bvh.h"

#include <algorithm>
#include <stdexcept>

namespace sdk {

namespace {

constexpr int kBins = 32;

struct BNode {
  Aabb box;
  int left = -1, right = -1;
  int first = 0, count = 0;  // leaf when left < 0
};

struct Builder {
  const std::vector<Aabb>& prims;
  std::vector<vec3> centers;
  std::vector<uint32_t> idx;
  std::vector<BNode> nodes;
  int maxLeaf;

  Builder(const std::vector<Aabb>& p, int ml) : prims(p), maxLeaf(ml) {
    centers.resize(p.size());
    idx.resize(p.size());
    for (size_t i = 0; i < p.size(); i++) {
      centers[i] = p[i].center();
      idx[i] = (uint32_t)i;
    }
  }

  int build(int first, int count) {
    int id = (int)nodes.size();
    nodes.emplace_back();
    Aabb box, cbox;
    for (int i = first; i < first + count; i++) {
      box.grow(prims[idx[i]]);
      cbox.grow(centers[idx[i]]);
    }
    nodes[id].box = box;
    if (count <= maxLeaf) {
      nodes[id].first = first;
      nodes[id].count = count;
      return id;
    }
    // Binned SAH over the centroid bounds, all three axes.
    int bestAxis = -1, bestSplit = -1;
    float bestCost = 3e38f;
    for (int axis = 0; axis < 3; axis++) {
      float lo = comp(cbox.lo, axis), hi = comp(cbox.hi, axis);
      if (!(hi > lo)) continue;
      float scale = kBins / (hi - lo);
      Aabb bb[kBins];
      int bc[kBins] = {0};
      for (int i = first; i < first + count; i++) {
        int b = std::min(kBins - 1, (int)((comp(centers[idx[i]], axis) - lo) * scale));
        bb[b].grow(prims[idx[i]]);
        bc[b]++;
      }
      float rightArea[kBins];
      int rightCount[kBins];
      Aabb acc;
      int n = 0;
      for (int b = kBins - 1; b > 0; b--) {
        acc.grow(bb[b]);
        n += bc[b];
        rightArea[b] = acc.area();
        rightCount[b] = n;
      }
      Aabb lacc;
      int ln = 0;
      for (int b = 0; b < kBins - 1; b++) {
        lacc.grow(bb[b]);
        ln += bc[b];
        int rn = rightCount[b + 1];
        if (ln == 0 || rn == 0) continue;
        float cost = lacc.area() * ln + rightArea[b + 1] * rn;
        if (cost < bestCost) {
          bestCost = cost;
          bestAxis = axis;
          bestSplit = b;
        }
      }
    }
    int mid;
    if (bestAxis < 0) {
      mid = first + count / 2;  // coincident centroids: split by index
    } else {
      float lo = comp(cbox.lo, bestAxis), hi = comp(cbox.hi, bestAxis);
      float scale = kBins / (hi - lo);
      auto it = std::stable_partition(idx.begin() + first, idx.begin() + first + count,
                                      [&](uint32_t p) {
                                        int b = std::min(kBins - 1,
                                                         (int)((comp(centers[p], bestAxis) - lo) * scale));
                                        return b <= bestSplit;
                                      });
      mid = (int)(it - idx.begin());
      if (mid == first || mid == first + count) mid = first + count / 2;
    }
    int l = build(first, mid - first);
    int r = build(mid, first + count - mid);
    nodes[id].left = l;
    nodes[id].right = r;
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
  b.nodes.reserve(prims.size() * 2);
  b.build(0, (int)prims.size());
  Bvh4 out;
  out.bounds = b.nodes[0].box;
  out.primOrder = b.idx;
  if (b.nodes[0].left < 0) {
    // Single leaf: wrap it in one 4-wide node with one used slot.
    BvhNode4 nd;
    for (int i = 0; i < 4; i++) {
      nd.lox[i] = nd.loy[i] = nd.loz[i] = 3e38f;
      nd.hix[i] = nd.hiy[i] = nd.hiz[i] = 3e38f;
      nd.child[i] = (int32_t)0xFFFFFFFFu;
      nd.count[i] = -1;
    }
    const BNode& c = b.nodes[0];
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
  out.nodes.reserve(b.nodes.size());
  out.nodes.emplace_back();
  Collapser col{b.nodes, out, b.nodes[0].box.area()};
  out.stackBound = col.emit(0, 0, 1);
  return out;
}

}  // namespace sdk
