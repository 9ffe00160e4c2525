// Two-level BVH construction (host, deterministic).
//
// Replaces the driver-side acceleration builds of the reference
// (src/accel.cpp:8-121: one GAS per quadric kind / mesh, one IAS over all
// objects). A binned-SAH binary tree is built top-down, then collapsed into
// 4-wide nodes whose child boxes are stored structure-of-arrays in one 128-byte
// record, so a traversal step is eight 16-byte loads from one cache line pair.
// The same input always produces the same bytes (no threads, stable
// partitions, fixed tie-breaks).
#pragma once

#include <stdint.h>

#include <vector>

#include "sd_math.h"

namespace sdk {

struct Aabb {
  vec3 lo{3e38f, 3e38f, 3e38f}, hi{-3e38f, -3e38f, -3e38f};
  void grow(vec3 p) {
    lo = vmin(lo, p);
    hi = vmax(hi, p);
  }
  void grow(const Aabb& b) {
    lo = vmin(lo, b.lo);
    hi = vmax(hi, b.hi);
  }
  bool empty() const { return lo.x > hi.x; }
  float area() const {
    if (empty()) return 0.0f;
    vec3 d = hi - lo;
    return 2.0f * (d.x * d.y + d.y * d.z + d.z * d.x);
  }
  vec3 center() const { return (lo + hi) * 0.5f; }
};

// Device node: 4 children, boxes SoA. count: 0 = inner (child = node index),
// > 0 = leaf with `count` primitives starting at `child`, -1 = empty slot.
struct alignas(16) BvhNode4 {
  float lox[4], hix[4], loy[4], hiy[4], loz[4], hiz[4];
  int32_t child[4];
  int32_t count[4];
};
static_assert(sizeof(BvhNode4) == 128, "node layout");

struct Bvh4 {
  std::vector<BvhNode4> nodes;    // root at index 0
  std::vector<uint32_t> primOrder; // leaf order -> input primitive index
  Aabb bounds;
  int stackBound = 0;   // max entries a depth-first traversal can hold
  int depth = 0;
  double sahCost = 0.0; // expected cost: 1 per node visit, 1 per primitive test
};

// prims: one box per primitive. maxLeaf: primitives per leaf (>= 1).
Bvh4 buildBvh4(const std::vector<Aabb>& prims, int maxLeaf);

// Points in structure-of-arrays form, padded to a multiple of 4 by repeating
// the last point.
struct SoaPoints {
  std::vector<float> x, y, z;
};
SoaPoints toSoa(const std::vector<vec3>& pts);

// Bounds of the points under `a`, each coordinate computed as Affine::point
// computes it (so the box equals the one grown from a.point(p)).
Aabb transformedBounds(const Affine& a, const SoaPoints& pts);

}  // namespace sdk
