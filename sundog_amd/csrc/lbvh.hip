// GPU build of the two-level BVH: a linear BVH (Karras, "Maximizing
// Parallelism in the Construction of BVHs, Octrees, and k-d Trees", HPG
// 2012) over 30-bit Morton codes of primitive centroids, refit bottom-up,
// then collapsed on the device into the same 4-wide 128-byte nodes the host
// SAH build emits (bvh.h). It replaces the driver-side builds of the
// reference (src/accel.cpp:8-121) where build time matters more than tree
// quality; the host SAH build stays the default.
//
// Collapse: binary internal nodes at even depth become 4-wide nodes whose
// slots are their grandchildren (or a child directly when that child is a
// leaf); a subtree of at most maxLeaf primitives (a contiguous range of the
// sorted order) becomes one leaf.
#include <hip/hip_runtime.h>

#include <cstring>

#include <algorithm>
#include <stdexcept>
#include <vector>

#include "session.h"

namespace sdk {

namespace {

constexpr int kLbvhBlock = 256;
constexpr uint32_t kLeafBit = 0x80000000u;  // binary child reference: leaf (sorted position)

struct BinNode {
  uint32_t left, right;  // child: internal index, or kLeafBit | sorted position
  uint32_t parent;       // internal index; 0xFFFFFFFF at the root
  uint32_t lo, hi;       // covered range of sorted positions
  float4 boxLo, boxHi;   // refit bounds
};

__device__ inline uint32_t expandBits(uint32_t v) {
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

// order-preserving float <-> uint mapping for atomic min/max
__device__ inline uint32_t fkey(float f) {
  uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ inline float funkey(uint32_t k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}

// Centroid bounds: per-block reduction, then atomics on ordered keys.
__global__ void k_centroid_bounds(const float4* __restrict__ boxes, int n, uint32_t* __restrict__ bounds) {
  __shared__ float s[6][kLbvhBlock];
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  float c[3] = {3e38f, 3e38f, 3e38f}, d[3] = {-3e38f, -3e38f, -3e38f};
  if (i < n) {
    float4 lo = boxes[2 * i], hi = boxes[2 * i + 1];
    c[0] = d[0] = 0.5f * (lo.x + hi.x);
    c[1] = d[1] = 0.5f * (lo.y + hi.y);
    c[2] = d[2] = 0.5f * (lo.z + hi.z);
  }
  for (int a = 0; a < 3; a++) {
    s[a][threadIdx.x] = c[a];
    s[3 + a][threadIdx.x] = d[a];
  }
  __syncthreads();
  for (int o = kLbvhBlock / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o)
      for (int a = 0; a < 3; a++) {
        s[a][threadIdx.x] = fminf(s[a][threadIdx.x], s[a][threadIdx.x + o]);
        s[3 + a][threadIdx.x] = fmaxf(s[3 + a][threadIdx.x], s[3 + a][threadIdx.x + o]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 3) atomicMin(&bounds[threadIdx.x], fkey(s[threadIdx.x][0]));
  else if (threadIdx.x < 6) atomicMax(&bounds[threadIdx.x], fkey(s[threadIdx.x][0]));
}

__global__ void k_bounds_init(uint32_t* b) {
  if (threadIdx.x < 3) b[threadIdx.x] = 0xFFFFFFFFu;
  else if (threadIdx.x < 6) b[threadIdx.x] = 0u;
}

// 64-bit keys: 30-bit Morton code of the centroid, then the primitive index
// (unique keys, so the order and the tree are deterministic).
__global__ void k_morton(const float4* __restrict__ boxes, int n, const uint32_t* __restrict__ bounds,
                         uint64_t* __restrict__ keys) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float lo[3] = {funkey(bounds[0]), funkey(bounds[1]), funkey(bounds[2])};
  float hi[3] = {funkey(bounds[3]), funkey(bounds[4]), funkey(bounds[5])};
  float4 a = boxes[2 * i], b = boxes[2 * i + 1];
  float c[3] = {0.5f * (a.x + b.x), 0.5f * (a.y + b.y), 0.5f * (a.z + b.z)};
  uint32_t q[3];
  for (int k = 0; k < 3; k++) {
    float ext = hi[k] - lo[k];
    float u = ext > 0.0f ? (c[k] - lo[k]) / ext : 0.5f;
    q[k] = (uint32_t)fminf(fmaxf(u * 1024.0f, 0.0f), 1023.0f);
  }
  uint32_t m = (expandBits(q[0]) << 2) | (expandBits(q[1]) << 1) | expandBits(q[2]);
  keys[i] = ((uint64_t)m << 32) | (uint32_t)i;
}

__device__ inline int delta(const uint64_t* keys, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  return __clzll((long long)(keys[i] ^ keys[j]));
}

// Karras' binary radix tree: internal node i covers [min(i,j), max(i,j)].
__global__ void k_karras(const uint64_t* __restrict__ keys, int n, BinNode* __restrict__ nodes,
                         uint32_t* __restrict__ leafParent) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
  int dmin = delta(keys, n, i, i - d);
  int lmax = 2;
  while (delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
  int l = 0;
  for (int t = lmax >> 1; t >= 1; t >>= 1)
    if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
  int j = i + l * d;
  int dnode = delta(keys, n, i, j);
  int s = 0;
  for (int t = (l + 1) >> 1;; t = (t + 1) >> 1) {
    if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
    if (t == 1) break;
  }
  int gamma = i + s * d + min(d, 0);
  int lo = min(i, j), hi = max(i, j);
  BinNode& nd = nodes[i];
  nd.lo = (uint32_t)lo;
  nd.hi = (uint32_t)hi;
  nd.left = lo == gamma ? (kLeafBit | (uint32_t)gamma) : (uint32_t)gamma;
  nd.right = hi == gamma + 1 ? (kLeafBit | (uint32_t)(gamma + 1)) : (uint32_t)(gamma + 1);
  if (nd.left & kLeafBit) leafParent[gamma] = (uint32_t)i;
  else nodes[gamma].parent = (uint32_t)i;
  if (nd.right & kLeafBit) leafParent[gamma + 1] = (uint32_t)i;
  else nodes[gamma + 1].parent = (uint32_t)i;
  if (i == 0) nd.parent = 0xFFFFFFFFu;
}

__device__ inline void childBox(const BinNode* nodes, const float4* boxes, const uint64_t* keys, uint32_t c,
                                float4& lo, float4& hi) {
  if (c & kLeafBit) {
    uint32_t prim = (uint32_t)keys[c & ~kLeafBit];
    lo = boxes[2 * prim];
    hi = boxes[2 * prim + 1];
  } else {
    lo = nodes[c].boxLo;
    hi = nodes[c].boxHi;
  }
}

// The refit reads boxes another thread wrote moments earlier, on another CU
// whose stores are not visible through this CU's L1: read them volatile
// (coherent loads that bypass it).
__device__ inline void childBoxCoherent(const BinNode* nodes, const float4* boxes, const uint64_t* keys, uint32_t c,
                                        float4& lo, float4& hi) {
  if (c & kLeafBit) {
    uint32_t prim = (uint32_t)keys[c & ~kLeafBit];
    lo = boxes[2 * prim];
    hi = boxes[2 * prim + 1];
  } else {
    const volatile float* a = reinterpret_cast<const volatile float*>(&nodes[c].boxLo);
    const volatile float* b = reinterpret_cast<const volatile float*>(&nodes[c].boxHi);
    lo = make_float4(a[0], a[1], a[2], 0.0f);
    hi = make_float4(b[0], b[1], b[2], 0.0f);
  }
}

// Bottom-up refit: the second child to arrive at a parent computes it.
__global__ void k_refit(BinNode* nodes, const float4* __restrict__ boxes, const uint64_t* __restrict__ keys, int n,
                        const uint32_t* __restrict__ leafParent, uint32_t* __restrict__ visits) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t p = leafParent[i];
  while (p != 0xFFFFFFFFu) {
    __threadfence();
    if (atomicAdd(&visits[p], 1u) == 0) return;  // the sibling finishes this node
    __threadfence();
    float4 al, ah, bl, bh;
    childBoxCoherent(nodes, boxes, keys, nodes[p].left, al, ah);
    childBoxCoherent(nodes, boxes, keys, nodes[p].right, bl, bh);
    nodes[p].boxLo = make_float4(fminf(al.x, bl.x), fminf(al.y, bl.y), fminf(al.z, bl.z), 0.0f);
    nodes[p].boxHi = make_float4(fmaxf(ah.x, bh.x), fmaxf(ah.y, bh.y), fmaxf(ah.z, bh.z), 0.0f);
    p = nodes[p].parent;
  }
}

__device__ inline bool collapsed(const BinNode& nd, int maxLeaf) { return (int)(nd.hi - nd.lo + 1) <= maxLeaf; }

// An internal node becomes a 4-wide node when it sits at even depth and is
// not inside (or itself) a collapsed leaf range.
__global__ void k_classify(const BinNode* __restrict__ nodes, int numInternal, int maxLeaf,
                           uint32_t* __restrict__ isNode4) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= numInternal) return;
  int depth = 0;
  bool inLeaf = collapsed(nodes[i], maxLeaf) && i != 0;
  for (uint32_t p = nodes[i].parent; p != 0xFFFFFFFFu; p = nodes[p].parent) {
    depth++;
    if (collapsed(nodes[p], maxLeaf)) inLeaf = true;
  }
  isNode4[i] = (!inLeaf && (depth & 1) == 0) ? 1u : 0u;
}

// Slot of a 4-wide node for binary child c of a node at even depth.
struct Slot {
  uint32_t ref;
  int32_t count;
  float4 lo, hi;
};

__device__ inline Slot makeSlot(const BinNode* nodes, const float4* boxes, const uint64_t* keys,
                                const uint32_t* node4Index, int maxLeaf, bool refIsPrim, uint32_t nodeBase,
                                uint32_t primBase, uint32_t c) {
  Slot s;
  childBox(nodes, boxes, keys, c, s.lo, s.hi);
  uint32_t first, cnt;
  if (c & kLeafBit) {
    first = c & ~kLeafBit;
    cnt = 1;
  } else if (collapsed(nodes[c], maxLeaf)) {
    first = nodes[c].lo;
    cnt = nodes[c].hi - nodes[c].lo + 1;
  } else {
    s.ref = nodeBase + node4Index[c];
    s.count = 0;
    return s;
  }
  uint32_t target = refIsPrim ? (uint32_t)keys[first] : primBase + first;
  s.ref = kRefLeaf | ((cnt - 1) << 27) | target;
  s.count = (int32_t)cnt;
  return s;
}

__global__ void k_emit(const BinNode* __restrict__ nodes, int numInternal, const float4* __restrict__ boxes,
                       const uint64_t* __restrict__ keys, const uint32_t* __restrict__ isNode4,
                       const uint32_t* __restrict__ node4Index, int maxLeaf, int refIsPrim, uint32_t nodeBase,
                       uint32_t primBase, BvhNode4* __restrict__ out, int32_t* __restrict__ fanout) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= numInternal || !isNode4[i]) return;
  uint32_t kids[4];
  int k = 0;
  const BinNode& nd = nodes[i];
  uint32_t ch[2] = {nd.left, nd.right};
  for (int a = 0; a < 2; a++) {
    uint32_t c = ch[a];
    if ((c & kLeafBit) || collapsed(nodes[c], maxLeaf)) {
      kids[k++] = c;
    } else {
      kids[k++] = nodes[c].left;
      kids[k++] = nodes[c].right;
    }
  }
  BvhNode4 o;
  for (int s = 0; s < 4; s++) {
    if (s < k) {
      Slot sl = makeSlot(nodes, boxes, keys, node4Index, maxLeaf, refIsPrim != 0, nodeBase, primBase, kids[s]);
      o.lox[s] = sl.lo.x; o.loy[s] = sl.lo.y; o.loz[s] = sl.lo.z;
      o.hix[s] = sl.hi.x; o.hiy[s] = sl.hi.y; o.hiz[s] = sl.hi.z;
      o.child[s] = (int32_t)sl.ref;
      o.count[s] = sl.count;
    } else {
      o.lox[s] = o.loy[s] = o.loz[s] = o.hix[s] = o.hiy[s] = o.hiz[s] = 3e38f;
      o.child[s] = (int32_t)kRefEmpty;
      o.count[s] = -1;
    }
  }
  out[node4Index[i]] = o;
  fanout[node4Index[i]] = k;
}

// Worst-case traversal stack: max over 4-wide root-to-node paths of the sum
// of (fan-out - 1), the entries a depth-first traversal can hold.
__global__ void k_stack_bound(const BinNode* __restrict__ nodes, int numInternal, const uint32_t* __restrict__ isNode4,
                              const uint32_t* __restrict__ node4Index, const int32_t* __restrict__ fanout,
                              int32_t* __restrict__ bound) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= numInternal || !isNode4[i]) return;
  int sum = 0;
  for (uint32_t p = (uint32_t)i; p != 0xFFFFFFFFu; p = nodes[p].parent)
    if (isNode4[p]) sum += fanout[node4Index[p]] - 1;
  atomicMax(bound, sum);
}

// Triangle pool in leaf order: 3 float4 per triangle, v0.w = triangle id.
__global__ void k_emit_tris(const uint64_t* __restrict__ keys, int n, const float* __restrict__ pos,
                            const uint32_t* __restrict__ idx, float4* __restrict__ tris) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  uint32_t t = (uint32_t)keys[p];
  uint32_t i0 = idx[3 * t], i1 = idx[3 * t + 1], i2 = idx[3 * t + 2];
  tris[3 * p] = make_float4(pos[3 * i0], pos[3 * i0 + 1], pos[3 * i0 + 2], __uint_as_float(t));
  tris[3 * p + 1] = make_float4(pos[3 * i1], pos[3 * i1 + 1], pos[3 * i1 + 2], 0.0f);
  tris[3 * p + 2] = make_float4(pos[3 * i2], pos[3 * i2 + 1], pos[3 * i2 + 2], 0.0f);
}

__global__ void k_tri_boxes(const float* __restrict__ pos, const uint32_t* __restrict__ idx, int n,
                            float4* __restrict__ boxes) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  float4 lo = make_float4(3e38f, 3e38f, 3e38f, 0.0f), hi = make_float4(-3e38f, -3e38f, -3e38f, 0.0f);
  for (int k = 0; k < 3; k++) {
    uint32_t v = idx[3 * t + k];
    float x = pos[3 * v], y = pos[3 * v + 1], z = pos[3 * v + 2];
    lo.x = fminf(lo.x, x); lo.y = fminf(lo.y, y); lo.z = fminf(lo.z, z);
    hi.x = fmaxf(hi.x, x); hi.y = fmaxf(hi.y, y); hi.z = fmaxf(hi.z, z);
  }
  boxes[2 * t] = lo;
  boxes[2 * t + 1] = hi;
}

// A one-slot 4-wide node holding a single leaf (fewer than two primitives
// or everything in one collapsed range).
__global__ void k_single_leaf(const float4* __restrict__ boxes, const uint64_t* __restrict__ keys, int n,
                              int refIsPrim, uint32_t primBase, BvhNode4* out) {
  if (threadIdx.x != 0) return;
  float4 lo = make_float4(3e38f, 3e38f, 3e38f, 0.0f), hi = make_float4(-3e38f, -3e38f, -3e38f, 0.0f);
  for (int p = 0; p < n; p++) {
    uint32_t prim = (uint32_t)keys[p];
    float4 a = boxes[2 * prim], b = boxes[2 * prim + 1];
    lo.x = fminf(lo.x, a.x); lo.y = fminf(lo.y, a.y); lo.z = fminf(lo.z, a.z);
    hi.x = fmaxf(hi.x, b.x); hi.y = fmaxf(hi.y, b.y); hi.z = fmaxf(hi.z, b.z);
  }
  BvhNode4 o;
  for (int s = 0; s < 4; s++) {
    o.lox[s] = o.loy[s] = o.loz[s] = o.hix[s] = o.hiy[s] = o.hiz[s] = 3e38f;
    o.child[s] = (int32_t)kRefEmpty;
    o.count[s] = -1;
  }
  o.lox[0] = lo.x; o.loy[0] = lo.y; o.loz[0] = lo.z;
  o.hix[0] = hi.x; o.hiy[0] = hi.y; o.hiz[0] = hi.z;
  uint32_t target = refIsPrim ? (uint32_t)keys[0] : primBase;
  o.child[0] = (int32_t)(kRefLeaf | ((uint32_t)(n - 1) << 27) | target);
  o.count[0] = n;
  out[0] = o;
}

inline unsigned grid(size_t n) { return (unsigned)((n + kLbvhBlock - 1) / kLbvhBlock); }

// World box of each mesh instance from its transformed vertices: one block
// per job, a min/max reduction over the mesh's vertices, padded by a few
// ulps of the magnitude as the host pads transformed boxes. Jobs with no
// vertices (quadric instances) keep the box the host wrote.
__global__ __launch_bounds__(256) void k_instance_boxes(const float* __restrict__ pos,
                                                        const InstanceBoxJob* __restrict__ jobs, int n,
                                                        float4* __restrict__ boxes) {
  __shared__ float red[6][256 / 64];
  const InstanceBoxJob& jb = jobs[blockIdx.x];
  if (jb.numVerts <= 0) return;  // uniform per block
  const float* m = jb.xf;
  float lo[3] = {3e38f, 3e38f, 3e38f}, hi[3] = {-3e38f, -3e38f, -3e38f};
  for (int v = threadIdx.x; v < jb.numVerts; v += blockDim.x) {
    const float* q = pos + 3 * ((size_t)jb.posOff + v);
    const float x = q[0], y = q[1], z = q[2];
    const float w[3] = {m[0] * x + m[1] * y + m[2] * z + m[3], m[4] * x + m[5] * y + m[6] * z + m[7],
                        m[8] * x + m[9] * y + m[10] * z + m[11]};
#pragma unroll
    for (int a = 0; a < 3; a++) {
      lo[a] = fminf(lo[a], w[a]);
      hi[a] = fmaxf(hi[a], w[a]);
    }
  }
#pragma unroll
  for (int a = 0; a < 3; a++) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      lo[a] = fminf(lo[a], __shfl_xor(lo[a], off, 64));
      hi[a] = fmaxf(hi[a], __shfl_xor(hi[a], off, 64));
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int a = 0; a < 3; a++) {
      red[a][wave] = lo[a];
      red[3 + a][wave] = hi[a];
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float l[3], h[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
      l[a] = red[a][0];
      h[a] = red[3 + a][0];
      for (int k = 1; k < (int)(blockDim.x / 64); k++) {
        l[a] = fminf(l[a], red[a][k]);
        h[a] = fmaxf(h[a], red[3 + a][k]);
      }
      const float mag = fmaxf(fabsf(l[a]), fabsf(h[a]));
      const float pad = mag * 2e-6f + 1e-7f;
      l[a] -= pad;
      h[a] += pad;
    }
    boxes[2 * (size_t)blockIdx.x] = make_float4(l[0], l[1], l[2], 0.0f);
    boxes[2 * (size_t)blockIdx.x + 1] = make_float4(h[0], h[1], h[2], 0.0f);
  }
}

// ---------------------------------------------------------------------------
// exclusive prefix sum of uint32 (device-wide): 2,048-element blocks scanned
// in LDS, their totals scanned the same way (recursively), then added back
// ---------------------------------------------------------------------------

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;  // consecutive elements per thread
constexpr int kScanChunk = kScanThreads * kScanItems;

__global__ __launch_bounds__(kScanThreads) void k_scan_block(const uint32_t* __restrict__ in,
                                                            uint32_t* __restrict__ out,
                                                            uint32_t* __restrict__ blockSums, int n) {
  __shared__ uint32_t part[kScanThreads];
  const int base = blockIdx.x * kScanChunk + threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  uint32_t sum = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    v[k] = base + k < n ? in[base + k] : 0u;
    sum += v[k];
  }
  part[threadIdx.x] = sum;
  __syncthreads();
  // inclusive scan of the per-thread sums (Hillis-Steele)
  for (int off = 1; off < kScanThreads; off <<= 1) {
    const uint32_t x = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0u;
    __syncthreads();
    part[threadIdx.x] += x;
    __syncthreads();
  }
  uint32_t run = threadIdx.x ? part[threadIdx.x - 1] : 0u;
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
  if (blockSums && threadIdx.x == kScanThreads - 1) blockSums[blockIdx.x] = part[kScanThreads - 1];
}

__global__ void k_scan_add(uint32_t* __restrict__ out, const uint32_t* __restrict__ blockOffsets, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] += blockOffsets[i / kScanChunk];
}

// out[i] = in[0] + ... + in[i - 1]. Temporary buffers go to `pool`, which
// must outlive the queued kernels.
void exclusiveScan(const uint32_t* in, uint32_t* out, int n, std::vector<DevBuf>& pool, hipStream_t s) {
  const int blocks = (n + kScanChunk - 1) / kScanChunk;
  if (blocks <= 1) {
    hipLaunchKernelGGL(k_scan_block, dim3(1), dim3(kScanThreads), 0, s, in, out, (uint32_t*)nullptr, n);
    HIP_OK(hipGetLastError());
    return;
  }
  pool.emplace_back();
  pool.back().alloc((size_t)blocks * 4);
  uint32_t* sums = pool.back().as<uint32_t>();
  pool.emplace_back();
  pool.back().alloc((size_t)blocks * 4);
  uint32_t* offsets = pool.back().as<uint32_t>();
  hipLaunchKernelGGL(k_scan_block, dim3(blocks), dim3(kScanThreads), 0, s, in, out, sums, n);
  HIP_OK(hipGetLastError());
  exclusiveScan(sums, offsets, blocks, pool, s);
  hipLaunchKernelGGL(k_scan_add, dim3((n + 255) / 256), dim3(256), 0, s, out, offsets, n);
  HIP_OK(hipGetLastError());
}

// ---------------------------------------------------------------------------
// LSD radix sort of the 64-bit keys by their Morton code (bits 32-61), four
// stable 8-bit passes. The keys arrive in primitive order with the index in
// the low word, so the stable sort gives exactly the order of the full
// 64-bit keys (code, then index). A pass: per-block digit counts
// (digit-major, counts[d * blocks + b]), their exclusive scan, then a stable
// scatter in which each block walks its 4,096 keys in tiles of 256 (one per
// thread, in order), ranks each key among the equal digits of its wave with
// eight ballots, and places it after the earlier waves and tiles.
// ---------------------------------------------------------------------------

constexpr int kSortThreads = 256;
constexpr int kSortTiles = 16;
constexpr int kSortChunk = kSortThreads * kSortTiles;

__global__ __launch_bounds__(kSortThreads) void k_radix_count(const uint64_t* __restrict__ keys, int n, int shift,
                                                             uint32_t* __restrict__ counts) {
  __shared__ uint32_t hist[256];
  hist[threadIdx.x] = 0u;
  __syncthreads();
  const int begin = blockIdx.x * kSortChunk;
  const int end = min(n, begin + kSortChunk);
  for (int i = begin + (int)threadIdx.x; i < end; i += kSortThreads)
    atomicAdd(&hist[(uint32_t)(keys[i] >> shift) & 0xFFu], 1u);
  __syncthreads();
  counts[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = hist[threadIdx.x];
}

__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(const uint64_t* __restrict__ in,
                                                               uint64_t* __restrict__ out, int n, int shift,
                                                               const uint32_t* __restrict__ offsets) {
  constexpr int kWaves = kSortThreads / 64;
  __shared__ uint32_t running[256];         // next output position of each digit
  __shared__ uint32_t waveHist[kWaves][256];  // this tile's digit counts per wave
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  running[threadIdx.x] = offsets[(size_t)threadIdx.x * gridDim.x + blockIdx.x];
  for (int w = 0; w < kWaves; w++) waveHist[w][threadIdx.x] = 0u;
  __syncthreads();
  const int begin = blockIdx.x * kSortChunk;
  const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));  // the lanes before this one
  for (int t = 0; t < kSortTiles && begin + t * kSortThreads < n; t++) {
    const int i = begin + t * kSortThreads + (int)threadIdx.x;
    const bool valid = i < n;
    const uint64_t key = valid ? in[i] : 0ull;
    const uint32_t d = (uint32_t)(key >> shift) & 0xFFu;
    // the lanes of this wave holding a valid key with the same digit
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      const uint64_t m = __ballot((d >> b) & 1u);
      peers &= ((d >> b) & 1u) ? m : ~m;
    }
    const uint32_t rank = (uint32_t)__popcll(peers & below);
    if (valid && rank == 0u) waveHist[wave][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    if (valid) {
      uint32_t pos = running[d] + rank;
      for (int w = 0; w < wave; w++) pos += waveHist[w][d];
      out[pos] = key;
    }
    __syncthreads();
    uint32_t add = 0u;
    for (int w = 0; w < kWaves; w++) {
      add += waveHist[w][threadIdx.x];
      waveHist[w][threadIdx.x] = 0u;
    }
    running[threadIdx.x] += add;
    __syncthreads();
  }
}

// Sorts n keys from `in` into `out` (both device buffers of n keys; `in` is
// overwritten).
void sortKeysByCode(uint64_t* in, uint64_t* out, int n, std::vector<DevBuf>& pool, hipStream_t s) {
  const int blocks = (n + kSortChunk - 1) / kSortChunk;
  pool.emplace_back();
  pool.back().alloc((size_t)n * 8);
  uint64_t* tmp = pool.back().as<uint64_t>();
  pool.emplace_back();
  pool.back().alloc((size_t)blocks * 256 * 4);
  uint32_t* counts = pool.back().as<uint32_t>();
  pool.emplace_back();
  pool.back().alloc((size_t)blocks * 256 * 4);
  uint32_t* offsets = pool.back().as<uint32_t>();
  // passes at bits 32, 40, 48, 56: in -> tmp -> in -> tmp -> out
  uint64_t* src[4] = {in, tmp, in, tmp};
  uint64_t* dst[4] = {tmp, in, tmp, out};
  for (int p = 0; p < 4; p++) {
    const int shift = 32 + 8 * p;
    hipLaunchKernelGGL(k_radix_count, dim3(blocks), dim3(kSortThreads), 0, s, src[p], n, shift, counts);
    HIP_OK(hipGetLastError());
    exclusiveScan(counts, offsets, blocks * 256, pool, s);
    hipLaunchKernelGGL(k_radix_scatter, dim3(blocks), dim3(kSortThreads), 0, s, src[p], dst[p], n, shift,
                       offsets);
    HIP_OK(hipGetLastError());
  }
}

}  // namespace

// Builds one tree into `out` (4-wide nodes at [0, numNodes)) from 2n float4
// primitive boxes (lo, hi per primitive). Leaf references are final: with
// refIsPrim they name the input primitive (the top level: one instance per
// leaf), otherwise primBase + a position in the sorted order (a BLAS whose
// triangles are emitted in that order from `sortedKeys`). Inner references
// are offset by nodeBase. Returns the node count; `stackBound` receives the
// worst-case traversal stack of the tree.
int lbvhBuild(const float4* boxes, int n, int maxLeaf, bool refIsPrim, uint32_t nodeBase, uint32_t primBase,
              DevBuf& out, DevBuf& sortedKeys, int& stackBound, hipStream_t s) {
  if (n <= 0) throw std::runtime_error("lbvh: no primitives");
  if (maxLeaf < 1 || maxLeaf > 16) throw std::runtime_error("lbvh: maxLeaf must be 1..16");
  DevBuf bounds, keysIn, nodes, leafParent, visits, isNode4, index4, fanout, bound;
  std::vector<DevBuf> pool;  // sort and scan temporaries, alive until the final synchronisation
  bounds.alloc(6 * 4);
  keysIn.alloc((size_t)n * 8);
  sortedKeys.alloc((size_t)n * 8);
  hipLaunchKernelGGL(k_bounds_init, dim3(1), dim3(64), 0, s, bounds.as<uint32_t>());
  hipLaunchKernelGGL(k_centroid_bounds, dim3(grid(n)), dim3(kLbvhBlock), 0, s, boxes, n, bounds.as<uint32_t>());
  hipLaunchKernelGGL(k_morton, dim3(grid(n)), dim3(kLbvhBlock), 0, s, boxes, n, bounds.as<uint32_t>(),
                     keysIn.as<uint64_t>());
  sortKeysByCode(keysIn.as<uint64_t>(), sortedKeys.as<uint64_t>(), n, pool, s);
  const uint64_t* keys = sortedKeys.as<uint64_t>();
  if (n <= maxLeaf || n == 1) {
    out.alloc(sizeof(BvhNode4));
    hipLaunchKernelGGL(k_single_leaf, dim3(1), dim3(64), 0, s, boxes, keys, n, refIsPrim ? 1 : 0, primBase,
                       out.as<BvhNode4>());
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(s));
    stackBound = 0;
    return 1;
  }
  const int ni = n - 1;
  nodes.alloc((size_t)ni * sizeof(BinNode));
  leafParent.alloc((size_t)n * 4);
  visits.alloc((size_t)ni * 4);
  isNode4.alloc((size_t)ni * 4);
  index4.alloc((size_t)ni * 4);
  HIP_OK(hipMemsetAsync(visits.p, 0, (size_t)ni * 4, s));
  hipLaunchKernelGGL(k_karras, dim3(grid(ni)), dim3(kLbvhBlock), 0, s, keys, n, nodes.as<BinNode>(),
                     leafParent.as<uint32_t>());
  hipLaunchKernelGGL(k_refit, dim3(grid(n)), dim3(kLbvhBlock), 0, s, nodes.as<BinNode>(), boxes, keys, n,
                     leafParent.as<uint32_t>(), visits.as<uint32_t>());
  hipLaunchKernelGGL(k_classify, dim3(grid(ni)), dim3(kLbvhBlock), 0, s, nodes.as<BinNode>(), ni, maxLeaf,
                     isNode4.as<uint32_t>());
  exclusiveScan(isNode4.as<uint32_t>(), index4.as<uint32_t>(), ni, pool, s);
  uint32_t lastIdx = 0, lastFlag = 0;
  HIP_OK(hipMemcpyAsync(&lastIdx, index4.as<uint32_t>() + ni - 1, 4, hipMemcpyDeviceToHost, s));
  HIP_OK(hipMemcpyAsync(&lastFlag, isNode4.as<uint32_t>() + ni - 1, 4, hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  const int num4 = (int)(lastIdx + lastFlag);
  out.alloc((size_t)num4 * sizeof(BvhNode4));
  fanout.alloc((size_t)num4 * 4);
  bound.alloc(4);
  HIP_OK(hipMemsetAsync(bound.p, 0, 4, s));
  hipLaunchKernelGGL(k_emit, dim3(grid(ni)), dim3(kLbvhBlock), 0, s, nodes.as<BinNode>(), ni, boxes, keys,
                     isNode4.as<uint32_t>(), index4.as<uint32_t>(), maxLeaf, refIsPrim ? 1 : 0, nodeBase, primBase,
                     out.as<BvhNode4>(), fanout.as<int32_t>());
  hipLaunchKernelGGL(k_stack_bound, dim3(grid(ni)), dim3(kLbvhBlock), 0, s, nodes.as<BinNode>(), ni,
                     isNode4.as<uint32_t>(), index4.as<uint32_t>(), fanout.as<int32_t>(), bound.as<int32_t>());
  HIP_OK(hipGetLastError());
  int32_t sb = 0;
  HIP_OK(hipMemcpyAsync(&sb, bound.p, 4, hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  stackBound = sb;
  return num4;
}

hipError_t launchInstanceBoxes(const float* pos, const InstanceBoxJob* jobs, int n, float4* boxes, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_instance_boxes, dim3(n), dim3(256), 0, s, pos, jobs, n, boxes);
  return hipGetLastError();
}

hipError_t launchTriBoxes(const float* pos, const uint32_t* idx, int n, float4* boxes, hipStream_t s) {
  hipLaunchKernelGGL(k_tri_boxes, dim3(grid(n)), dim3(kLbvhBlock), 0, s, pos, idx, n, boxes);
  return hipGetLastError();
}

hipError_t launchEmitTris(const uint64_t* sortedKeys, int n, const float* pos, const uint32_t* idx, float4* tris,
                          hipStream_t s) {
  hipLaunchKernelGGL(k_emit_tris, dim3(grid(n)), dim3(kLbvhBlock), 0, s, sortedKeys, n, pos, idx, tris);
  return hipGetLastError();
}

}  // namespace sdk
