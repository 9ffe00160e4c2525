// GPU builds of the two-level BVH over 30-bit Morton codes of primitive
// centroids: a linear BVH (Karras, "Maximizing Parallelism in the
// Construction of BVHs, Octrees, and k-d Trees", HPG 2012) refit bottom-up,
// or PLOC clustering (Meister and Bittner, IEEE TVCG 2018) whose merges
// follow the surface-area cost. Either binary tree is collapsed on the
// device into the same 4-wide 128-byte nodes the host SAH build emits
// (bvh.h), opening the largest child first; a subtree of at most maxLeaf
// primitives (a contiguous range of the leaf order) becomes one leaf. They
// replace the driver-side builds of the reference (src/accel.cpp:8-121).
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

// The same scan on buffers allocated once for a largest n, for builds that
// scan every round (allocating per call would cost more than the scans).
struct ScanScratch {
  std::vector<DevBuf> sums, offsets;  // one pair per level of block totals
  void reserve(int n) {
    sums.clear();
    offsets.clear();
    for (int blocks = (n + kScanChunk - 1) / kScanChunk; blocks > 1; blocks = (blocks + kScanChunk - 1) / kScanChunk) {
      sums.emplace_back();
      sums.back().alloc((size_t)blocks * 4);
      offsets.emplace_back();
      offsets.back().alloc((size_t)blocks * 4);
    }
  }
};

void exclusiveScan(const uint32_t* in, uint32_t* out, int n, ScanScratch& ws, hipStream_t s, size_t level = 0) {
  const int blocks = (n + kScanChunk - 1) / kScanChunk;
  if (blocks <= 1) {
    hipLaunchKernelGGL(k_scan_block, dim3(1), dim3(kScanThreads), 0, s, in, out, (uint32_t*)nullptr, n);
    HIP_OK(hipGetLastError());
    return;
  }
  if (level >= ws.sums.size() || ws.sums[level].bytes < (size_t)blocks * 4)
    throw std::runtime_error("scan: workspace reserved for fewer elements");
  uint32_t* sums = ws.sums[level].as<uint32_t>();
  uint32_t* offsets = ws.offsets[level].as<uint32_t>();
  hipLaunchKernelGGL(k_scan_block, dim3(blocks), dim3(kScanThreads), 0, s, in, out, sums, n);
  HIP_OK(hipGetLastError());
  exclusiveScan(sums, offsets, blocks, ws, s, level + 1);
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

// ---------------------------------------------------------------------------
// PLOC: parallel locally-ordered clustering (Meister and Bittner, "Parallel
// Locally-Ordered Clustering for Bounding Volume Hierarchy Construction",
// IEEE TVCG 2018). The clusters start as the Morton-sorted primitives. In
// each round every cluster finds the cluster within `radius` positions of
// it whose merged box has the smallest surface area (ties: the lower
// position), mutual nearest neighbours merge into a new binary node, and the
// surviving clusters are compacted in order. Some pair is always mutual (the
// lowest-positioned cluster of a closest pair picks its partner, which picks
// it back), so every round merges at least one pair. Internal nodes are
// numbered from the root down (the last merge is node 0, the root the
// collapse expects), and a final pass orders the leaves depth-first so that
// every subtree covers a contiguous range of positions, the form the 4-wide
// collapse reads.
// ---------------------------------------------------------------------------

constexpr int kPlocRadiusMax = 32;  // neighbours searched on either side (1..32)

// Half the surface area of a box; NaN (a box with NaN bounds) counts as +inf
// so that the neighbour search stays a total order.
__device__ inline float halfArea(float4 lo, float4 hi) {
  const float dx = hi.x - lo.x, dy = hi.y - lo.y, dz = hi.z - lo.z;
  const float a = dx * dy + dy * dz + dz * dx;
  return a < __builtin_inff() ? a : __builtin_inff();
}
__device__ inline float4 min4(float4 a, float4 b) {
  return make_float4(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z), 0.0f);
}
__device__ inline float4 max4(float4 a, float4 b) {
  return make_float4(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z), 0.0f);
}

// Initial clusters: one leaf per sorted position, with its primitive's box.
__global__ void k_ploc_init(const float4* __restrict__ boxes, const uint64_t* __restrict__ keys, int n,
                            uint32_t* __restrict__ clusters, float4* __restrict__ cbox) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const uint32_t prim = (uint32_t)keys[p];
  clusters[p] = kLeafBit | (uint32_t)p;
  cbox[2 * p] = boxes[2 * prim];
  cbox[2 * p + 1] = boxes[2 * prim + 1];
}

// Nearest neighbour of every cluster within the radius; the block's clusters
// and the radius on either side are staged in LDS.
// Clustering state kept on the device, so that several rounds run between
// host synchronisations (a round sized for more clusters than are left only
// skips the extra positions; once one cluster is left, rounds do nothing).
struct PlocState {
  uint32_t m;        // clusters left
  uint32_t created;  // merges so far (internal nodes written)
  uint32_t error;    // a round that merged nothing or lost count
  uint32_t rounds;   // rounds that merged
};

__global__ void k_ploc_state_init(PlocState* st, uint32_t n) {
  if (threadIdx.x == 0) *st = PlocState{n, 0u, 0u, 0u};
}

__global__ __launch_bounds__(kLbvhBlock) void k_ploc_nn(const float4* __restrict__ cbox,
                                                        const PlocState* __restrict__ st, int radius,
                                                        uint32_t* __restrict__ nn) {
  const int m = (int)st->m;
  if (m <= 1) return;  // uniform: every thread of the block returns
  __shared__ float4 s[2 * (kLbvhBlock + 2 * kPlocRadiusMax)];
  const int span = kLbvhBlock + 2 * radius;
  const int base = (int)blockIdx.x * kLbvhBlock - radius;
  for (int k = threadIdx.x; k < span; k += blockDim.x) {
    const int j = base + k;
    if (j >= 0 && j < m) {
      s[2 * k] = cbox[2 * j];
      s[2 * k + 1] = cbox[2 * j + 1];
    }
  }
  __syncthreads();
  const int i = (int)blockIdx.x * kLbvhBlock + threadIdx.x;
  if (i >= m) return;
  const float4 lo = s[2 * (i - base)], hi = s[2 * (i - base) + 1];
  float best = __builtin_inff();
  int bestJ = -1;
  const int j0 = max(0, i - radius), j1 = min(m - 1, i + radius);
  for (int j = j0; j <= j1; j++) {
    if (j == i) continue;
    const int k = j - base;
    const float a = halfArea(min4(lo, s[2 * k]), max4(hi, s[2 * k + 1]));
    if (bestJ < 0 || a < best) {
      best = a;
      bestJ = j;
    }
  }
  nn[i] = (uint32_t)bestJ;
}

// merge[i]: cluster i is the lower of a mutual pair (it becomes the new
// node); keep[i]: cluster i survives the round (it is not the upper one).
__global__ void k_ploc_flags(const uint32_t* __restrict__ nn, int mUpper, const PlocState* __restrict__ st,
                             uint32_t* __restrict__ merge, uint32_t* __restrict__ keep) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= mUpper) return;
  const int m = (int)st->m;
  if (m <= 1 || i >= m) {  // positions past the clusters left scan as zeros
    merge[i] = 0u;
    keep[i] = 0u;
    return;
  }
  const uint32_t j = nn[i];
  const bool mutual = nn[j] == (uint32_t)i;
  merge[i] = (mutual && (uint32_t)i < j) ? 1u : 0u;
  keep[i] = (mutual && (uint32_t)i > j) ? 0u : 1u;
}

// The round's totals (merges, surviving clusters) into the state, checked:
// a round merges at least one pair and every merge removes one cluster.
__global__ void k_ploc_totals(const uint32_t* __restrict__ merge, const uint32_t* __restrict__ mergePos,
                              const uint32_t* __restrict__ keep, const uint32_t* __restrict__ keepPos, int mUpper,
                              int numInternal, PlocState* __restrict__ st) {
  if (threadIdx.x != 0) return;
  const uint32_t m = st->m;
  if (m <= 1) return;
  const uint32_t merges = mergePos[mUpper - 1] + merge[mUpper - 1];
  const uint32_t kept = keepPos[mUpper - 1] + keep[mUpper - 1];
  if (merges == 0 || kept != m - merges || st->created + merges > (uint32_t)numInternal) {
    st->error = 1u;
    st->m = 1u;  // stop the remaining rounds
    return;
  }
  st->created += merges;
  st->m = kept;
  st->rounds += 1u;
}

// New binary nodes for the merged pairs and the compacted cluster list.
// `created` merges happened in earlier rounds; node ids count down from
// numInternal - 1, so the final merge is node 0.
__global__ void k_ploc_merge(const uint32_t* __restrict__ clusters, const float4* __restrict__ cbox,
                             const uint32_t* __restrict__ nn, const uint32_t* __restrict__ merge,
                             const uint32_t* __restrict__ mergePos, const uint32_t* __restrict__ keep,
                             const uint32_t* __restrict__ keepPos, const PlocState* __restrict__ st,
                             int numInternal,
                             BinNode* __restrict__ nodes, uint32_t* __restrict__ leafParent,
                             uint32_t* __restrict__ count, uint32_t* __restrict__ outClusters,
                             float4* __restrict__ outBox) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int m = (int)st->m;
  if (m <= 1 || i >= m || !keep[i]) return;
  const uint32_t created = st->created;  // updated by k_ploc_totals after this kernel
  const uint32_t dst = keepPos[i];
  if (!merge[i]) {
    outClusters[dst] = clusters[i];
    outBox[2 * dst] = cbox[2 * i];
    outBox[2 * dst + 1] = cbox[2 * i + 1];
    return;
  }
  const uint32_t j = nn[i];
  if (created + mergePos[i] >= (uint32_t)numInternal) return;  // cannot happen: n - 1 merges in all
  const uint32_t id = (uint32_t)(numInternal - 1) - (created + mergePos[i]);
  const uint32_t a = clusters[i], b = clusters[j];
  const float4 lo = min4(cbox[2 * i], cbox[2 * j]), hi = max4(cbox[2 * i + 1], cbox[2 * j + 1]);
  BinNode& nd = nodes[id];
  nd.left = a;
  nd.right = b;
  nd.parent = 0xFFFFFFFFu;  // set when this node is merged in a later round
  nd.boxLo = lo;
  nd.boxHi = hi;
  count[id] = ((a & kLeafBit) ? 1u : count[a]) + ((b & kLeafBit) ? 1u : count[b]);
  if (a & kLeafBit) leafParent[a & ~kLeafBit] = id;
  else nodes[a].parent = id;
  if (b & kLeafBit) leafParent[b & ~kLeafBit] = id;
  else nodes[b].parent = id;
  outClusters[dst] = id;
  outBox[2 * dst] = lo;
  outBox[2 * dst + 1] = hi;
}

__device__ inline uint32_t subtreeCount(const uint32_t* count, uint32_t c) {
  return (c & kLeafBit) ? 1u : count[c];
}

// Depth-first position of a subtree's first leaf: the sizes of the left
// siblings along its path to the root (at most numInternal steps).
__device__ inline uint32_t dfsOffset(const BinNode* nodes, const uint32_t* count, uint32_t child, uint32_t parent,
                                     int numInternal) {
  uint32_t off = 0;
  for (int step = 0; parent != 0xFFFFFFFFu && step < numInternal; step++) {
    const BinNode& nd = nodes[parent];
    if (nd.right == child) off += subtreeCount(count, nd.left);
    child = parent;
    parent = nd.parent;
  }
  return off;
}

__global__ void k_ploc_leaf_order(const BinNode* __restrict__ nodes, const uint32_t* __restrict__ count,
                                  const uint32_t* __restrict__ leafParent, int n, uint32_t* __restrict__ dfsPos) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  dfsPos[p] = dfsOffset(nodes, count, kLeafBit | (uint32_t)p, leafParent[p], n - 1);
}

// Each internal node's range of depth-first positions, and its leaf
// references renamed to depth-first positions (only leaf references change,
// and the walk compares internal ids only).
__global__ void k_ploc_ranges(BinNode* __restrict__ nodes, const uint32_t* __restrict__ count,
                              const uint32_t* __restrict__ dfsPos, int numInternal) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= numInternal) return;
  const uint32_t lo = dfsOffset(nodes, count, (uint32_t)i, nodes[i].parent, numInternal);
  nodes[i].lo = lo;
  nodes[i].hi = lo + count[i] - 1;
  const uint32_t l = nodes[i].left, r = nodes[i].right;
  if (l & kLeafBit) nodes[i].left = kLeafBit | dfsPos[l & ~kLeafBit];
  if (r & kLeafBit) nodes[i].right = kLeafBit | dfsPos[r & ~kLeafBit];
}

__global__ void k_ploc_keys(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ dfsPos, int n,
                            uint64_t* __restrict__ out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const uint32_t q = dfsPos[p];
  if (q < (uint32_t)n) out[q] = keys[p];  // a permutation of [0, n) for a well-formed tree
}

// The PLOC tree over n >= 2 Morton-sorted primitives: n - 1 internal nodes
// with depth-first ranges, and the primitives' keys in depth-first order
// (replacing `keys`).
void plocTree(const float4* boxes, DevBuf& keys, int n, int radius, DevBuf& nodes, hipStream_t s) {
  if (radius < 1 || radius > kPlocRadiusMax) throw std::runtime_error("ploc: radius must be 1..32");
  constexpr int kRoundsPerSync = 4;
  const int ni = n - 1;
  DevBuf clA, clB, bxA, bxB, nn, merge, keep, mergePos, keepPos, count, leafParent, state, dfsPos, keysDfs;
  clA.alloc((size_t)n * 4);
  clB.alloc((size_t)n * 4);
  bxA.alloc((size_t)n * 2 * sizeof(float4));
  bxB.alloc((size_t)n * 2 * sizeof(float4));
  nn.alloc((size_t)n * 4);
  merge.alloc((size_t)n * 4);
  keep.alloc((size_t)n * 4);
  mergePos.alloc((size_t)n * 4);
  keepPos.alloc((size_t)n * 4);
  count.alloc((size_t)ni * 4);
  leafParent.alloc((size_t)n * 4);
  state.alloc(sizeof(PlocState));
  nodes.alloc((size_t)ni * sizeof(BinNode));
  ScanScratch ws;
  ws.reserve(n);
  hipLaunchKernelGGL(k_ploc_state_init, dim3(1), dim3(64), 0, s, state.as<PlocState>(), (uint32_t)n);
  hipLaunchKernelGGL(k_ploc_init, dim3(grid(n)), dim3(kLbvhBlock), 0, s, boxes, keys.as<uint64_t>(), n,
                     clA.as<uint32_t>(), bxA.as<float4>());
  HIP_OK(hipGetLastError());
  PlocState* st = state.as<PlocState>();
  int mUpper = n;  // clusters left as of the last synchronisation (an upper bound afterwards)
  DevBuf *cl = &clA, *clOut = &clB, *bx = &bxA, *bxOut = &bxB;
  PlocState h{};
  for (int batch = 0;; batch++) {
    if (batch * kRoundsPerSync > n + 64) throw std::runtime_error("ploc: clustering did not converge");
    for (int r = 0; r < kRoundsPerSync; r++) {
      hipLaunchKernelGGL(k_ploc_nn, dim3(grid(mUpper)), dim3(kLbvhBlock), 0, s, bx->as<float4>(), st, radius,
                         nn.as<uint32_t>());
      hipLaunchKernelGGL(k_ploc_flags, dim3(grid(mUpper)), dim3(kLbvhBlock), 0, s, nn.as<uint32_t>(), mUpper, st,
                         merge.as<uint32_t>(), keep.as<uint32_t>());
      HIP_OK(hipGetLastError());
      exclusiveScan(merge.as<uint32_t>(), mergePos.as<uint32_t>(), mUpper, ws, s);
      exclusiveScan(keep.as<uint32_t>(), keepPos.as<uint32_t>(), mUpper, ws, s);
      hipLaunchKernelGGL(k_ploc_merge, dim3(grid(mUpper)), dim3(kLbvhBlock), 0, s, cl->as<uint32_t>(),
                         bx->as<float4>(), nn.as<uint32_t>(), merge.as<uint32_t>(), mergePos.as<uint32_t>(),
                         keep.as<uint32_t>(), keepPos.as<uint32_t>(), st, ni, nodes.as<BinNode>(),
                         leafParent.as<uint32_t>(), count.as<uint32_t>(), clOut->as<uint32_t>(),
                         bxOut->as<float4>());
      hipLaunchKernelGGL(k_ploc_totals, dim3(1), dim3(64), 0, s, merge.as<uint32_t>(), mergePos.as<uint32_t>(),
                         keep.as<uint32_t>(), keepPos.as<uint32_t>(), mUpper, ni, st);
      HIP_OK(hipGetLastError());
      std::swap(cl, clOut);
      std::swap(bx, bxOut);
    }
    HIP_OK(hipMemcpyAsync(&h, state.p, sizeof h, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    if (h.error) throw std::runtime_error("ploc: inconsistent round");
    if (h.m <= 1) break;
    mUpper = (int)h.m;
  }
  const uint32_t created = h.created;
  if (created != (uint32_t)ni) throw std::runtime_error("ploc: incomplete tree");
  dfsPos.alloc((size_t)n * 4);
  keysDfs.alloc((size_t)n * 8);
  hipLaunchKernelGGL(k_ploc_leaf_order, dim3(grid(n)), dim3(kLbvhBlock), 0, s, nodes.as<BinNode>(),
                     count.as<uint32_t>(), leafParent.as<uint32_t>(), n, dfsPos.as<uint32_t>());
  hipLaunchKernelGGL(k_ploc_ranges, dim3(grid(ni)), dim3(kLbvhBlock), 0, s, nodes.as<BinNode>(),
                     count.as<uint32_t>(), dfsPos.as<uint32_t>(), ni);
  hipLaunchKernelGGL(k_ploc_keys, dim3(grid(n)), dim3(kLbvhBlock), 0, s, keys.as<uint64_t>(), dfsPos.as<uint32_t>(),
                     n, keysDfs.as<uint64_t>());
  HIP_OK(hipGetLastError());
  HIP_OK(hipStreamSynchronize(s));
  keys = std::move(keysDfs);
}

// ---------------------------------------------------------------------------
// Collapse of a binary tree (LBVH or PLOC) into 4-wide nodes, top down, one
// level of 4-wide nodes per round: a node's slots start as its binary
// root's two children, and while there are fewer than four, the slot with
// the largest surface area that is an inner binary node (not a collapsed
// leaf range) is replaced by its two children (the greedy opening of wide
// BVH builders). Inner slots become the next level's 4-wide nodes, numbered
// level by level (the root is node 0).
// ---------------------------------------------------------------------------

struct Slots {
  uint32_t c[4];
  int k;
};

__device__ inline bool openable(const BinNode* nodes, uint32_t c, int maxLeaf) {
  return !(c & kLeafBit) && !collapsed(nodes[c], maxLeaf);
}

__device__ inline Slots expandSlots(const BinNode* nodes, uint32_t r, int maxLeaf) {
  Slots s;
  s.k = 2;
  s.c[0] = nodes[r].left;
  s.c[1] = nodes[r].right;
  s.c[2] = s.c[3] = 0;
  while (s.k < 4) {
    int best = -1;
    float bestA = -1.0f;
    for (int i = 0; i < s.k; i++) {
      if (!openable(nodes, s.c[i], maxLeaf)) continue;
      const float a = halfArea(nodes[s.c[i]].boxLo, nodes[s.c[i]].boxHi);
      if (a > bestA) {
        bestA = a;
        best = i;
      }
    }
    if (best < 0) break;
    const uint32_t c = s.c[best];
    s.c[best] = nodes[c].left;
    s.c[s.k++] = nodes[c].right;
  }
  return s;
}

// Pass 1 of a level: how many of each 4-wide node's slots are inner nodes.
__global__ void k_wide_count(const BinNode* __restrict__ nodes, const uint32_t* __restrict__ frontier, int nF,
                             int maxLeaf, uint32_t* __restrict__ innerCount) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nF) return;
  const Slots s = expandSlots(nodes, frontier[f], maxLeaf);
  uint32_t c = 0;
  for (int i = 0; i < s.k; i++) c += openable(nodes, s.c[i], maxLeaf) ? 1u : 0u;
  innerCount[f] = c;
}

__global__ void k_wide_total(const uint32_t* __restrict__ innerCount, const uint32_t* __restrict__ innerPos, int nF,
                             uint32_t* __restrict__ total) {
  if (threadIdx.x == 0) total[0] = innerPos[nF - 1] + innerCount[nF - 1];
}

// Pass 2 of a level: write the level's 4-wide nodes (indices levelBase + f)
// and the next level's frontier (indices levelBase + nF + position). acc[f]
// is the sum of (fan-out - 1) over the node's 4-wide ancestors; the largest
// acc + fan-out - 1 of any node bounds the traversal stack.
__global__ void k_wide_emit(const BinNode* __restrict__ nodes, const float4* __restrict__ boxes,
                            const uint64_t* __restrict__ keys, const uint32_t* __restrict__ frontier,
                            const uint32_t* __restrict__ acc, const uint32_t* __restrict__ innerPos, int nF,
                            uint32_t levelBase, int maxLeaf, int refIsPrim, uint32_t nodeBase, uint32_t primBase,
                            BvhNode4* __restrict__ out, uint32_t* __restrict__ nextFrontier,
                            uint32_t* __restrict__ nextAcc, int32_t* __restrict__ bound) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nF) return;
  const Slots s = expandSlots(nodes, frontier[f], maxLeaf);
  const uint32_t myAcc = acc[f] + (uint32_t)(s.k - 1);
  uint32_t next = innerPos[f];
  BvhNode4 o;
  for (int i = 0; i < 4; i++) {
    if (i >= s.k) {
      o.lox[i] = o.loy[i] = o.loz[i] = o.hix[i] = o.hiy[i] = o.hiz[i] = 3e38f;
      o.child[i] = (int32_t)kRefEmpty;
      o.count[i] = -1;
      continue;
    }
    const uint32_t c = s.c[i];
    float4 lo, hi;
    childBox(nodes, boxes, keys, c, lo, hi);
    o.lox[i] = lo.x; o.loy[i] = lo.y; o.loz[i] = lo.z;
    o.hix[i] = hi.x; o.hiy[i] = hi.y; o.hiz[i] = hi.z;
    if (openable(nodes, c, maxLeaf)) {
      o.child[i] = (int32_t)(nodeBase + levelBase + (uint32_t)nF + next);
      o.count[i] = 0;
      nextFrontier[next] = c;
      nextAcc[next] = myAcc;
      next++;
    } else {
      uint32_t first, cnt;
      if (c & kLeafBit) {
        first = c & ~kLeafBit;
        cnt = 1;
      } else {
        first = nodes[c].lo;
        cnt = nodes[c].hi - nodes[c].lo + 1;
      }
      const uint32_t target = refIsPrim ? (uint32_t)keys[first] : primBase + first;
      o.child[i] = (int32_t)(kRefLeaf | ((cnt - 1) << 27) | target);
      o.count[i] = (int32_t)cnt;
    }
  }
  out[levelBase + f] = o;
  atomicMax(bound, (int32_t)myAcc);
}

// 4-wide nodes of the binary tree over n >= 2 primitives (root: binary node
// 0, not a collapsed range) into `out`; returns their number.
int collapseWide(const BinNode* nodes, const float4* boxes, const uint64_t* keys, int n, int maxLeaf, bool refIsPrim,
                 uint32_t nodeBase, uint32_t primBase, DevBuf& out, int& stackBound, hipStream_t s) {
  const int ni = n - 1;
  DevBuf fA, fB, aA, aB, innerCount, innerPos, total, bound;
  ScanScratch ws;
  ws.reserve(ni);
  fA.alloc((size_t)ni * 4);
  fB.alloc((size_t)ni * 4);
  aA.alloc((size_t)ni * 4);
  aB.alloc((size_t)ni * 4);
  innerCount.alloc((size_t)ni * 4);
  innerPos.alloc((size_t)ni * 4);
  total.alloc(4);
  bound.alloc(4);
  out.alloc((size_t)ni * sizeof(BvhNode4));  // every 4-wide node opens at least one binary node
  HIP_OK(hipMemsetAsync(fA.p, 0, 4, s));     // the root
  HIP_OK(hipMemsetAsync(aA.p, 0, 4, s));
  HIP_OK(hipMemsetAsync(bound.p, 0, 4, s));
  DevBuf *fr = &fA, *frNext = &fB, *ac = &aA, *acNext = &aB;
  uint32_t levelBase = 0;
  int nF = 1;
  while (nF > 0) {
    if (levelBase + (uint32_t)nF > (uint32_t)ni)
      throw std::runtime_error("bvh collapse: more 4-wide nodes than binary nodes");
    hipLaunchKernelGGL(k_wide_count, dim3(grid(nF)), dim3(kLbvhBlock), 0, s, nodes, fr->as<uint32_t>(), nF, maxLeaf,
                       innerCount.as<uint32_t>());
    HIP_OK(hipGetLastError());
    exclusiveScan(innerCount.as<uint32_t>(), innerPos.as<uint32_t>(), nF, ws, s);
    hipLaunchKernelGGL(k_wide_total, dim3(1), dim3(64), 0, s, innerCount.as<uint32_t>(), innerPos.as<uint32_t>(), nF,
                       total.as<uint32_t>());
    hipLaunchKernelGGL(k_wide_emit, dim3(grid(nF)), dim3(kLbvhBlock), 0, s, nodes, boxes, keys, fr->as<uint32_t>(),
                       ac->as<uint32_t>(), innerPos.as<uint32_t>(), nF, levelBase, maxLeaf, refIsPrim ? 1 : 0,
                       nodeBase, primBase, out.as<BvhNode4>(), frNext->as<uint32_t>(), acNext->as<uint32_t>(),
                       bound.as<int32_t>());
    HIP_OK(hipGetLastError());
    uint32_t nNext = 0;
    HIP_OK(hipMemcpyAsync(&nNext, total.p, 4, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    levelBase += (uint32_t)nF;
    nF = (int)nNext;
    std::swap(fr, frNext);
    std::swap(ac, acNext);
  }
  int32_t sb = 0;
  HIP_OK(hipMemcpyAsync(&sb, bound.p, 4, hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  stackBound = sb;
  return (int)levelBase;
}

// Builds one tree into `out` (4-wide nodes at [0, numNodes)) from 2n float4
// primitive boxes (lo, hi per primitive). Leaf references are final: with
// refIsPrim they name the input primitive (the top level: one instance per
// leaf), otherwise primBase + a position in the sorted order (a BLAS whose
// triangles are emitted in that order from `sortedKeys`). Inner references
// are offset by nodeBase. Returns the node count; `stackBound` receives the
// worst-case traversal stack of the tree. method 0 builds Karras' radix
// tree over the Morton order, 1 the PLOC tree (above) with the given search
// radius.
int lbvhBuild(const float4* boxes, int n, int maxLeaf, bool refIsPrim, uint32_t nodeBase, uint32_t primBase,
              DevBuf& out, DevBuf& sortedKeys, int& stackBound, hipStream_t s, int method, int plocRadius) {
  if (method != 0 && method != 1) throw std::runtime_error("lbvh: method must be 0 (LBVH) or 1 (PLOC)");
  if (n <= 0) throw std::runtime_error("lbvh: no primitives");
  if (maxLeaf < 1 || maxLeaf > 16) throw std::runtime_error("lbvh: maxLeaf must be 1..16");
  DevBuf bounds, keysIn, nodes, leafParent, visits;
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
  if (method == 1) {
    plocTree(boxes, sortedKeys, n, plocRadius, nodes, s);
    keys = sortedKeys.as<uint64_t>();
  } else {
    nodes.alloc((size_t)ni * sizeof(BinNode));
    leafParent.alloc((size_t)n * 4);
    visits.alloc((size_t)ni * 4);
    HIP_OK(hipMemsetAsync(visits.p, 0, (size_t)ni * 4, s));
    hipLaunchKernelGGL(k_karras, dim3(grid(ni)), dim3(kLbvhBlock), 0, s, keys, n, nodes.as<BinNode>(),
                       leafParent.as<uint32_t>());
    hipLaunchKernelGGL(k_refit, dim3(grid(n)), dim3(kLbvhBlock), 0, s, nodes.as<BinNode>(), boxes, keys, n,
                       leafParent.as<uint32_t>(), visits.as<uint32_t>());
    HIP_OK(hipGetLastError());
  }
  const int num4 = collapseWide(nodes.as<BinNode>(), boxes, keys, n, maxLeaf, refIsPrim, nodeBase, primBase, out,
                                stackBound, s);
  HIP_OK(hipStreamSynchronize(s));
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
