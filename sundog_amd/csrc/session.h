// Resident render session: device scene tables, two-level BVH, film and
// launch state of one scene on one GPU. Shared by the session entry points
// and the serving loop.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/sundog_api.h"
#include "bvh.h"
#include "dev.h"
#include "prepare.h"

namespace sdk {

hipError_t launchRender(const LaunchParams& P, bool countWork, int waves, int groups, hipStream_t s);
hipError_t launchTrace(const LaunchParams& P, const float* rays, int n, bool brute, int numInst, float* t,
                       int* inst, int* prim, int* counters, hipStream_t s);
hipError_t launchShadow(const LaunchParams& P, const float* rays, int n, float* tr, hipStream_t s);
hipError_t launchEnvBins(const LaunchParams& P, const float* u, int n, int* row, int* col, hipStream_t s,
                         int mode = 0);
hipError_t launchLightSample(const LaunchParams& P, int light, const float* xs, int n, uint64_t seed0, uint64_t seq,
                            float* out, hipStream_t s);
hipError_t launchFlameField(const LaunchParams& P, int flame, const float* pts, int n, float* sigma, float* em,
                            hipStream_t s);
hipError_t launchPcg(uint64_t seed, uint64_t seq, int n, uint32_t* out, hipStream_t s);
hipError_t launchQueueReset(QueueState* q, int n, hipStream_t s);
hipError_t launchQueuePrep(QueueState* q, RequestDev* reqs, const int32_t* firstSample, int numReq,
                           hipStream_t s);
hipError_t launchPersistent(const SceneTable& S, const RequestDev* reqs, const WorkItem* items, int numItems,
                            QueueState* q, int tileSize, int waves, int blocks, uint32_t* spill, int spillStride,
                            int stackLds, float* sampleBuf, int64_t sampleBufWaves, float* groupBuf,
                            unsigned long long* rayCounter, hipStream_t s);
hipError_t launchFoldGroups(const RequestDev* reqs, const WorkItem* items, int numItems, int tileSize,
                            const float* groupBuf, hipStream_t s);
int persistentBlocksPerCU(int waves, int stackLds);
hipError_t launchWfShade(const LaunchParams& P, const WfParams& W, int blocks, hipStream_t s);
hipError_t launchWfClosest(const LaunchParams& P, const WfParams& W, int blocks, int waves, hipStream_t s);
hipError_t launchWfShadow(const LaunchParams& P, const WfParams& W, int blocks, int waves, hipStream_t s);
hipError_t launchWfClear(WfCounters* c, int what, hipStream_t s);
int wfBlocksPerCU(int which, int waves);
int stackLdsEntries();
int stackLdsDefault();
int stackLdsMax();
struct DevBuf;
int lbvhBuild(const float4* boxes, int n, int maxLeaf, bool refIsPrim, uint32_t nodeBase, uint32_t primBase,
              DevBuf& out, DevBuf& sortedKeys, int& stackBound, hipStream_t s, int method = 0,
              int plocRadius = 16);
// One world-box job of the GPU rebuild: the object-to-world transform and the
// object's vertices in the position pool (numVerts 0: keep the host's box).
struct InstanceBoxJob {
  float xf[12];
  int32_t posOff, numVerts, pad[2];
};
hipError_t launchInstanceBoxes(const float* pos, const InstanceBoxJob* jobs, int n, float4* boxes, hipStream_t s);
hipError_t launchTriBoxes(const float* pos, const uint32_t* idx, int n, float4* boxes, hipStream_t s);
hipError_t launchEmitTris(const uint64_t* sortedKeys, int n, const float* pos, const uint32_t* idx, float4* tris,
                          hipStream_t s);
int renderBlockThreads();
constexpr int kWfBlock = 256;  // threads per block of the wavefront kernels
constexpr size_t kStackEntryBytes = 8;  // traversal stack entry: node reference + entry distance

// Auto occupancy for the megakernel: its variants (2 to 5 waves per SIMD)
// give bit-identical films, so the first passes of a render can time them
// on real work. After one untimed warm-up pass, candidates are timed in the
// order 5, 4, 3, 2 (cost per path); the trials stop at the first variant
// more than 5% slower than the best so far, and later passes use the best.
// Large frames favour many waves (latency hiding); small frames, whose grids
// cannot fill the chip, favour the spill-free low-occupancy variants.
// Sample groups for a megakernel launch of `blocks` 16x16 tiles and `count`
// samples on a device of `cus` CUs: up to 4 groups, each rendering a share
// of the samples in its own block per tile, until the grid has 4 blocks per
// CU (a 256x256 frame has one tile per CU). More than one group needs the
// launch's samples to fit one shared pool.
inline int sampleGroupsFor(size_t blocks, int count, int cus) {
  if (count > kPoolSamples || blocks == 0) return 1;
  const size_t want = (size_t)std::max(cus, 1) * 4;
  int g = 1;
  while (g < 4 && g < count && blocks * (size_t)g < want) g++;
  return g;
}

struct OccupancyPicker {
  static constexpr int kCandidates[4] = {5, 4, 3, 2};
  int rounds = 1;           // timed passes per candidate (the fastest counts)
  bool exhaustive = false;  // time every candidate instead of stopping early
  int chosen = 0;           // the variant settled on (0: still timing)
  int next = 0;             // index of the candidate being timed
  int round = 0;            // timed passes of that candidate so far
  bool warm = false;        // the untimed warm-up pass has run
  double msPerPath[6] = {};
  // The variant for the next pass; trial is set when that pass is timed.
  int pick(bool& trial) {
    trial = false;
    if (chosen != 0) return chosen;
    if (!warm) {
      warm = true;
      return kCandidates[0];
    }
    trial = true;
    return kCandidates[next];
  }
  void record(int w, double ms) {
    msPerPath[w] = round == 0 ? ms : std::min(msPerPath[w], ms);
    if (++round < rounds) return;
    round = 0;
    int best = kCandidates[0];
    for (int i = 0; i <= next; i++)
      if (msPerPath[kCandidates[i]] < msPerPath[best]) best = kCandidates[i];
    next++;
    if (next == 4 || (!exhaustive && msPerPath[w] > 1.05 * msPerPath[best])) chosen = best;
  }
  void reset() {
    const int r = rounds;
    const bool ex = exhaustive;
    *this = OccupancyPicker{};
    rounds = r;
    exhaustive = ex;
  }
};

// Sample groups per tile for the megakernel, chosen like the occupancy
// variant from timed passes (after it, at the chosen variant): 1, 2, 3, 4,
// 6, 8, 16 groups until one is more than 5% slower than the best so far. Frames
// whose cost varies a lot between tiles (flames) want many groups: the
// expensive tiles' samples spread over more blocks, so the grid does not end
// on a few long blocks.
struct GroupPicker {
  static constexpr int kN = 7;
  static constexpr int kCandidates[kN] = {1, 2, 3, 4, 6, 8, 16};
  int rounds = 1;
  bool exhaustive = false;
  int chosen = 0, next = 0, round = 0;
  double msPerPath[kPoolSamples + 1] = {};
  int best() const {
    int b = kCandidates[0];
    for (int i = 0; i < next && i < kN; i++)
      if (msPerPath[kCandidates[i]] < msPerPath[b]) b = kCandidates[i];
    return b;
  }
  // Groups for a pass of `count` samples; trial is set when the pass is timed.
  int pick(bool& trial, int count) {
    trial = false;
    if (chosen == 0 && next > 0 && next < kN && kCandidates[next] > count) chosen = best();
    if (chosen != 0) return std::min(chosen, std::max(count, 1));
    trial = true;
    return kCandidates[next];
  }
  void record(int g, double ms) {
    msPerPath[g] = round == 0 ? ms : std::min(msPerPath[g], ms);
    if (++round < rounds) return;
    round = 0;
    next++;
    const int b = best();
    if (next == kN || (!exhaustive && msPerPath[g] > 1.05 * msPerPath[b])) chosen = b;
  }
  void reset() {
    const int r = rounds;
    const bool ex = exhaustive;
    *this = GroupPicker{};
    rounds = r;
    exhaustive = ex;
  }
};

#define HIP_OK(expr)                                                                          \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess)                                                                     \
      throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e_) + " at " + \
                               #expr);                                                        \
  } while (0)

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) {
    o.p = nullptr;
    o.bytes = 0;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      bytes = o.bytes;
      o.p = nullptr;
      o.bytes = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  void alloc(size_t n) {
    release();
    if (n == 0) n = 16;
    HIP_OK(hipMalloc(&p, n));
    bytes = n;
  }
  template <class T>
  void upload(const std::vector<T>& v, hipStream_t s) {
    alloc(v.size() * sizeof(T));
    if (!v.empty()) HIP_OK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
  }
  template <class T>
  T* as() const {
    return reinterpret_cast<T*>(p);
  }
};

// Environment CDFs over luminance * sin(theta), prefix sums in double, stored
// as float with the last entry pinned to 1 (reference src/env_light.cpp:27-59).
inline bool buildEnvCdf(const HostImage& img, std::vector<float>& marg, std::vector<float>& cond) {
  int w = img.width, h = img.height;
  marg.assign(h + 1, 0.0f);
  cond.assign((size_t)h * (w + 1), 0.0f);
  double total = 0.0;
  for (int row = 0; row < h; row++) {
    float sinT = sinf(kPi * (row + 0.5f) / h);
    float* cdf = &cond[(size_t)row * (w + 1)];
    double sum = 0.0;
    cdf[0] = 0.0f;
    for (int col = 0; col < w; col++) {
      const float* p = &img.rgba[4 * ((size_t)row * w + col)];
      float lum = 0.2126f * p[0] + 0.7152f * p[1] + 0.0722f * p[2];
      sum += (double)fmaxf(lum, 0.0f) * sinT;
      cdf[col + 1] = (float)sum;
    }
    if (sum > 0.0) {
      for (int c = 1; c <= w; c++) cdf[c] = (float)(cdf[c] / sum);
    } else {
      for (int c = 0; c <= w; c++) cdf[c] = (float)c / w;
    }
    cdf[w] = 1.0f;
    total += sum;
    marg[row + 1] = (float)total;
  }
  bool degenerate = total <= 0.0;
  if (!degenerate) {
    for (int r = 1; r <= h; r++) marg[r] = (float)(marg[r] / total);
  } else {
    for (int r = 0; r <= h; r++) marg[r] = (float)r / h;
  }
  marg[h] = 1.0f;
  return !degenerate;
}

// 16-ary lookup table of a CDF of n entries (the device's cdfBin16): its
// keys cdf[1..n-2] in 16-float leaf blocks, over a mid level holding the
// first key of every leaf block after the first, under a top block holding
// the first entry of every mid block after the first; unused slots are +inf.
// Returns false when the top level would need more than one block.
struct CdfSearchShape {
  int leaf, mid;
  int blocks() const { return 1 + mid + leaf; }
};
inline CdfSearchShape cdfSearchShape(int n) {
  const int keys = n > 2 ? n - 2 : 0;
  CdfSearchShape sh;
  sh.leaf = keys > 0 ? (keys + 15) / 16 : 1;
  sh.mid = sh.leaf > 1 ? (sh.leaf - 1 + 15) / 16 : 1;
  return sh;
}
inline bool buildCdfSearch(const float* cdf, int n, float* out) {
  const CdfSearchShape sh = cdfSearchShape(n);
  if (sh.mid - 1 > 16) return false;
  const int keys = n > 2 ? n - 2 : 0;
  const float inf = __builtin_inff();
  for (int i = 0; i < 16 * sh.blocks(); i++) out[i] = inf;
  float* top = out;
  float* mid = out + 16;
  float* leaf = out + 16 * (1 + sh.mid);
  for (int k = 0; k < keys; k++) leaf[k] = cdf[1 + k];
  for (int b = 1; b < sh.leaf; b++) mid[b - 1] = leaf[16 * b];
  for (int c = 1; c < sh.mid; c++) top[c - 1] = mid[16 * c];
  return true;
}

inline Aabb quadricBox(int kind) {
  const float e = 1e-4f;
  Aabb b;
  switch (kind) {
    case GK_RECT:
    case GK_DISK:
      b.lo = v3(-1, -e, -1);
      b.hi = v3(1, e, 1);
      break;
    case GK_PARABOLA:
      b.lo = v3(-1, -e, -1);
      b.hi = v3(1, 0.5f + e, 1);
      break;
    default:
      b.lo = v3(-1, -1, -1);
      b.hi = v3(1, 1, 1);
  }
  return b;
}

// Widen a transformed box by a few ulps of its magnitude so that rounding
// (of the forward transform here, of the float inverse on the ray side)
// never shrinks it below the surface it bounds.
inline Aabb padTransformed(Aabb r) {
  vec3 mag = vmax(vec3{fabsf(r.lo.x), fabsf(r.lo.y), fabsf(r.lo.z)}, vec3{fabsf(r.hi.x), fabsf(r.hi.y), fabsf(r.hi.z)});
  vec3 pad = mag * 2e-6f + v3(1e-7f);
  r.lo = r.lo - pad;
  r.hi = r.hi + pad;
  return r;
}

inline Aabb transformBox(const Affine& m, const Aabb& b) {
  Aabb r;
  for (int i = 0; i < 8; i++) {
    vec3 c = vec3{(i & 1) ? b.hi.x : b.lo.x, (i & 2) ? b.hi.y : b.lo.y, (i & 4) ? b.hi.z : b.lo.z};
    r.grow(m.point(c));
  }
  return padTransformed(r);
}

// World box of every object. A mesh instance gets the bounds of its
// transformed vertices (those of its triangles' convex hull): transformed
// object-box corners grow with rotation and non-uniform scale (38% more
// surface area on average on the instanced-icosphere scene), and every
// extra unit of area is extra rays entering the instance for nothing.
// Quadrics keep their transformed boxes. Split over host threads.
inline std::vector<Aabb> objectWorldBoxes(const Scene& S, const std::vector<HostMesh>& meshes,
                                          const std::vector<Aabb>& meshBounds) {
  std::vector<Aabb> out(S.objects.size());
  std::vector<SoaPoints> soa(meshes.size());
  for (const Object& o : S.objects)
    if (o.geomKind == GK_MESH && soa[o.meshId].x.empty() && !meshes[o.meshId].positions.empty())
      soa[o.meshId] = toSoa(meshes[o.meshId].positions);
  auto work = [&](size_t begin, size_t end) {
    for (size_t i = begin; i < end; i++) {
      const Object& o = S.objects[i];
      if (o.geomKind != GK_MESH) {
        out[i] = transformBox(o.xform, quadricBox(o.geomKind));
        continue;
      }
      if (meshes[o.meshId].positions.empty()) {
        out[i] = transformBox(o.xform, meshBounds[o.meshId]);
        continue;
      }
      out[i] = padTransformed(transformedBounds(o.xform, soa[o.meshId]));
    }
  };
  const size_t n = out.size();
  const size_t threads = std::min<size_t>(std::max(1u, std::thread::hardware_concurrency()), 16);
  if (n < 64 || threads == 1) {
    work(0, n);
    return out;
  }
  std::vector<std::thread> pool;
  const size_t chunk = (n + threads - 1) / threads;
  for (size_t b = 0; b < n; b += chunk) pool.emplace_back(work, b, std::min(n, b + chunk));
  for (std::thread& t : pool) t.join();
  return out;
}

inline void relocate(std::vector<BvhNode4>& nodes, size_t from, size_t count, uint32_t nodeBase,
                     uint32_t primBase, const std::vector<uint32_t>* leafMap) {
  for (size_t i = from; i < from + count; i++) {
    for (int c = 0; c < 4; c++) {
      uint32_t r = (uint32_t)nodes[i].child[c];
      if (r == kRefEmpty) continue;
      if (r & kRefLeaf) {
        uint32_t first = r & 0x07FFFFFFu, hi = r & 0xF8000000u;
        uint32_t nf = leafMap ? (*leafMap)[first] : first + primBase;
        if (nf >= (1u << 27)) throw std::runtime_error("bvh: primitive index exceeds the packed range");
        nodes[i].child[c] = (int32_t)(hi | nf);
      } else {
        nodes[i].child[c] = (int32_t)(r + nodeBase);
      }
    }
  }
}

}  // namespace sdk

using namespace sdk;

struct sundog_session {
  hipStream_t stream = nullptr;
  Prepared prep;
  LaunchParams P{};
  int kernel = SUNDOG_KERNEL_MEGA;
  bool countWork = false;
  // Occupancy variant of the render kernel: 2 to 5 waves per SIMD, or 0 to
  // choose per scene from timed passes (autoPick).
  int waves = 0;
  OccupancyPicker autoPick;
  int wfWaves = 4;  // occupancy variant of the wavefront query kernels (4 or 5)
  int numInstances = 0;
  int stackNeed = 0;
  int spillRows = 0;
  DevBuf textures, texels, materials, lights, flames, instances, nodes, tris, meshes;
  DevBuf positions, indices, normals, uvs, lightPos, lightCdf;
  DevBuf envTex, margCdf, condCdf, envSearch;
  int envSearchShape[4] = {0, 0, 0, 0};
  DevBuf accum, albedo, normal, rayCounter, workCounters, spill, sampleBuf;
  size_t spillThreads = 0;
  // wavefront variant
  DevBuf wfState, wfClosest, wfShadow, wfBuckets, wfCnt;
  int wfSlots = 0;
  uint32_t* wfHostCnt = nullptr;  // pinned copy of the counters
  int wfIterations = 0;           // queue rounds of the last pass
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double renderMs = 0.0;
  double buildMs = 0.0;
  int cuCount = 0;  // CUs of the session's device (queried at the first pass)
  int sampleGroups = 0;  // megakernel sample groups per tile; 0: chosen by groupPick
  GroupPicker groupPick;
  uint64_t paths = 0;
  int tlasNodes = 0, blasNodes = 0;
  // 0: host binned SAH, 1: GPU LBVH, 2: GPU LBVH meshes + host SAH top,
  // 3: GPU PLOC, 4: GPU PLOC meshes + host SAH top
  int bvhBuilder = 0;
  int plocRadius = 16;       // neighbours searched on either side by the PLOC build
  double bvhGpuMs = 0.0;     // wall time of the last GPU build
  std::vector<DevInstance> instHost;
  std::vector<int> meshPosOff, meshIdxOff;
  std::vector<Aabb> meshBox;

  ~sundog_session() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (wfHostCnt) (void)hipHostFree(wfHostCnt);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
  }

  uint64_t deviceBytes() const {
    const DevBuf* all[] = {&textures, &texels, &materials, &lights, &flames, &instances, &nodes, &tris,
                           &meshes, &positions, &indices, &normals, &uvs, &lightPos, &lightCdf, &envTex,
                           &margCdf, &condCdf, &envSearch, &accum, &albedo, &normal, &rayCounter, &workCounters, &spill,
                           &sampleBuf, &wfState, &wfClosest, &wfShadow, &wfBuckets, &wfCnt};
    uint64_t b = 0;
    for (const DevBuf* d : all) b += d->bytes;
    return b;
  }

  // Spill rows cover the deepest possible stack beyond the smallest LDS
  // stack any kernel of this session uses.
  int spillRowsFor(int ldsEntries) const { return std::max(0, stackNeed - std::min(ldsEntries, stackLdsEntries())); }

  int pickWaves(bool& trial) {
    trial = false;
    return waves != 0 ? waves : autoPick.pick(trial);
  }
  void recordTrial(int w, double msPerPath) { autoPick.record(w, msPerPath); }
  void resetAuto() {
    autoPick.reset();
    groupPick.reset();
  }
  // Sample groups for a pass: fixed by tuning, the frame-size rule while the
  // occupancy variant is being timed or for the work-counting pass, else the
  // group picker's choice or candidate (gTrial: the pass is a group trial).
  int pickGroups(size_t tiles, int count, bool wavesTrial, bool& gTrial) {
    gTrial = false;
    if (count > kPoolSamples) return 1;
    if (sampleGroups > 0) return std::min(sampleGroups, count);
    if (wavesTrial || countWork || (waves == 0 && autoPick.chosen == 0))
      return sampleGroupsFor(tiles, count, cuCount);
    return groupPick.pick(gTrial, count);
  }
  int currentGroups() const { return sampleGroups > 0 ? sampleGroups : groupPick.chosen; }
  // The variant renders use now (the fixed one, the auto choice, or 0 while
  // auto mode is still timing candidates).
  int currentWaves() const { return waves != 0 ? waves : autoPick.chosen; }

  // Per-sample result region of every wave of a render grid (the
  // shared-sample megakernel), grown on demand.
  void ensureSampleBuf(size_t waves) {
    const size_t need = waves * (size_t)kPoolPlanes * kPoolSamples * 64 * sizeof(float);
    if (need > sampleBuf.bytes) {
      HIP_OK(hipStreamSynchronize(stream));
      sampleBuf.alloc(need);
    }
    P.sampleBuf = sampleBuf.as<float>();
    P.sampleBufWaves = (int64_t)waves;
  }

  void ensureSpill(size_t threads) {
    spillRows = spillRowsFor(P.stackLds);
    if (spillRows <= 0) {
      if (!spill.p) spill.alloc(16);
      P.spill = spill.as<uint32_t>();
      P.spillStride = 0;
      P.spillRows = 0;
      return;
    }
    const size_t need = (size_t)spillRows * threads * kStackEntryBytes;
    if (need > spill.bytes) {
      HIP_OK(hipStreamSynchronize(stream));
      spill.alloc(need);
    }
    P.spill = spill.as<uint32_t>();
    P.spillStride = (int)threads;
    P.spillRows = spillRows;
  }

  void build(const Scene& s) {
    auto t0 = std::chrono::steady_clock::now();
    prepareScene(s, prep);
    const Scene& S = prep.scene;
    HIP_OK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    HIP_OK(hipEventCreate(&ev0));
    HIP_OK(hipEventCreate(&ev1));

    // textures + texel pool
    std::vector<DevTexture> dt;
    std::vector<float4> tx;
    for (size_t i = 0; i < S.textures.size(); i++) {
      const Texture& t = S.textures[i];
      DevTexture d{};
      d.kind = t.kind;
      d.a[0] = t.a.x; d.a[1] = t.a.y; d.a[2] = t.a.z;
      d.b[0] = t.b.x; d.b[1] = t.b.y; d.b[2] = t.b.z;
      d.scale[0] = t.scale.x;
      d.scale[1] = t.scale.y;
      d.width = t.width;
      d.imgOffset = -1;
      if (t.kind == TX_IMAGE) {
        const HostImage& im = prep.texImages[i];
        d.imgOffset = (int)tx.size();
        d.imgW = im.width;
        d.imgH = im.height;
        for (size_t k = 0; k < (size_t)im.width * im.height; k++)
          tx.push_back(make_float4(im.rgba[4 * k], im.rgba[4 * k + 1], im.rgba[4 * k + 2], im.rgba[4 * k + 3]));
      }
      dt.push_back(d);
    }
    textures.upload(dt, stream);
    texels.upload(tx, stream);

    std::vector<DevMaterial> dm;
    for (const Material& m : S.materials) {
      DevMaterial d{};
      d.kind = m.kind;
      d.texId = m.texId;
      d.color[0] = m.color.x; d.color[1] = m.color.y; d.color[2] = m.color.z;
      d.roughness = m.roughness;
      d.ior = m.ior;
      d.intensity = m.intensity;
      d.twoSided = m.twoSided;
      d.absorb[0] = m.absorb.x; d.absorb[1] = m.absorb.y; d.absorb[2] = m.absorb.z;
      d.waveAmp = m.waveAmp;
      d.waveFreq = m.waveFreq;
      d.waveOctaves = m.waveOctaves;
      dm.push_back(d);
    }
    materials.upload(dm, stream);

    // mesh pools
    std::vector<float> pos, nrm, uv;
    std::vector<uint32_t> idx;
    std::vector<DevMesh> dmesh;
    for (const HostMesh& m : prep.meshes) {
      DevMesh d{};
      d.posOffset = (int)(pos.size() / 3);
      d.idxOffset = (int)idx.size();
      d.numTris = (int)m.numTris();
      d.numVerts = (int)m.positions.size();
      for (const vec3& p : m.positions) { pos.push_back(p.x); pos.push_back(p.y); pos.push_back(p.z); }
      idx.insert(idx.end(), m.indices.begin(), m.indices.end());
      d.nrmOffset = -1;
      if (!m.normals.empty()) {
        d.nrmOffset = (int)(nrm.size() / 3);
        for (const vec3& n : m.normals) { nrm.push_back(n.x); nrm.push_back(n.y); nrm.push_back(n.z); }
      }
      d.uvOffset = -1;
      if (!m.uvs.empty()) {
        d.uvOffset = (int)(uv.size() / 2);
        for (const vec2& t : m.uvs) { uv.push_back(t.x); uv.push_back(t.y); }
      }
      dmesh.push_back(d);
    }

    // lights (+ mesh-light pools)
    std::vector<DevLight> dl;
    std::vector<float> lpos, lcdf;
    for (size_t i = 0; i < S.lights.size(); i++) {
      const Light& l = S.lights[i];
      DevLight d{};
      d.kind = l.kind;
      d.p[0] = l.p.x; d.p[1] = l.p.y; d.p[2] = l.p.z;
      d.u[0] = l.u.x; d.u[1] = l.u.y; d.u[2] = l.u.z;
      d.v[0] = l.v.x; d.v[1] = l.v.y; d.v[2] = l.v.z;
      d.n[0] = l.n.x; d.n[1] = l.n.y; d.n[2] = l.n.z;
      d.L[0] = l.L.x; d.L[1] = l.L.y; d.L[2] = l.L.z;
      d.radius = l.radius;
      d.area = l.area;
      d.dir[0] = l.dir.x; d.dir[1] = l.dir.y; d.dir[2] = l.dir.z;
      d.twoSided = l.twoSided;
      d.texId = l.texId;
      d.flameId = l.flameId;
      d.mPosOffset = d.mIdxOffset = d.mUvOffset = d.mCdfOffset = -1;
      d.mNumTris = 0;
      d.mNgSign = 1.0f;
      dl.push_back(d);
    }
    for (const MeshLightPatch& p : prep.meshLights) {
      DevLight& d = dl[p.lightId];
      const Object& o = S.objects[p.objectId];
      d.mPosOffset = (int)(lpos.size() / 3);
      for (const vec3& w : p.worldPos) { lpos.push_back(w.x); lpos.push_back(w.y); lpos.push_back(w.z); }
      d.mIdxOffset = dmesh[o.meshId].idxOffset;
      d.mUvOffset = dmesh[o.meshId].uvOffset;
      d.mCdfOffset = (int)lcdf.size();
      lcdf.insert(lcdf.end(), p.cdf.begin(), p.cdf.end());
      d.mNumTris = (int)(p.cdf.size() - 1);
      d.area = p.area;
      d.mNgSign = p.ngSign;
    }
    lights.upload(dl, stream);
    lightPos.upload(lpos, stream);
    lightCdf.upload(lcdf, stream);

    std::vector<DevFlame> df;
    for (const Flame& f : S.flames) {
      DevFlame d{};
      d.base[0] = f.base.x; d.base[1] = f.base.y; d.base[2] = f.base.z;
      d.height = f.height;
      d.radius = f.radius;
      d.intensity = f.intensity;
      d.sigma = f.sigma;
      d.noiseScale = f.noiseScale;
      d.seed = f.seed;
      d.octaves = f.octaves;
      d.step = f.step;
      d.bounds = f.bounds;
      df.push_back(d);
    }
    flames.upload(df, stream);

    // ---- acceleration structures ----
    std::vector<BvhNode4> allNodes;
    std::vector<float4> triPool;
    std::vector<uint32_t> blasRoot(prep.meshes.size(), 0);
    std::vector<Aabb> meshBounds(prep.meshes.size());
    std::vector<Bvh4> blases(prep.meshes.size());
    int blasStack = 0;
    std::vector<bool> meshUsed(prep.meshes.size(), false);
    for (const Object& o : S.objects)
      if (o.geomKind == GK_MESH) meshUsed[o.meshId] = true;
    for (size_t mi = 0; mi < prep.meshes.size(); mi++) {
      if (!meshUsed[mi]) continue;
      const HostMesh& m = prep.meshes[mi];
      std::vector<Aabb> boxes(m.numTris());
      for (size_t t = 0; t < m.numTris(); t++) {
        boxes[t].grow(m.positions[m.indices[3 * t]]);
        boxes[t].grow(m.positions[m.indices[3 * t + 1]]);
        boxes[t].grow(m.positions[m.indices[3 * t + 2]]);
      }
      blases[mi] = buildBvh4(boxes, 4);
      meshBounds[mi] = blases[mi].bounds;
      blasStack = std::max(blasStack, blases[mi].stackBound);
    }

    // top level over all objects
    std::vector<Aabb> instBoxes = objectWorldBoxes(S, prep.meshes, meshBounds);
    int tlasStack = 0;
    if (!instBoxes.empty()) {
      Bvh4 tlas = buildBvh4(instBoxes, 1);
      tlasStack = tlas.stackBound;
      allNodes = tlas.nodes;
      relocate(allNodes, 0, allNodes.size(), 0, 0, &tlas.primOrder);
    } else {
      BvhNode4 nd;
      for (int i = 0; i < 4; i++) {
        nd.lox[i] = nd.loy[i] = nd.loz[i] = nd.hix[i] = nd.hiy[i] = nd.hiz[i] = 3e38f;
        nd.child[i] = (int32_t)kRefEmpty;
        nd.count[i] = -1;
      }
      allNodes.push_back(nd);
    }
    tlasNodes = (int)allNodes.size();
    for (size_t mi = 0; mi < prep.meshes.size(); mi++) {
      if (!meshUsed[mi]) continue;
      const Bvh4& b = blases[mi];
      uint32_t nodeBase = (uint32_t)allNodes.size();
      uint32_t primBase = (uint32_t)(triPool.size() / 3);
      allNodes.insert(allNodes.end(), b.nodes.begin(), b.nodes.end());
      relocate(allNodes, nodeBase, b.nodes.size(), nodeBase, primBase, nullptr);
      blasRoot[mi] = nodeBase;
      const HostMesh& m = prep.meshes[mi];
      for (uint32_t t : b.primOrder) {
        vec3 a = m.positions[m.indices[3 * t]], bb = m.positions[m.indices[3 * t + 1]],
             c = m.positions[m.indices[3 * t + 2]];
        float4 v0 = make_float4(a.x, a.y, a.z, 0.0f);
        int32_t tid = (int32_t)t;
        memcpy(&v0.w, &tid, 4);
        triPool.push_back(v0);
        triPool.push_back(make_float4(bb.x, bb.y, bb.z, 0.0f));
        triPool.push_back(make_float4(c.x, c.y, c.z, 0.0f));
      }
    }
    blasNodes = (int)allNodes.size() - tlasNodes;
    stackNeed = tlasStack + 1 + blasStack;
    P.stackLds = stackLdsDefault();
    spillRows = spillRowsFor(P.stackLds);

    std::vector<DevInstance> di;
    for (size_t i = 0; i < S.objects.size(); i++) {
      const Object& o = S.objects[i];
      DevInstance d{};
      memcpy(d.inv, prep.inverses[i].m, sizeof d.inv);
      d.geomKind = o.geomKind;
      d.meshId = o.meshId;
      d.blasRoot = o.geomKind == GK_MESH ? blasRoot[o.meshId] : 0;
      d.mats = (uint32_t)o.matFront | ((uint32_t)o.matBack << 16);
      d.lightId = o.lightId;
      d.cutoutTex = o.cutoutTexId;
      bool masked = o.matFront == kMatNone || o.matBack == kMatNone || o.cutoutTexId >= 0;
      bool transmissive = materialTransmissive(S, o.matFront) || materialTransmissive(S, o.matBack);
      d.flags = (masked ? INST_MASKED : 0) | (transmissive ? INST_TRANSMISSIVE : 0);
      di.push_back(d);
    }
    numInstances = (int)di.size();
    instances.upload(di, stream);
    instHost = di;
    meshBox = meshBounds;
    meshPosOff.clear();
    meshIdxOff.clear();
    for (const DevMesh& dm : dmesh) {
      meshPosOff.push_back(dm.posOffset);
      meshIdxOff.push_back(dm.idxOffset);
    }
    nodes.upload(allNodes, stream);
    tris.upload(triPool, stream);
    meshes.upload(dmesh, stream);
    positions.upload(pos, stream);
    indices.upload(idx, stream);
    normals.upload(nrm, stream);
    uvs.upload(uv, stream);

    // environment
    int envImportance = 0;
    if (S.bgKind == BG_ENVMAP) {
      std::vector<float> marg, cond;
      bool ok = buildEnvCdf(prep.env, marg, cond);
      if (!ok && S.env.importance)
        fprintf(stderr, "envmap: %s has zero luminance, falling back to uniform sampling\n",
                S.env.file.c_str());
      envImportance = (S.env.importance && ok) ? 1 : 0;
      std::vector<float4> et((size_t)prep.env.width * prep.env.height);
      for (size_t k = 0; k < et.size(); k++)
        et[k] = make_float4(prep.env.rgba[4 * k], prep.env.rgba[4 * k + 1], prep.env.rgba[4 * k + 2],
                            prep.env.rgba[4 * k + 3]);
      envTex.upload(et, stream);
      margCdf.upload(marg, stream);
      condCdf.upload(cond, stream);
      const int ew = prep.env.width, eh = prep.env.height;
      const CdfSearchShape ms = cdfSearchShape(eh + 1), rs = cdfSearchShape(ew + 1);
      std::vector<float> table(16 * ((size_t)ms.blocks() + (size_t)eh * rs.blocks()));
      bool fits = buildCdfSearch(marg.data(), eh + 1, table.data());
      for (int r = 0; fits && r < eh; r++)
        fits = buildCdfSearch(&cond[(size_t)r * (ew + 1)], ew + 1, &table[16 * ((size_t)ms.blocks() + (size_t)r * rs.blocks())]);
      if (fits) {
        envSearch.upload(table, stream);
        envSearchShape[0] = ms.blocks();
        envSearchShape[1] = ms.mid;
        envSearchShape[2] = rs.blocks();
        envSearchShape[3] = rs.mid;
      }
    }

    // film
    size_t npx = (size_t)S.render.width * S.render.height;
    accum.alloc(npx * 16);
    albedo.alloc(npx * 16);
    normal.alloc(npx * 16);
    rayCounter.alloc(8);
    workCounters.alloc(32);

    // launch parameters
    LaunchParams& p = P;
    p.width = S.render.width;
    p.height = S.render.height;
    p.sppTotal = S.render.spp;
    p.maxDepth = S.render.maxDepth;
    p.clampVal = S.render.clampVal;
    p.seed = S.render.seed;
    p.countRays = 0;
    p.transparentShadows = S.render.transparentShadows ? 1 : 0;
    p.numLights = (int)S.lights.size();
    p.numFlames = (int)S.flames.size();
    p.accum = accum.as<float4>();
    p.albedo = albedo.as<float4>();
    p.normal = normal.as<float4>();
    p.rayCounter = rayCounter.as<unsigned long long>();
    p.workCounters = workCounters.as<unsigned long long>();
    Camera c = makeCamera(S.camera, S.render.width, S.render.height);
    auto st3 = [](float* d, vec3 v) { d[0] = v.x; d[1] = v.y; d[2] = v.z; };
    st3(p.cam.origin, c.origin);
    st3(p.cam.lowerLeft, c.lowerLeft);
    st3(p.cam.horizontal, c.horizontal);
    st3(p.cam.vertical, c.vertical);
    st3(p.cam.u, c.u);
    st3(p.cam.v, c.v);
    st3(p.cam.w, c.w);
    p.cam.lensRadius = c.lensRadius;
    p.bgKind = S.bgKind;
    st3(p.bgA, S.bgA);
    st3(p.bgB, S.bgB);
    p.hasEnv = S.bgKind == BG_ENVMAP ? 1 : 0;
    p.envTex = envTex.as<float4>();
    p.margCdf = margCdf.as<float>();
    p.condCdf = condCdf.as<float>();
    p.envW = prep.env.width;
    p.envH = prep.env.height;
    p.envRotate = S.env.rotateDeg * kPi / 180.0f;
    p.envIntensity = S.env.intensity;
    p.envImportance = envImportance;
    p.envSearch = envSearch.p ? envSearch.as<float4>() : nullptr;
    p.envSearchMargBlocks = envSearchShape[0];
    p.envSearchMargMid = envSearchShape[1];
    p.envSearchRowBlocks = envSearchShape[2];
    p.envSearchRowMid = envSearchShape[3];
    p.textures = textures.as<DevTexture>();
    p.texels = texels.as<float4>();
    p.materials = materials.as<DevMaterial>();
    p.lights = lights.as<DevLight>();
    p.flames = flames.as<DevFlame>();
    p.instances = instances.as<DevInstance>();
    p.nodes = nodes.as<BvhNode4>();
    p.tris = tris.as<float4>();
    p.meshes = meshes.as<DevMesh>();
    p.positions = positions.as<float>();
    p.indices = indices.as<uint32_t>();
    p.normals = normals.as<float>();
    p.uvs = uvs.as<float>();
    p.lightPos = lightPos.as<float>();
    p.lightCdf = lightCdf.as<float>();
    p.tlasRoot = 0;
    p.instFlags = 0;
    for (const DevInstance& d : instHost) p.instFlags |= d.flags;
    reset();
    HIP_OK(hipStreamSynchronize(stream));
    buildMs = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }

  void reset() {
    size_t npx = (size_t)P.width * P.height;
    HIP_OK(hipMemsetAsync(accum.p, 0, npx * 16, stream));
    HIP_OK(hipMemsetAsync(albedo.p, 0, npx * 16, stream));
    HIP_OK(hipMemsetAsync(normal.p, 0, npx * 16, stream));
    HIP_OK(hipMemsetAsync(rayCounter.p, 0, 8, stream));
    HIP_OK(hipMemsetAsync(workCounters.p, 0, 32, stream));
    renderMs = 0.0;
    paths = 0;
  }

  // Rebuilds both levels with the GPU LBVH (lbvh.hip) from the resident
  // mesh pools: the top level first (instance boxes from the meshes' bounds,
  // one instance per leaf), then every used mesh's BLAS (up to 4 triangles
  // per leaf, triangles re-emitted in leaf order). Closest hits do not depend
  // on the tree, so films stay bit-identical wherever shadow rays are opaque.
  // GPU rebuild of the BVH. sahTop: the top level is built with the host's
  // binned SAH over the instance boxes the GPU computed (4,096 boxes take
  // about a millisecond), the mesh levels on the GPU. method: 0 Karras'
  // LBVH, 1 PLOC clustering with plocRadius (both collapsed to the same
  // 4-wide nodes).
  void buildBvhGpu(bool sahTop = false, int method = 0) {
    HIP_OK(hipStreamSynchronize(stream));
    auto t0 = std::chrono::steady_clock::now();
    const Scene& S = prep.scene;
    // world boxes: quadrics from their transformed boxes on the host, mesh
    // instances from their transformed vertices on the GPU
    std::vector<float4> ib;
    std::vector<InstanceBoxJob> jobs(S.objects.size());
    for (size_t i = 0; i < S.objects.size(); i++) {
      const Object& o = S.objects[i];
      InstanceBoxJob& jb = jobs[i];
      memcpy(jb.xf, o.xform.m, sizeof(jb.xf));
      jb.posOff = jb.numVerts = 0;
      Aabb w;
      if (o.geomKind == GK_MESH) {
        jb.posOff = meshPosOff[o.meshId];
        jb.numVerts = (int)prep.meshes[o.meshId].positions.size();
        if (jb.numVerts == 0) w = transformBox(o.xform, meshBox[o.meshId]);
      } else {
        w = transformBox(o.xform, quadricBox(o.geomKind));
      }
      ib.push_back(make_float4(w.lo.x, w.lo.y, w.lo.z, 0.0f));
      ib.push_back(make_float4(w.hi.x, w.hi.y, w.hi.z, 0.0f));
    }
    DevBuf instBoxes, tlas, tlasKeys, boxJobs;
    instBoxes.upload(ib, stream);
    boxJobs.upload(jobs, stream);
    HIP_OK(launchInstanceBoxes(positions.as<float>(), boxJobs.as<InstanceBoxJob>(), (int)jobs.size(),
                               instBoxes.as<float4>(), stream));
    int tlasStack = 0;
    int nT = 0;
    if (sahTop) {
      std::vector<float4> hb(ib.size());
      HIP_OK(hipMemcpyAsync(hb.data(), instBoxes.p, hb.size() * sizeof(float4), hipMemcpyDeviceToHost, stream));
      HIP_OK(hipStreamSynchronize(stream));
      std::vector<Aabb> boxes(S.objects.size());
      for (size_t i = 0; i < boxes.size(); i++) {
        boxes[i].lo = vec3{hb[2 * i].x, hb[2 * i].y, hb[2 * i].z};
        boxes[i].hi = vec3{hb[2 * i + 1].x, hb[2 * i + 1].y, hb[2 * i + 1].z};
      }
      Bvh4 top = buildBvh4(boxes, 1);
      std::vector<BvhNode4> tn = top.nodes;
      relocate(tn, 0, tn.size(), 0, 0, &top.primOrder);
      tlas.upload(tn, stream);
      nT = (int)tn.size();
      tlasStack = top.stackBound;
    } else {
      nT = lbvhBuild(instBoxes.as<float4>(), (int)S.objects.size(), 1, true, 0, 0, tlas, tlasKeys, tlasStack, stream,
                     method, plocRadius);
    }
    struct Blas {
      DevBuf nodes, keys;
      int num = 0, stack = 0;
      uint32_t nodeBase = 0, primBase = 0;
      int mesh = -1, tris = 0;
    };
    std::vector<Blas> bl;
    uint32_t nodeBase = (uint32_t)nT, primBase = 0;
    int blasStack = 0;
    std::vector<uint32_t> root(prep.meshes.size(), 0);
    for (size_t mi = 0; mi < prep.meshes.size(); mi++) {
      bool used = false;
      for (const Object& o : S.objects) used = used || (o.geomKind == GK_MESH && o.meshId == (int)mi);
      if (!used) continue;
      const int n = (int)prep.meshes[mi].numTris();
      DevBuf boxes;
      boxes.alloc((size_t)n * 2 * sizeof(float4));
      const float* pos = positions.as<float>() + 3 * (size_t)meshPosOff[mi];
      const uint32_t* idx = indices.as<uint32_t>() + meshIdxOff[mi];
      HIP_OK(launchTriBoxes(pos, idx, n, boxes.as<float4>(), stream));
      bl.emplace_back();
      Blas& b = bl.back();
      b.mesh = (int)mi;
      b.tris = n;
      b.nodeBase = nodeBase;
      b.primBase = primBase;
      b.num = lbvhBuild(boxes.as<float4>(), n, 4, false, nodeBase, primBase, b.nodes, b.keys, b.stack, stream, method,
                        plocRadius);
      root[mi] = nodeBase;
      nodeBase += (uint32_t)b.num;
      primBase += (uint32_t)n;
      blasStack = std::max(blasStack, b.stack);
    }
    DevBuf newNodes, newTris;
    newNodes.alloc((size_t)nodeBase * sizeof(BvhNode4));
    newTris.alloc(std::max<size_t>(16, (size_t)primBase * 3 * sizeof(float4)));
    HIP_OK(hipMemcpyAsync(newNodes.p, tlas.p, (size_t)nT * sizeof(BvhNode4), hipMemcpyDeviceToDevice, stream));
    for (Blas& b : bl) {
      HIP_OK(hipMemcpyAsync(newNodes.as<BvhNode4>() + b.nodeBase, b.nodes.p, (size_t)b.num * sizeof(BvhNode4),
                            hipMemcpyDeviceToDevice, stream));
      const float* pos = positions.as<float>() + 3 * (size_t)meshPosOff[b.mesh];
      const uint32_t* idx = indices.as<uint32_t>() + meshIdxOff[b.mesh];
      HIP_OK(launchEmitTris(b.keys.as<uint64_t>(), b.tris, pos, idx, newTris.as<float4>() + 3 * (size_t)b.primBase,
                            stream));
    }
    for (size_t i = 0; i < S.objects.size(); i++)
      if (S.objects[i].geomKind == GK_MESH) instHost[i].blasRoot = root[S.objects[i].meshId];
    instances.upload(instHost, stream);
    HIP_OK(hipStreamSynchronize(stream));
    nodes = std::move(newNodes);
    tris = std::move(newTris);
    P.nodes = nodes.as<BvhNode4>();
    P.tris = tris.as<float4>();
    P.instances = instances.as<DevInstance>();
    tlasNodes = nT;
    blasNodes = (int)nodeBase - nT;
    stackNeed = tlasStack + 1 + blasStack;
    spillRows = spillRowsFor(P.stackLds);
    bvhBuilder = method == 1 ? (sahTop ? 4 : 3) : (sahTop ? 2 : 1);
    bvhGpuMs = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }

  // While the occupancy variant is being timed (auto mode), a pass over a
  // large frame runs as trial sub-passes of a quarter of its samples, so the
  // search costs about one pass's worth of samples instead of several whole
  // passes spent partly on slow variants (a 64-spp render in 16-spp passes
  // would otherwise end before it uses the variant it found); once a variant
  // is chosen, the rest of the pass is one launch. Progressive passes give
  // the film of a single pass bit for bit, so the split changes no result.
  // Frames whose quarter pass is below kTrialMinPaths paths are not split:
  // such launches are too short to time reliably.
  static constexpr double kTrialMinPaths = 2.0e6;
  void renderPass(int offset, int count) {
    if (count <= 0) return;
    if (offset < 0) throw std::runtime_error("render: negative sample offset");
    int done = 0;
    const int sub = (count + 3) / 4;
    if (kernel != SUNDOG_KERNEL_WAVEFRONT && !countWork && count >= 4 &&
        (double)P.width * P.height * sub >= kTrialMinPaths)
      while (done < count && waves == 0 && autoPick.chosen == 0) {
        const int c = std::min(sub, count - done);
        launchPass(offset + done, c);
        done += c;
      }
    if (done < count) launchPass(offset + done, count - done);
  }

  void launchPass(int offset, int count) {
    if (kernel == SUNDOG_KERNEL_WAVEFRONT) {
      wavefrontPass(offset, count);
      return;
    }
    size_t gx = (P.width + 15) / 16, gy = (P.height + 15) / 16;
    if (cuCount == 0) {
      int dev = 0;
      HIP_OK(hipGetDevice(&dev));
      HIP_OK(hipDeviceGetAttribute(&cuCount, hipDeviceAttributeMultiprocessorCount, dev));
    }
    bool trial = false, gTrial = false;
    const int w = countWork ? 2 : pickWaves(trial);
    const int groups = pickGroups(gx * gy, count, trial, gTrial);
    ensureSpill(gx * gy * groups * renderBlockThreads());
    ensureSampleBuf(gx * gy * (renderBlockThreads() / 64));
    P.sampleOffset = offset;
    P.sppThisLaunch = count;
    HIP_OK(hipEventRecord(ev0, stream));
    HIP_OK(launchRender(P, countWork, w, groups, stream));
    HIP_OK(hipEventRecord(ev1, stream));
    HIP_OK(hipEventSynchronize(ev1));
    float ms = 0.0f;
    HIP_OK(hipEventElapsedTime(&ms, ev0, ev1));
    if (trial) recordTrial(w, ms / ((double)P.width * P.height * count));
    if (gTrial) groupPick.record(groups, ms / ((double)P.width * P.height * count));
    renderMs += ms;
    paths += (uint64_t)P.width * P.height * count;
  }

  // Wavefront pass: shade -> (closest-hit queries, shadow queries) rounds
  // until no slot has a query left. The host reads the queue sizes after
  // each round to know when the pass is over.
  void wavefrontPass(int offset, int count) {
    if (countWork) throw std::runtime_error("render: work counting needs the megakernel");
    const int slots = P.width * P.height;
    if (slots != wfSlots) {
      HIP_OK(hipStreamSynchronize(stream));
      wfState.alloc((size_t)kWfPlanes * slots * sizeof(float4));
      wfClosest.alloc((size_t)slots * 4);
      wfShadow.alloc((size_t)slots * 4);
      wfBuckets.alloc((size_t)kWfKeys * slots * 4);
      wfCnt.alloc(sizeof(WfCounters));
      if (!wfHostCnt) HIP_OK(hipHostMalloc((void**)&wfHostCnt, sizeof(WfCounters), hipHostMallocDefault));
      wfSlots = slots;
    }
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    int cus = 0;
    HIP_OK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int bShade = cus * wfBlocksPerCU(0, 0), bClosest = cus * wfBlocksPerCU(1, wfWaves),
              bShadow = cus * wfBlocksPerCU(2, wfWaves);
    ensureSpill((size_t)std::max(bClosest, bShadow) * renderBlockThreads());
    P.sampleOffset = offset;
    P.sppThisLaunch = count;
    WfParams W{};
    W.state = wfState.as<float4>();
    W.numSlots = slots;
    W.closestQ = wfClosest.as<uint32_t>();
    W.shadowQ = wfShadow.as<uint32_t>();
    W.buckets = wfBuckets.as<uint32_t>();
    W.cnt = wfCnt.as<WfCounters>();
    W.rayCounter = P.countRays ? rayCounter.as<unsigned long long>() : nullptr;
    HIP_OK(hipEventRecord(ev0, stream));
    HIP_OK(launchWfClear(W.cnt, 0, stream));
    HIP_OK(launchWfClear(W.cnt, 1, stream));
    W.first = 1;
    HIP_OK(launchWfShade(P, W, std::min(bShade, (slots + kWfBlock - 1) / kWfBlock), stream));
    W.first = 0;
    wfIterations = 0;
    while (true) {
      HIP_OK(launchWfClear(W.cnt, 0, stream));
      HIP_OK(launchWfClosest(P, W, bClosest, wfWaves, stream));
      HIP_OK(launchWfShadow(P, W, bShadow, wfWaves, stream));
      HIP_OK(launchWfClear(W.cnt, 1, stream));
      HIP_OK(hipMemcpyAsync(wfHostCnt, W.cnt, sizeof(WfCounters), hipMemcpyDeviceToHost, stream));
      HIP_OK(hipStreamSynchronize(stream));
      uint64_t pending = 0;
      for (int k = 0; k < kWfKeys; k++) pending += ((const WfCounters*)wfHostCnt)->bucketN[k];
      if (pending == 0) break;
      wfIterations++;
      int blocks = (int)std::min<uint64_t>((uint64_t)bShade, (pending + kWfBlock - 1) / kWfBlock);
      HIP_OK(launchWfShade(P, W, blocks, stream));
    }
    HIP_OK(hipEventRecord(ev1, stream));
    HIP_OK(hipEventSynchronize(ev1));
    float ms = 0.0f;
    HIP_OK(hipEventElapsedTime(&ms, ev0, ev1));
    renderMs += ms;
    paths += (uint64_t)P.width * P.height * count;
  }
};
