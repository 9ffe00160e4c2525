// Device-resident scene tables and launch parameters for the gfx950 kernels.
//
// The information content follows the reference's descriptors
// (device/params.h:20-145, HitRecordData :148-159); the layout is this
// library's own: fixed 16-byte-aligned records, all scene arrays in HBM
// pools, and an instance table that carries the world->object transform
// (the reference leaves that inverse to OptiX).
#pragma once

#include <stdint.h>

#include "bvh.h"
#include "kinds.h"
#include "sd_math.h"

namespace sdk {

constexpr uint32_t kRefLeaf = 0x80000000u;
constexpr uint32_t kRefEmpty = 0xFFFFFFFFu;
constexpr uint32_t kRefMarker = 0xFFFFFFFEu;  // "return to the top level"
constexpr int kMedMax = 4;

struct alignas(16) DevTexture {
  int32_t kind;
  float a[3];
  float b[3];
  float scale[2];
  float width;
  int32_t imgOffset;  // texel index into the float4 texel pool
  int32_t imgW, imgH;
  int32_t pad[2];
};

struct alignas(16) DevMaterial {
  int32_t kind, texId;
  float color[3];
  float roughness, ior, intensity;
  int32_t twoSided;
  float absorb[3];
  float waveAmp, waveFreq;
  int32_t waveOctaves;
  int32_t pad;
};

struct alignas(16) DevLight {
  int32_t kind;
  float p[3], u[3], v[3], n[3], L[3];
  float radius, area;
  float dir[3];
  int32_t twoSided, texId, flameId;
  // mesh lights
  int32_t mPosOffset;   // world-space vertex pool offset (vec3)
  int32_t mIdxOffset;   // index pool offset (uint32 triplets, mesh-local vertex ids)
  int32_t mUvOffset;    // uv pool offset or -1
  int32_t mCdfOffset;   // cdf pool offset (numTris + 1 floats)
  int32_t mNumTris;
  float mNgSign;
  int32_t pad[2];
};

struct alignas(16) DevFlame {
  float base[3];
  float height, radius, intensity, sigma, noiseScale;
  uint32_t seed;
  int32_t octaves;
  float step;
  int32_t bounds;
};

enum : int32_t { INST_MASKED = 1, INST_TRANSMISSIVE = 2 };

struct alignas(16) DevInstance {
  float inv[12];        // world -> object, row-major 3x4
  int32_t geomKind;
  uint32_t blasRoot;    // packed node reference of the mesh's BLAS root
  int32_t meshId;
  uint32_t mats;        // front | back << 16 (0xFFFF = pass-through)
  int32_t lightId;
  int32_t cutoutTex;
  int32_t flags;        // INST_*
  int32_t pad;
};
static_assert(sizeof(DevInstance) == 80, "instance layout");

struct alignas(16) DevMesh {
  int32_t posOffset, idxOffset, nrmOffset, uvOffset;  // pools; -1 = absent
  int32_t numTris, numVerts;
  int32_t pad[2];
};

struct CameraDev {
  float origin[3], lowerLeft[3], horizontal[3], vertical[3], u[3], v[3], w[3];
  float lensRadius;
};

struct LaunchParams {
  int32_t width, height, sppTotal, sampleOffset, sppThisLaunch, maxDepth;
  int32_t passSpp;  // samples per pass of a graph-replayed request (k_queue_prep)
  float clampVal;
  uint32_t seed;
  int32_t countRays, transparentShadows;
  int32_t numLights, numFlames;

  float4* accum;
  float4* albedo;
  float4* normal;
  unsigned long long* rayCounter;
  unsigned long long* workCounters;  // node visits, triangle tests, quadric tests, instance transforms
  // Per-sample results of the shared-sample megakernel: kPoolPlanes x
  // kPoolSamples x 64 floats per wave of the render grid; sampleBufWaves
  // regions (launchRender refuses a grid with more waves).
  float* sampleBuf;
  int64_t sampleBufWaves;

  CameraDev cam;
  int32_t bgKind;
  float bgA[3], bgB[3];

  const float4* envTex;
  const float* margCdf;
  const float* condCdf;
  int32_t envW, envH;
  float envRotate, envIntensity;
  int32_t envImportance;
  int32_t hasEnv;
  // 16-ary lookup tables of the CDFs (buildCdfSearch): the marginal's
  // table, then one table per row; null when a CDF is too long for three
  // levels (the lookups then bisect the CDFs)
  const float4* envSearch;
  int32_t envSearchMargBlocks, envSearchMargMid;  // 16-float blocks of the marginal table; its mid-level blocks
  int32_t envSearchRowBlocks, envSearchRowMid;    // the same for each row's table

  const DevTexture* textures;
  const float4* texels;
  const DevMaterial* materials;
  const DevLight* lights;
  const DevFlame* flames;
  const DevInstance* instances;
  const BvhNode4* nodes;
  const float4* tris;        // 3 float4 per leaf-ordered triangle; v0.w = triangle id bits
  const DevMesh* meshes;
  const float* positions;    // vec3 pool
  const uint32_t* indices;   // uint32 pool
  const float* normals;      // vec3 pool
  const float* uvs;          // vec2 pool
  const float* lightPos;     // world-space mesh-light vertices (vec3 pool)
  const float* lightCdf;
  uint32_t* spill;           // traversal-stack overflow, 8-byte entries, [entry][thread]
  int32_t spillStride;       // threads covered by one spill row
  int32_t spillRows;
  int32_t stackLds;          // LDS stack entries per lane of the megakernel
  uint32_t tlasRoot;         // packed reference of the top-level root
  int32_t instFlags;         // OR of every instance's flags (INST_MASKED, INST_TRANSMISSIVE)
};

// Persistent tile-queue work item: one tile of one request for one pass.
struct WorkItem {
  int32_t request;
  int32_t tile;   // tile index in row-major order over the request's frame
  int32_t group;  // sample group of the pass (0 when the pass is not split)
  int32_t pad;
};

// The persistent kernel takes the resident scenes' parameters by value (a
// kernel argument, so every scene pointer is known to be global memory) and
// reads only these per-request records from device memory.
constexpr int kMaxScenesPerLaunch = 4;

// Shared-sample megakernel: a wave draws the paths of its 64 pixels from one
// pool of kPoolSamples samples per pixel at a time; each finished path leaves
// kPoolPlanes floats (radiance, first-hit albedo, first-hit normal).
constexpr int kPoolSamples = 16;
constexpr int kPoolPlanes = 9;

struct SceneTable {
  LaunchParams scene[kMaxScenesPerLaunch];
};

struct RequestDev {
  int32_t scene;  // index into the launch's SceneTable
  int32_t width, height, sppTotal, sampleOffset, sppThisLaunch, passSpp;
  // Sample groups of a pass: 1 folds each tile's samples in the kernel;
  // more split a pass's samples over that many work items per tile, which
  // write per-sample results to the pool regions of the batch's group
  // buffer starting at regionBase (tiles x pools per tile of them), folded
  // afterwards by k_fold_groups.
  int32_t groups, regionBase;
  uint32_t seed;
  float4* accum;
  float4* albedo;
  float4* normal;
  CameraDev cam;
};

struct QueueState {
  uint32_t next;  // work-item counter
  uint32_t pass;  // pass counter advanced by the queue-prep kernel
  uint32_t pad[2];
};

// Wavefront variant. Path state lives in kWfPlanes float4 planes of
// numSlots entries (plane j of slot i at state[j * numSlots + i]; one slot
// per pixel), so every state access of a wave is one coalesced 1 KiB
// transaction. Between kernels, slots move through three lists: closest-hit
// queries, shadow queries, and kWfKeys shading buckets (one per material
// kind of the surface to shade, plus misses and new samples) that order the
// shading kernel's work by material.
constexpr int kWfKeys = 16;
constexpr int kWfPlanes = 16;

struct WfCounters {
  uint32_t closestN, shadowN, pad0, pad1;
  uint32_t bucketN[kWfKeys];
};

struct WfParams {
  float4* state;
  int32_t numSlots;
  int32_t first;       // first shading pass of a launch: every slot, no list
  uint32_t* closestQ;  // numSlots entries
  uint32_t* shadowQ;   // numSlots entries
  uint32_t* buckets;   // kWfKeys lists of numSlots entries
  WfCounters* cnt;
  unsigned long long* rayCounter;
};

}  // namespace sdk
