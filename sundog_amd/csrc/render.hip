// gfx950 path-tracing kernels: two-level BVH4 traversal with an LDS stack,
// analytic quadrics and watertight triangles, BSDFs / lights / environment /
// flames, and the path megakernel with its progressive film update.
//
// Path semantics (draw order, MIS weights, clamp and sanitize placement,
// medium stack, transparent shadows) restate the reference raygen program
// device/programs.cu:311-572 and its helpers; each block cites its source.
// Traversal, memory layout and launch shape are this library's own: one
// work-item per pixel, 16x16-pixel workgroups of four 8x8 waves, 4-wide
// nodes fetched as eight 16-byte loads, and a per-lane stack in LDS.
#include <hip/hip_runtime.h>

#include "dev.h"
#include "shade.h"

namespace sdk {

#define DEV __device__ __forceinline__

constexpr int kBlock = 256;
constexpr int kStackLdsMax = 24;  // largest runtime LDS stack (megakernel, persistent kernel)
constexpr int kStackLds = 16;  // 8-byte (reference, entry distance) LDS entries per lane (fixed-size kernels)
// Default LDS stack of the megakernel: with 5 waves per SIMD, 12 entries
// (24 KiB per block) leave LDS for all 20 waves of a CU; measured fastest at
// 1080p on the instanced scene (deeper entries spill to HBM).
constexpr int kStackLdsMegakernel = 12;

// Slab test against the 4 child boxes of one node. Returns per-child entry
// distance, or +inf when missed. Each slab distance carries the rounding of
// the subtraction and the product and the <= 1 ulp error of the hardware
// reciprocal that makes `inv` (4u relative in all); tFar is widened by twice
// that, 8u, and a little more, so boxes stay conservative under rounding
// (Ize, "Robust BVH Ray Traversal", JCGT 2013, with gamma(4) for gamma(3)).
DEV void boxes4(const BvhNode4& nd, vec3 org, vec3 inv, float tmin, float tmax, float dist[4]) {
  const float4* q = reinterpret_cast<const float4*>(&nd);
  float4 lx = q[0], hx = q[1], ly = q[2], hy = q[3], lz = q[4], hz = q[5];
  float lxs[4] = {lx.x, lx.y, lx.z, lx.w}, hxs[4] = {hx.x, hx.y, hx.z, hx.w};
  float lys[4] = {ly.x, ly.y, ly.z, ly.w}, hys[4] = {hy.x, hy.y, hy.z, hy.w};
  float lzs[4] = {lz.x, lz.y, lz.z, lz.w}, hzs[4] = {hz.x, hz.y, hz.z, hz.w};
#pragma unroll
  for (int i = 0; i < 4; i++) {
    float t0x = (lxs[i] - org.x) * inv.x, t1x = (hxs[i] - org.x) * inv.x;
    float t0y = (lys[i] - org.y) * inv.y, t1y = (hys[i] - org.y) * inv.y;
    float t0z = (lzs[i] - org.z) * inv.z, t1z = (hzs[i] - org.z) * inv.z;
    float tn = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fmaxf(fminf(t0z, t1z), tmin));
    float tf = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fminf(fmaxf(t0z, t1z), tmax));
    dist[i] = (tn <= tf * 1.0000006f) ? tn : __builtin_inff();
  }
}

// ---------------------------------------------------------------------------
// traversal stack: LDS for the first kStackLds entries, a global spill
// region (entry-major, one column per thread) beyond.
// ---------------------------------------------------------------------------

// Stack entries are (node reference, entry distance) pairs: a popped entry
// whose box starts beyond the current closest hit is dropped without
// fetching its node. The LDS column is held as an LDS-address-space pointer:
// with a generic pointer the compiler merges the LDS and spill branches of
// pop() into one flat load, which waits on every outstanding global load of
// the lane.
typedef __attribute__((address_space(3))) uint64_t lds_u64;

struct Stack {
  lds_u64* lds;     // this lane's column base: lds[i * kBlock]
  uint64_t* spill;  // this lane's spill base: spill[i * stride]
  int stride;
  int sp;
  int cap;  // LDS entries per lane; deeper entries go to the spill region
  DEV void push(uint32_t ref, float t) {
    uint64_t e = (uint64_t)ref | ((uint64_t)__float_as_uint(t) << 32);
    if (sp < cap) lds[sp * kBlock] = e;
    else spill[(size_t)(sp - cap) * stride] = e;
    sp++;
  }
  DEV uint2 pop() {
    sp--;
    uint64_t e = sp < cap ? lds[sp * kBlock] : spill[(size_t)(sp - cap) * stride];
    return make_uint2((uint32_t)e, (uint32_t)(e >> 32));
  }
};

DEV lds_u64* ldsColumn(uint64_t* sharedArray) { return (lds_u64*)(sharedArray + threadIdx.x); }

// Per-query traversal work, for the algorithmic-FLOP count of the roofline.
struct WorkCount {
  uint32_t nodes = 0, tris = 0, quads = 0, insts = 0;
};

// Upper t bound for primitive candidates: a hair above the current hit so
// that exact ties still reach the (t, instance, primitive) tie rule.
DEV float candBound(float t) {
#pragma clang fp contract(off)
  return t * 1.0000004f + 1e-30f;
}

struct Hit {
  float t;
  int inst;
  int prim;
  float b1, b2;
};

DEV bool closer(float t, int inst, int prim, const Hit& h) {
  return t < h.t || (t == h.t && (inst < h.inst || (inst == h.inst && prim < h.prim)));
}

struct ShadePoint {
  vec3 nFront;  // world unit normal of the canonical front (or ng for emissive mesh lights)
  float u, v;
  bool frontface;
  vec3 ngWorld;
};

// Hit-side resolve (device/programs.cu:135-190): world normal via the
// inverse transpose, uv, front side by the sign of the object-space cosine.
DEV ShadePoint shadePoint(const LaunchParams& P, const DevInstance& in, int prim, float b1, float b2,
                          vec3 oObj, vec3 dObj, float t, bool forShading) {
  ShadePoint sp;
  if (in.geomKind != GK_MESH) {
    vec3 p = oObj + t * dObj;
    vec3 n;
    quadricSurface(in.geomKind, p, n, sp.u, sp.v);
    sp.frontface = dot(n, dObj) < 0.0f;
    sp.nFront = normalize(xfNormal(in.inv, n));
    sp.ngWorld = sp.nFront;
    return sp;
  }
  const DevMesh& m = P.meshes[in.meshId];
  const uint32_t* tri = P.indices + m.idxOffset + 3 * prim;
  uint32_t i0 = tri[0], i1 = tri[1], i2 = tri[2];
  const float* pos = P.positions + 3 * m.posOffset;
  vec3 p0 = ld3i(pos, i0), p1 = ld3i(pos, i1), p2 = ld3i(pos, i2);
  vec3 ngObj = cross(p1 - p0, p2 - p0);
  sp.frontface = dot(ngObj, dObj) < 0.0f;
  vec3 ng = normalize(xfNormal(in.inv, ngObj));
  sp.ngWorld = ng;
  vec3 ns = ng;
  float b0 = 1.0f - b1 - b2;
  if (m.nrmOffset >= 0) {
    const float* nn = P.normals + 3 * m.nrmOffset;
    vec3 n = ld3i(nn, i0) * b0 + ld3i(nn, i1) * b1 + ld3i(nn, i2) * b2;
    ns = normalize(xfNormal(in.inv, n));
    if (dot(ns, ng) < 0.0f) ns = -ns;
  }
  sp.nFront = ns;
  if (m.uvOffset >= 0) {
    const float* uv = P.uvs + 2 * m.uvOffset;
    sp.u = uv[2 * i0] * b0 + uv[2 * i1] * b1 + uv[2 * i2] * b2;
    sp.v = uv[2 * i0 + 1] * b0 + uv[2 * i1 + 1] * b1 + uv[2 * i2 + 1] * b2;
  } else {
    sp.u = b1;
    sp.v = b2;
  }
  if (forShading && in.lightId >= 0) {
    uint32_t side = sp.frontface ? (in.mats & 0xFFFFu) : (in.mats >> 16);
    if (side != 0xFFFFu && P.materials[side].kind == MT_EMISSIVE) sp.nFront = ng;
  }
  return sp;
}

// Radiance-ray mask (device/programs.cu:218-239): pass-through faces and
// alpha cutouts are skipped.
DEV bool maskAccepts(const LaunchParams& P, const DevInstance& in, int prim, float b1, float b2, vec3 oObj,
                     vec3 dObj, float t) {
  if (!(in.flags & INST_MASKED)) return true;
  ShadePoint sp = shadePoint(P, in, prim, b1, b2, oObj, dObj, t, false);
  uint32_t side = sp.frontface ? (in.mats & 0xFFFFu) : (in.mats >> 16);
  if (side == 0xFFFFu) return false;
  if (in.cutoutTex >= 0) {
    float4 c = evalTexture(P, in.cutoutTex, sp.u, sp.v);
    if (c.w < 0.5f) return false;
  }
  return true;
}

struct RayState {
  vec3 o, d, inv;
};

DEV RayState makeRay(vec3 o, vec3 d) {
  RayState r;
  r.o = o;
  r.d = d;
  // only the slab test reads it, whose widening covers the reciprocal's
  // error; a zero component gives an infinite slope as with a division
  r.inv = vec3{__builtin_amdgcn_rcpf(d.x), __builtin_amdgcn_rcpf(d.y), __builtin_amdgcn_rcpf(d.z)};
  return r;
}

// Closest hit with masking. Ties in t resolve to the smaller (instance,
// primitive) so the answer does not depend on traversal order. kMasked =
// false compiles the query for scenes without masked instances (no
// pass-through faces, no cutouts), where maskAccepts always accepts.
template <bool kMasked>
DEV bool traceClosestT(const LaunchParams& P, Stack& st, vec3 o, vec3 d, float tmin, float tmax, Hit& hit,
                       vec3& oObjOut, vec3& dObjOut, WorkCount& wc) {
  hit.t = tmax;
  hit.inst = 0x7fffffff;
  hit.prim = 0x7fffffff;
  hit.b1 = hit.b2 = 0.0f;
  bool found = false;
  RayState world = makeRay(o, d);
  RayState obj = world;
  RayTri rt{2, 0, 1, 0.0f, 0.0f, 1.0f};  // set on entering a mesh instance, before any triangle test
  int curInst = -1;
  bool inBlas = false;
  const int base = st.sp;
  uint32_t cur = P.tlasRoot;
  while (true) {
    if (!(cur & kRefLeaf)) {
      const RayState& r = inBlas ? obj : world;
      wc.nodes++;
      float dist[4];
      boxes4(P.nodes[cur], r.o, r.inv, tmin, hit.t, dist);
      const int32_t* refs = &P.nodes[cur].child[0];
      int4 rc = *reinterpret_cast<const int4*>(refs);
      uint32_t ref[4] = {(uint32_t)rc.x, (uint32_t)rc.y, (uint32_t)rc.z, (uint32_t)rc.w};
      // sort 4 (dist, ref) pairs ascending
#define SD_CSWAP(a, b)                                   \
  if (dist[b] < dist[a]) {                               \
    float td = dist[a]; dist[a] = dist[b]; dist[b] = td; \
    uint32_t tr = ref[a]; ref[a] = ref[b]; ref[b] = tr;  \
  }
      SD_CSWAP(0, 1) SD_CSWAP(2, 3) SD_CSWAP(0, 2) SD_CSWAP(1, 3) SD_CSWAP(1, 2)
#undef SD_CSWAP
      // hit children are sorted first (misses and empty slots carry +inf);
      // the farther ones go on the stack far-to-near with their distances
      const bool h1 = dist[1] < __builtin_inff() && ref[1] != kRefEmpty;
      const bool h2 = dist[2] < __builtin_inff() && ref[2] != kRefEmpty;
      const bool h3 = dist[3] < __builtin_inff() && ref[3] != kRefEmpty;
      if (h3) st.push(ref[3], dist[3]);
      if (h2) st.push(ref[2], dist[2]);
      if (h1) st.push(ref[1], dist[1]);
      if (dist[0] < __builtin_inff() && ref[0] != kRefEmpty) {
        cur = ref[0];
        continue;
      }
    } else if (!inBlas) {
      int ii = (int)(cur & 0x07FFFFFFu);
      const DevInstance& in = P.instances[ii];
      vec3 oo = xfPoint(in.inv, o), dd = xfVector(in.inv, d);
      wc.insts++;
      if (in.geomKind == GK_MESH) {
        st.push(kRefMarker, -__builtin_inff());
        obj = makeRay(oo, dd);
        rt = rayTriSetup(dd);
        curInst = ii;
        inBlas = true;
        cur = in.blasRoot;
        continue;
      }
      wc.quads++;
      float ts[2];
      int n = quadricRoots(in.geomKind, oo, dd, tmin, candBound(hit.t), ts);
      for (int k = 0; k < n; k++) {
        if (closer(ts[k], ii, k, hit) && (!kMasked || maskAccepts(P, in, k, 0.0f, 0.0f, oo, dd, ts[k]))) {
          hit.t = ts[k];
          hit.inst = ii;
          hit.prim = k;
          oObjOut = oo;
          dObjOut = dd;
          found = true;
          break;  // roots ascend: the first accepted one is the closest
        }
      }
    } else {
      int first = (int)(cur & 0x07FFFFFFu);
      int cnt = (int)((cur >> 27) & 0xFu) + 1;
      const DevInstance& in = P.instances[curInst];
      for (int k = 0; k < cnt; k++) {
        wc.tris++;
        const float4* tv = P.tris + 3 * (size_t)(first + k);
        float4 a = tv[0], b = tv[1], c = tv[2];
        float t, b1, b2;
        if (triHit(rt, obj.o, f4xyz(a), f4xyz(b), f4xyz(c), tmin, candBound(hit.t), t, b1, b2)) {
          int prim = __float_as_int(a.w);
          if (closer(t, curInst, prim, hit) && (!kMasked || maskAccepts(P, in, prim, b1, b2, obj.o, obj.d, t))) {
            hit.t = t;
            hit.inst = curInst;
            hit.prim = prim;
            hit.b1 = b1;
            hit.b2 = b2;
            oObjOut = obj.o;
            dObjOut = obj.d;
            found = true;
          }
        }
      }
    }
    // pop; an entry whose box starts beyond the current closest hit cannot
    // hold a closer hit (the entry distance is within 4u of exact, the
    // widening here is 6.7u)
    while (true) {
      if (st.sp == base) return found;
      uint2 e = st.pop();
      cur = e.x;
      if (cur == kRefMarker) {
        inBlas = false;
        continue;
      }
      if (__uint_as_float(e.y) <= hit.t * 1.0000004f) break;
    }
  }
}

// Shadow transmittance along (tmin, tmax) (device/programs.cu:70-85,
// :247-292): opaque => 0; with transparent shadows, each glass / water
// crossing multiplies (1 - F) and adds its signed sigma * t; every primitive
// is visited once, so each crossing counts once. kSurfaces = false compiles
// the query for scenes without masked or transmissive instances, where every
// primitive hit occludes.
template <bool kSurfaces>
DEV vec3 traceShadowT(const LaunchParams& P, Stack& st, vec3 o, vec3 d, float tmax, WorkCount& wc) {
  const float tmin = 0.0f;
  float trF = 1.0f;
  vec3 tau = v3(0.0f);
  RayState world = makeRay(o, d);
  RayState obj = world;
  RayTri rt{2, 0, 1, 0.0f, 0.0f, 1.0f};  // set on entering a mesh instance, before any triangle test
  int curInst = -1;
  bool inBlas = false;
  const int base = st.sp;
  uint32_t cur = P.tlasRoot;
  bool occluded = false;
  while (!occluded) {
    if (!(cur & kRefLeaf)) {
      const RayState& r = inBlas ? obj : world;
      wc.nodes++;
      float dist[4];
      boxes4(P.nodes[cur], r.o, r.inv, tmin, tmax, dist);
      int4 rc = *reinterpret_cast<const int4*>(&P.nodes[cur].child[0]);
      uint32_t ref[4] = {(uint32_t)rc.x, (uint32_t)rc.y, (uint32_t)rc.z, (uint32_t)rc.w};
      int nh = 0;
      uint32_t next = 0;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        if (dist[i] < __builtin_inff() && ref[i] != kRefEmpty) {
          if (nh == 0) next = ref[i];
          else st.push(ref[i], 0.0f);
          nh++;
        }
      }
      if (nh > 0) {
        cur = next;
        continue;
      }
    } else {
      int cnt = 1, first = (int)(cur & 0x07FFFFFFu);
      if (inBlas) cnt = (int)((cur >> 27) & 0xFu) + 1;
      if (!inBlas) {
        const DevInstance& in = P.instances[first];
        vec3 oo = xfPoint(in.inv, o), dd = xfVector(in.inv, d);
        wc.insts++;
        if (in.geomKind == GK_MESH) {
          st.push(kRefMarker, 0.0f);
          obj = makeRay(oo, dd);
          rt = rayTriSetup(dd);
          curInst = first;
          inBlas = true;
          cur = in.blasRoot;
          continue;
        }
        obj.o = oo;
        obj.d = dd;
        curInst = first;
      }
      const DevInstance& in = P.instances[curInst];
      for (int k = 0; k < cnt * (inBlas ? 1 : 2) && !occluded; k++) {
        float t, b1 = 0.0f, b2 = 0.0f;
        int prim;
        bool h;
        if (inBlas) {
          wc.tris++;
          const float4* tv = P.tris + 3 * (size_t)(first + k);
          float4 a = tv[0], b = tv[1], c = tv[2];
          h = triHit(rt, obj.o, f4xyz(a), f4xyz(b), f4xyz(c), tmin, tmax, t, b1, b2);
          prim = __float_as_int(a.w);
        } else {
          if (k == 0) wc.quads++;
          float ts[2];
          int n = quadricRoots(in.geomKind, obj.o, obj.d, tmin, tmax, ts);
          h = k < n;
          t = h ? ts[k] : 0.0f;
          prim = k;
        }
        if (!h) continue;
        if (!kSurfaces) {
          occluded = true;
          break;
        }
        uint32_t side;
        ShadePoint sp;
        bool needSp = (in.flags & (INST_MASKED | INST_TRANSMISSIVE)) != 0;
        if (needSp) {
          sp = shadePoint(P, in, prim, b1, b2, obj.o, obj.d, t, false);
          side = sp.frontface ? (in.mats & 0xFFFFu) : (in.mats >> 16);
        } else {
          side = in.mats & 0xFFFFu;  // opaque object: both faces occlude
        }
        if (side == 0xFFFFu) continue;
        if (in.cutoutTex >= 0) {
          float4 c = evalTexture(P, in.cutoutTex, sp.u, sp.v);
          if (c.w < 0.5f) continue;
        }
        const DevMaterial& mat = P.materials[side];
        if ((mat.kind != MT_DIELECTRIC && mat.kind != MT_WATER) || !P.transparentShadows) {
          occluded = true;
          break;
        }
        vec3 n = sp.frontface ? sp.nFront : -sp.nFront;
        float eta = sp.frontface ? 1.0f / mat.ior : mat.ior;
        vec3 refr;
        if (!refract(d, n, eta, refr)) continue;  // TIR: outside Snell's window
        float f0 = (1.0f - eta) / (1.0f + eta);
        f0 = f0 * f0;
        float c = sp.frontface ? -dot(d, n) : -dot(refr, n);
        float F = schlick(c, f0);
        trF = trF * (1.0f - F);
        float s = sp.frontface ? -t : t;
        tau.x = tau.x + mat.absorb[0] * s;
        tau.y = tau.y + mat.absorb[1] * s;
        tau.z = tau.z + mat.absorb[2] * s;
      }
    }
    if (occluded) break;
    while (true) {
      if (st.sp == base) {
        cur = 0xFFFFFFFFu;
        break;
      }
      cur = st.pop().x;
      if (cur != kRefMarker) break;
      inBlas = false;
    }
    if (cur == 0xFFFFFFFFu) break;
  }
  st.sp = base;
  if (occluded) return v3(0.0f);
  tau = vec3{fmaxf(0.0f, tau.x), fmaxf(0.0f, tau.y), fmaxf(0.0f, tau.z)};
  vec3 tr = v3(trF);
  if (maxc(tau) > 0.0f) tr = tr * vexp(-tau);
  return tr;
}

// The queries for a launch's scene: the specialisations above when no
// instance needs surface tests inside traversal.
DEV bool traceClosest(const LaunchParams& P, Stack& st, vec3 o, vec3 d, float tmin, float tmax, Hit& hit,
                      vec3& oObjOut, vec3& dObjOut, WorkCount& wc) {
  if (P.instFlags & INST_MASKED) return traceClosestT<true>(P, st, o, d, tmin, tmax, hit, oObjOut, dObjOut, wc);
  return traceClosestT<false>(P, st, o, d, tmin, tmax, hit, oObjOut, dObjOut, wc);
}
DEV vec3 traceShadow(const LaunchParams& P, Stack& st, vec3 o, vec3 d, float tmax, WorkCount& wc) {
  if (P.instFlags & (INST_MASKED | INST_TRANSMISSIVE)) return traceShadowT<true>(P, st, o, d, tmax, wc);
  return traceShadowT<false>(P, st, o, d, tmax, wc);
}

// ---------------------------------------------------------------------------
// path megakernel (device/programs.cu:311-572)
// ---------------------------------------------------------------------------

DEV vec3 evalBackground(const LaunchParams& P, vec3 dir) {
  if (P.bgKind == BG_ENVMAP) return envRadiance(P, dir);
  if (P.bgKind == BG_GRADIENT) {
    float t = 0.5f * (dir.y + 1.0f);
    return lerp(ld3(P.bgA), ld3(P.bgB), t);
  }
  return ld3(P.bgA);
}

// Medium stack: four 16-bit material ids packed in two words. The
// reference stores (sigma, ior) per level (device/programs.cu:364-367,
// :531-542); both are per-material constants, so the id carries the same
// information in 2 registers instead of 16.
DEV uint32_t medGet(uint32_t lo, uint32_t hi, int level) {
  uint32_t w = level < 2 ? lo : hi;
  return (level & 1) ? (w >> 16) : (w & 0xFFFFu);
}
DEV void medSet(uint32_t& lo, uint32_t& hi, int level, uint32_t id) {
  uint32_t& w = level < 2 ? lo : hi;
  w = (level & 1) ? ((w & 0xFFFFu) | (id << 16)) : ((w & 0xFFFF0000u) | id);
}

DEV float4 runningMean(float4 m, vec3 x, int s) {
  return make_float4(m.x + (x.x - m.x) / (s + 1), m.y + (x.y - m.y) / (s + 1), m.z + (x.z - m.z) / (s + 1),
                     1.0f);
}

// ---------------------------------------------------------------------------
// path state machine (device/programs.cu:311-572). A path advances through
// stage functions between ray queries: pathBegin (camera ray of the next
// sample), pathOnHit (after a closest-hit query: media, flames, background,
// emission, AOVs, NEE light sample), pathOnShadow (after the NEE shadow
// query: flame transmittance, MIS-weighted light contribution) and
// pathScatter (BSDF sample, medium stack, Russian roulette). The megakernel
// keeps the state in registers and traces inline; the wavefront kernels keep
// it in HBM between kernels. Both run the same stage functions, so both give
// the same film bits.
//
// The NEE shadow query splits a bounce: flame transmittance along the
// shadow ray draws random numbers only when the ray is unblocked
// (device/programs.cu:493-505), so everything after the query waits for it.
// ---------------------------------------------------------------------------

enum PathNext { PN_TRACE = 0, PN_SHADOW = 1, PN_END = 2, PN_SCATTER = 3 };

struct PathState {
  vec3 o, d;  // the path ray of the next closest-hit query
  vec3 L, beta, prevX;
  float prevPdf;
  uint64_t rngState;  // PCG32 state; the increment is (seed << 1) | 1
  uint32_t medLo, medHi;
  int depth, medDepth;
  int s, si;  // sample index; samples begun in this launch
  bool specular;
  vec3 aovA, aovN;  // this sample's first-hit albedo and camera-space normal
};

// What pathOnShadow and pathScatter need from the hit being shaded.
struct SurfState {
  vec3 x, ns, albedo;
  float etaRel;  // outside / inside IOR seen from the incoming ray
  uint32_t matId;
  bool frontface;
  // NEE light sample waiting for its shadow query
  vec3 wi, Li;
  float pdf, dist, cosS;
  int k;
  bool isDelta, transmitNee;
};

DEV int numStrata(const LaunchParams& P) {
  int nStrata = (int)floorf(sqrtf((float)P.sppTotal) + 0.5f);
  if (nStrata * nStrata > P.sppTotal) nStrata--;
  return nStrata < 1 ? 1 : nStrata;
}
DEV int neeStrategies(const LaunchParams& P) { return P.numLights + (P.hasEnv ? 1 : 0); }

// Starts the next sample of the pixel: stratified jitter over the total
// spp, thin-lens or pinhole camera ray (device/programs.cu:316-350). Returns
// false when the launch's samples are all begun.
// Camera ray and fresh path state of sample s of pixel (px, py).
DEV void pathStart(const LaunchParams& P, PathState& ps, int px, int py, int s) {
  const int pixel = py * P.width + px;
  const int nStrata = numStrata(P);
  ps.s = s;
  Pcg rng = Pcg::make(((uint64_t)pixel << 32) ^ (uint64_t)ps.s, (uint64_t)P.seed);
  float jx, jy;
  if (ps.s < nStrata * nStrata) {
    jx = ((ps.s % nStrata) + rng.rnd()) / nStrata;
    jy = ((ps.s / nStrata) + rng.rnd()) / nStrata;
  } else {
    jx = rng.rnd();
    jy = rng.rnd();
  }
  float u = (px + jx) / P.width;
  float v = ((P.height - 1 - py) + jy) / P.height;
  vec3 org = ld3(P.cam.origin);
  if (P.cam.lensRadius > 0.0f) {
    vec2 rd = concentricDisk(rng.rnd2()) * P.cam.lensRadius;
    org = org + ld3(P.cam.u) * rd.x + ld3(P.cam.v) * rd.y;
  }
  ps.d = normalize(ld3(P.cam.lowerLeft) + u * ld3(P.cam.horizontal) + v * ld3(P.cam.vertical) - org);
  ps.o = org;
  ps.prevX = org;
  ps.L = v3(0.0f);
  ps.beta = v3(1.0f);
  ps.prevPdf = 0.0f;
  ps.specular = true;
  ps.medDepth = 0;
  ps.depth = 0;
  ps.aovA = v3(0.0f);
  ps.aovN = v3(0.0f);
  ps.rngState = rng.state;
}

// The next of the launch's samples of pixel (px, py), in order; false once
// all sppThisLaunch have begun.
DEV bool pathBegin(const LaunchParams& P, PathState& ps, int px, int py) {
  if (ps.si == P.sppThisLaunch) return false;
  pathStart(P, ps, px, py, P.sampleOffset + ps.si);
  ps.si++;
  return true;
}

DEV Pcg pathRng(const LaunchParams& P, const PathState& ps) {
  Pcg r;
  r.state = ps.rngState;
  r.inc = ((uint64_t)P.seed << 1u) | 1u;
  return r;
}

// BSDF sample, medium stack update, throughput, Russian roulette from depth
// 4 with q = clamp(max beta, 0.05, 0.95) (device/programs.cu:524-557).
DEV int pathScatter(const LaunchParams& P, PathState& ps, const SurfState& sf, Pcg& rng) {
  const Closure cl = makeClosure(P.materials[sf.matId], sf.albedo, sf.etaRel);
  BsdfSample bs = sampleClosure(cl, ps.d, sf.ns, rng);
  if (!bs.valid) return PN_END;
  // a refracted ray enters (front face) or leaves (back face) the medium
  if ((cl.lobe == LK_INTERFACE || cl.lobe == LK_GLASS) && dot(bs.wi, sf.ns) < 0.0f) {
    if (sf.frontface) {
      if (ps.medDepth < kMedMax) medSet(ps.medLo, ps.medHi, ps.medDepth, sf.matId);
      ps.medDepth++;
    } else if (ps.medDepth > 0) {
      ps.medDepth--;
    }
  }
  ps.beta *= bs.weight;
  ps.specular = bs.isDelta;
  ps.prevPdf = bs.isDelta ? 0.0f : bs.pdf;
  ps.prevX = sf.x;
  vec3 offN = dot(bs.wi, sf.ns) >= 0.0f ? sf.ns : -sf.ns;
  ps.o = offsetRay(sf.x, offN);
  ps.d = bs.wi;
  if (ps.depth >= 4) {
    float q = clampf(maxc(ps.beta), 0.05f, 0.95f);
    if (rng.rnd() >= q) return PN_END;
    ps.beta *= 1.0f / q;
  }
  if (maxc(ps.beta) <= 0.0f) return PN_END;
  ps.depth++;
  return ps.depth < P.maxDepth ? PN_TRACE : PN_END;
}

// The sample's first-hit AOVs: into the path state, or, when aovOut is set,
// straight into the sample's slot of a pooled sample buffer (planes 3..8 at
// a stride of kPoolSamples x 64 floats), so they are not carried in
// registers through the rest of the path.
DEV void storeAov(PathState& ps, float* aovOut, vec3 a, vec3 n) {
  if (aovOut) {
    constexpr int kS = kPoolSamples * 64;
    aovOut[0] = a.x;
    aovOut[kS] = a.y;
    aovOut[2 * kS] = a.z;
    aovOut[3 * kS] = n.x;
    aovOut[4 * kS] = n.y;
    aovOut[5 * kS] = n.z;
  } else {
    ps.aovA = a;
    ps.aovN = n;
  }
}

// After the closest-hit query of ps.o/ps.d (isHit, hit; oObj/dObj is the
// ray in the hit instance's object space). Returns PN_SHADOW when an NEE
// shadow query (shadowRay) must run before the bounce can finish; with
// deferScatter, PN_SCATTER when the bounce only needs its BSDF sample
// (pathScatter from ps.rngState), otherwise the bounce is finished here.
DEV int pathOnHit(const LaunchParams& P, PathState& ps, bool isHit, const Hit& hit, vec3 oObj, vec3 dObj,
                  SurfState& sf, float* aovOut = nullptr, bool deferScatter = false) {
  const int nStrat = neeStrategies(P);
  Pcg rng = pathRng(P, ps);
  if (ps.medDepth > 0) {
    const DevMaterial& mm = P.materials[medGet(ps.medLo, ps.medHi, min(ps.medDepth, kMedMax) - 1)];
    vec3 sigma = ld3(mm.absorb);
    if (maxc(sigma) > 0.0f) ps.beta *= vexp(-sigma * (isHit ? hit.t : 1e4f));
  }
  if (P.numFlames > 0) {
    vec3 Lv = v3(0.0f);
    const float T = marchVolumes<true>(P, ps.o, ps.d, isHit ? hit.t : 1e16f, rng, -1, Lv);
    vec3 cv = ps.beta * Lv;
    if (ps.depth >= 1 && P.clampVal > 0.0f) cv = vmin(cv, v3(P.clampVal));
    ps.L += cv;
    ps.beta *= v3(T);
  }
  if (!isHit) {
    vec3 bg = evalBackground(P, ps.d);
    float w = 1.0f;
    if (P.hasEnv && !ps.specular) {
      float pdfE = envPdf(P, ps.d) / nStrat;
      w = ps.prevPdf / (ps.prevPdf + pdfE);
    }
    vec3 c = ps.beta * bg * w;
    if (ps.depth >= 1 && P.clampVal > 0.0f) c = vmin(c, v3(P.clampVal));
    ps.L += c;
    if (ps.depth == 0) storeAov(ps, aovOut, vec3{fminf(bg.x, 1.0f), fminf(bg.y, 1.0f), fminf(bg.z, 1.0f)}, v3(0.0f));
    ps.rngState = rng.state;
    return PN_END;
  }

  const DevInstance& in = P.instances[hit.inst];
  ShadePoint sp = shadePoint(P, in, hit.prim, hit.b1, hit.b2, oObj, dObj, hit.t, true);
  sf.frontface = sp.frontface;
  sf.matId = sp.frontface ? (in.mats & 0xFFFFu) : (in.mats >> 16);
  vec3 ns = sp.frontface ? sp.nFront : -sp.nFront;
  sf.x = ps.o + hit.t * ps.d;
  const DevMaterial& mat = P.materials[sf.matId];
  sf.albedo = surfaceColor(P, mat, sp.u, sp.v);
  if (mat.kind == MT_WATER && mat.waveAmp > 0.0f)
    ns = rippleNormal(sf.x, ns, -ps.d, mat.waveAmp, mat.waveFreq, mat.waveOctaves);
  sf.ns = ns;

  if (ps.depth == 0)
    storeAov(ps, aovOut, sf.albedo, vec3{dot(ns, ld3(P.cam.u)), dot(ns, ld3(P.cam.v)), dot(ns, ld3(P.cam.w))});

  if (mat.kind == MT_EMISSIVE) {
    if (sp.frontface || mat.twoSided) {
      vec3 Le = sf.albedo * mat.intensity;
      float w = 1.0f;
      if (!ps.specular && in.lightId >= 0 && P.numLights > 0) {
        float pdfL = lightPdfSolidAngle(P.lights[in.lightId], ps.prevX, sf.x, ns) / nStrat;
        w = ps.prevPdf / (ps.prevPdf + pdfL);
      }
      vec3 c = ps.beta * Le * w;
      if (ps.depth >= 1 && P.clampVal > 0.0f) c = vmin(c, v3(P.clampVal));
      ps.L += c;
    }
    ps.rngState = rng.state;
    return PN_END;
  }

  // the IOR outside an interface is that of the enclosing medium
  float etaExt = 1.0f;
  if (mat.kind == MT_DIELECTRIC || mat.kind == MT_WATER) {
    int outer = min(sp.frontface ? ps.medDepth : ps.medDepth - 1, kMedMax);
    if (outer > 0) etaExt = P.materials[medGet(ps.medLo, ps.medHi, outer - 1)].ior;
  }
  sf.etaRel = sp.frontface ? etaExt / mat.ior : mat.ior / etaExt;

  if (nStrat > 0 && !materialIsDelta(mat)) {
    int k = min((int)(rng.rnd() * nStrat), nStrat - 1);
    LightSample ls = (k < P.numLights) ? sampleLight(P, P.lights[k], sf.x, rng) : sampleEnv(P, rng);
    float cosS = ls.valid ? dot(ls.wi, ns) : 0.0f;
    bool transmitNee = cosS < 0.0f && mat.kind == MT_DIELECTRIC;
    if (ls.valid && (cosS > 0.0f || transmitNee)) {
      sf.wi = ls.wi;
      sf.Li = ls.Li;
      sf.pdf = ls.pdf;
      sf.dist = ls.dist;
      sf.cosS = cosS;
      sf.k = k;
      sf.isDelta = ls.isDelta;
      sf.transmitNee = transmitNee;
      ps.rngState = rng.state;
      return PN_SHADOW;
    }
  }
  if (deferScatter) {
    ps.rngState = rng.state;
    return PN_SCATTER;
  }
  int r = pathScatter(P, ps, sf, rng);
  ps.rngState = rng.state;
  return r;
}

// The NEE shadow ray of a suspended bounce.
DEV void shadowRay(const SurfState& sf, vec3& so, vec3& wi, float& tmax) {
  so = offsetRay(sf.x, sf.transmitNee ? -sf.ns : sf.ns);
  wi = sf.wi;
  tmax = sf.dist * 0.999f;
}

// MIS-weighted light contribution of the suspended NEE sample before the
// shadow query's transmittance is applied (device/programs.cu:481-524).
// Draws no random numbers.
// One fused closure evaluation gives both the BSDF value and its density.
DEV vec3 neeUnshadowed(const LaunchParams& P, const PathState& ps, const SurfState& sf) {
  const int nStrat = neeStrategies(P);
  const Closure cl = makeClosure(P.materials[sf.matId], sf.albedo, sf.etaRel);
  const LobeEval e = evalClosureAt(cl, sf.ns, -ps.d, sf.wi);
  const float pdfLe = (sf.isDelta ? 1.0f : sf.pdf) / nStrat;
  const float w = sf.isDelta ? 1.0f : pdfLe / (pdfLe + e.pdf);
  return ps.beta * e.f * fabsf(sf.cosS) * sf.Li * w / pdfLe;
}

// Whether the random numbers a bounce draws after its shadow query depend on
// that query's outcome: only the flame transmittance of a transparent
// shadow ray, which is marched (and draws) when the ray is unoccluded.
DEV bool shadowDraws(const LaunchParams& P) { return P.numFlames > 0 && P.transparentShadows; }

// After the shadow query (transmittance tr): MIS-weighted light
// contribution (device/programs.cu:481-524), then the BSDF sample.
DEV int pathOnShadow(const LaunchParams& P, PathState& ps, const SurfState& sf, vec3 tr) {
  Pcg rng = pathRng(P, ps);
  if (maxc(tr) > 0.0f) {
    if (shadowDraws(P)) {
      vec3 so, wi;
      float tmaxS;
      shadowRay(sf, so, wi, tmaxS);
      int skip = (sf.k < P.numLights) ? P.lights[sf.k].flameId : -1;
      tr = tr * volumeTransmittance(P, so, wi, tmaxS, rng, skip);
    }
    vec3 c = neeUnshadowed(P, ps, sf);
    c = c * tr;
    if (ps.depth >= 1 && P.clampVal > 0.0f) c = vmin(c, v3(P.clampVal));
    ps.L += sanitize(c);
  }
  int r = pathScatter(P, ps, sf, rng);
  ps.rngState = rng.state;
  return r;
}

// Wave-level sums of the ray / work counters, one 64-bit atomic per wave.
// Every lane of the wave must reach this call.
template <bool kCountWork>
DEV void flushCounters(bool countRays, unsigned long long* rayCounter, unsigned long long* workCounters,
                       uint32_t rays, const WorkCount& wc) {
  if (kCountWork) {
    uint32_t work[4] = {wc.nodes, wc.tris, wc.quads, wc.insts};
#pragma unroll
    for (int c = 0; c < 4; c++) {
      unsigned long long v = work[c];
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
      if ((threadIdx.x & 63) == 0 && v) atomicAdd(workCounters + c, v);
    }
  }
  if (countRays) {
    uint32_t r = rays;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) r += __shfl_xor(r, off, 64);
    if ((threadIdx.x & 63) == 0 && r) atomicAdd(rayCounter, (unsigned long long)r);
  }
}

// Shared-sample driver for one wave of the render grid: the wave's 64 pixels
// x kPoolSamples samples form one pool of paths, and a lane whose path ends
// takes the next path of the pool from an LDS counter, whatever its pixel.
// The wave then idles only for its last paths, not for the pixel whose 16
// paths happened to be longest. Each path's sanitized radiance and first-hit
// AOVs go to the wave's region of P.sampleBuf; once the pool is drained each
// lane folds its own pixel's samples in sample order, the same running-mean
// operations in the same order as a pixel rendered alone, so films are
// bit-identical to per-pixel rendering. Launches of more than kPoolSamples
// samples drain one pool per kPoolSamples. slotPixel(j, px, py) names the
// pixel of slot j (0..63) and returns false for a slot outside the image.
// With fold == false the wave renders only the launch's samples
// [gBegin, gEnd) (at most kPoolSamples) into slots gBegin * 64 + item of its
// region and leaves the fold to k_fold_samples, which runs once every
// sample group of the launch has written its slots.
template <class SlotPixel>
DEV void renderWavePooled(const LaunchParams& P, SlotPixel slotPixel, uint32_t* poolNext, float* waveBuf, Stack& st,
                          WorkCount& wc, uint32_t& rays, int gBegin, int gEnd, bool fold) {
  const int lane = threadIdx.x & 63;
  int myPx = 0, myPy = 0;
  const bool mine = slotPixel(lane, myPx, myPy);
  for (int c0 = gBegin; c0 < gEnd; c0 += kPoolSamples) {
    const int n = min(kPoolSamples, gEnd - c0);
    const int items = n * 64;  // item = sample-in-pool * 64 + slot
    if (lane == 0) *poolNext = 64;  // item `lane` is each lane's first
    // this pool's slots: from 0 when the wave folds each pool itself, else
    // at the group's first sample so that groups do not overlap
    float* poolBuf = waveBuf + (fold ? 0 : c0 * 64);
    int item = lane;
    PathState ps;
    ps.si = 0;
    ps.medLo = ps.medHi = 0;
    // claim items until one lies inside the image (or the pool is empty)
    auto claim = [&]() -> bool {
      while (item < items) {
        int px = 0, py = 0;
        if (slotPixel(item & 63, px, py)) {
          pathStart(P, ps, px, py, P.sampleOffset + c0 + (item >> 6));
          return true;
        }
        item = (int)atomicAdd(poolNext, 1u);
      }
      return false;
    };
    bool alive = claim();
    while (alive) {
      Hit hit;
      vec3 oObj = v3(0.0f), dObj = v3(0.0f);
      bool isHit = traceClosest(P, st, ps.o, ps.d, 0.0f, 1e16f, hit, oObj, dObj, wc);
      rays++;
      SurfState sf;
      // first-hit AOVs go straight to the sample's slot
      int r = pathOnHit(P, ps, isHit, hit, oObj, dObj, sf, poolBuf + 3 * (kPoolSamples * 64) + item, true);
      // The NEE sample's unshadowed contribution is computed before its
      // shadow query. Unless the flame march of a transparent shadow ray
      // draws random numbers after the query (which the BSDF sample's
      // draws must follow), the BSDF sample is taken before the shadow
      // query too, so the surface state is dead while the shadow ray
      // traverses. Same operations on the same values, in the same draw
      // order, as pathOnShadow.
      vec3 so = v3(0.0f), wi = v3(0.0f), cNee = v3(0.0f);
      float tmaxS = 0.0f;
      bool clampNee = false, pending = false;
      if (r == PN_SHADOW) {
        shadowRay(sf, so, wi, tmaxS);
        rays++;
        cNee = neeUnshadowed(P, ps, sf);
        clampNee = ps.depth >= 1 && P.clampVal > 0.0f;
        if (shadowDraws(P)) {
          vec3 tr = traceShadow(P, st, so, wi, tmaxS, wc);
          if (maxc(tr) > 0.0f) {
            Pcg rng = pathRng(P, ps);
            const int skip = (sf.k < P.numLights) ? P.lights[sf.k].flameId : -1;
            tr = tr * volumeTransmittance(P, so, wi, tmaxS, rng, skip);
            ps.rngState = rng.state;
            vec3 c = cNee * tr;
            if (clampNee) c = vmin(c, v3(P.clampVal));
            ps.L += sanitize(c);
          }
        } else {
          pending = true;
        }
        r = PN_SCATTER;
      }
      if (r == PN_SCATTER) {
        Pcg rng = pathRng(P, ps);
        r = pathScatter(P, ps, sf, rng);
        ps.rngState = rng.state;
      }
      if (pending) {
        const vec3 tr = traceShadow(P, st, so, wi, tmaxS, wc);
        if (maxc(tr) > 0.0f) {
          vec3 c = cNee * tr;
          if (clampNee) c = vmin(c, v3(P.clampVal));
          ps.L += sanitize(c);
        }
      }
      if (r == PN_END) {
        const vec3 L = sanitize(ps.L);
        const float v[3] = {L.x, L.y, L.z};
#pragma unroll
        for (int k = 0; k < 3; k++) poolBuf[k * (kPoolSamples * 64) + item] = v[k];
        item = (int)atomicAdd(poolNext, 1u);
        alive = claim();
      }
    }
    // every lane of the wave is past its last path: make the wave's
    // per-sample stores visible to the lane that folds each pixel
    __threadfence_block();
    if (!fold) continue;
    // (the film is only touched here, so no film value is live while paths
    // are traced)
    if (mine) {
      const int pixel = myPy * P.width + myPx;
      float4 fAcc = P.accum[pixel], fAlb = P.albedo[pixel], fNrm = P.normal[pixel];
      for (int si = 0; si < n; si++) {
        float x[kPoolPlanes];
#pragma unroll
        for (int k = 0; k < kPoolPlanes; k++) x[k] = waveBuf[k * (kPoolSamples * 64) + si * 64 + lane];
        const int s = P.sampleOffset + c0 + si;
        fAcc = runningMean(fAcc, vec3{x[0], x[1], x[2]}, s);
        fAlb = runningMean(fAlb, vec3{x[3], x[4], x[5]}, s);
        fNrm = runningMean(fNrm, vec3{x[6], x[7], x[8]}, s);
      }
      P.accum[pixel] = fAcc;
      P.albedo[pixel] = fAlb;
      P.normal[pixel] = fNrm;
    }
    // the next pool reuses the counter and the buffer region
    __threadfence_block();
  }
}

// kWaves: minimum waves per SIMD the register allocation must allow (2 fits
// the kernel without spills; 3 and 4 trade scratch spills for occupancy).
template <bool kCountWork, int kWaves>
__global__ __launch_bounds__(kBlock, kWaves) void k_render(LaunchParams P) {
  extern __shared__ uint64_t s_stack[];  // P.stackLds * kBlock entries
  __shared__ uint32_t s_poolNext[kBlock / 64];
  uint32_t rays = 0;
  WorkCount wc;
  Stack st;
  st.lds = ldsColumn(s_stack);
  st.spill = reinterpret_cast<uint64_t*>(P.spill) +
             ((size_t)(blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x) * kBlock + threadIdx.x;
  st.stride = P.spillStride;
  st.sp = 0;
  st.cap = P.stackLds;
  const int wave = threadIdx.x >> 6;
  const size_t waveGlobal = (size_t)(blockIdx.y * gridDim.x + blockIdx.x) * (kBlock / 64) + wave;
  // sample group blockIdx.z of gridDim.z: a contiguous share of the launch's
  // samples (one group: all of them, folded here)
  const int groups = (int)gridDim.z;
  const int per = (P.sppThisLaunch + groups - 1) / groups;
  const int gBegin = min(P.sppThisLaunch, (int)blockIdx.z * per);
  const int gEnd = min(P.sppThisLaunch, gBegin + per);
  // 16x16-pixel blocks; each wave covers an 8x8 quarter, slot j at (j % 8, j / 8)
  const int x0 = blockIdx.x * 16 + (wave & 1) * 8, y0 = blockIdx.y * 16 + (wave >> 1) * 8;
  auto slotPixel = [&](int j, int& px, int& py) -> bool {
    px = x0 + (j & 7);
    py = y0 + (j >> 3);
    return px < P.width && py < P.height;
  };
  renderWavePooled(P, slotPixel, &s_poolNext[wave], P.sampleBuf + waveGlobal * (kPoolPlanes * kPoolSamples * 64), st,
                   wc, rays, gBegin, gEnd, groups == 1);
  flushCounters<kCountWork>(P.countRays != 0, P.rayCounter, P.workCounters, rays, wc);
}

// The fold of a launch rendered in several sample groups: each pixel's
// samples, in sample order, into its running means (the same operations as
// the in-kernel fold of renderWavePooled).
__global__ void k_fold_samples(LaunchParams P) {
  const int px = blockIdx.x * 16 + (threadIdx.x & 15), py = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (px >= P.width || py >= P.height) return;
  const int tx = px & 15, ty = py & 15;
  const int wave = (ty >> 3) * 2 + (tx >> 3), lane = (ty & 7) * 8 + (tx & 7);
  const size_t waveGlobal = (size_t)(blockIdx.y * gridDim.x + blockIdx.x) * (kBlock / 64) + wave;
  const float* buf = P.sampleBuf + waveGlobal * (kPoolPlanes * kPoolSamples * 64);
  const int pixel = py * P.width + px;
  float4 fAcc = P.accum[pixel], fAlb = P.albedo[pixel], fNrm = P.normal[pixel];
  for (int si = 0; si < P.sppThisLaunch; si++) {
    float x[kPoolPlanes];
#pragma unroll
    for (int k = 0; k < kPoolPlanes; k++) x[k] = buf[k * (kPoolSamples * 64) + si * 64 + lane];
    const int s = P.sampleOffset + si;
    fAcc = runningMean(fAcc, vec3{x[0], x[1], x[2]}, s);
    fAlb = runningMean(fAlb, vec3{x[3], x[4], x[5]}, s);
    fNrm = runningMean(fNrm, vec3{x[6], x[7], x[8]}, s);
  }
  P.accum[pixel] = fAcc;
  P.albedo[pixel] = fAlb;
  P.normal[pixel] = fNrm;
}

// ---------------------------------------------------------------------------
// persistent tile-queue kernel: one launch drains (request, tile) work items
// of many requests. Each wave takes the next item from a global counter and
// renders that tile's pixels for the request's current pass; a pixel is only
// ever touched by the wave holding its tile, so films stay bit-identical to a
// render of the request alone.
// ---------------------------------------------------------------------------

// Prepares one graph-replayable pass: resets the work counter and writes
// each request's sample offset for this pass (offset = first + pass * count,
// clipped to the request's total). Runs as a kernel, never a memset, so it
// can sit inside a captured graph.
// Zeroes the pass and item counters of n queues (a kernel so that it can
// open a captured graph).
__global__ void k_queue_reset(QueueState* q, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    q[i].next = 0;
    q[i].pass = 0;
  }
}

__global__ void k_queue_prep(QueueState* q, RequestDev* reqs, const int32_t* firstSample, int numReq) {
  __shared__ uint32_t s_pass;
  if (threadIdx.x == 0) s_pass = q->pass;
  __syncthreads();
  for (int r = threadIdx.x; r < numReq; r += blockDim.x) {
    RequestDev& R = reqs[r];
    int chunk = R.passSpp > 0 ? R.passSpp : 1;
    int off = firstSample[r] + (int)s_pass * chunk;
    R.sampleOffset = off;
    R.sppThisLaunch = max(0, min(chunk, R.sppTotal - off));
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    q->next = 0;
    q->pass = s_pass + 1;
  }
}

// reqs[r].sppThisLaunch is already clipped to the request's remaining
// samples (by the host or k_queue_prep). The request index is wave-uniform,
// so the scene is picked from the kernel-argument table with scalar loads
// and only the request's own fields come from device memory.
template <int kWaves>
__global__ __launch_bounds__(kBlock, kWaves) void k_persistent(SceneTable S, const RequestDev* __restrict__ reqs,
                                                               const WorkItem* __restrict__ items, int numItems,
                                                               QueueState* q, int tileSize, uint32_t* spill,
                                                               int spillStride, int stackLds, float* sampleBuf,
                                                               float* groupBuf, unsigned long long* rayCounter) {
  extern __shared__ uint64_t s_stack[];  // stackLds * kBlock entries
  __shared__ uint32_t s_poolNext[kBlock / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  float* waveBuf = sampleBuf + ((size_t)blockIdx.x * (kBlock / 64) + wave) * (kPoolPlanes * kPoolSamples * 64);
  Stack st;
  st.lds = ldsColumn(s_stack);
  st.spill = reinterpret_cast<uint64_t*>(spill) + (size_t)blockIdx.x * kBlock + threadIdx.x;
  st.stride = spillStride;
  st.sp = 0;
  st.cap = stackLds;
  WorkCount wc;
  uint32_t rays = 0;
  const int pxPerTile = tileSize * tileSize;
  while (true) {
    uint32_t item = 0;
    if (lane == 0) item = atomicAdd(&q->next, 1u);
    item = __builtin_amdgcn_readfirstlane(__shfl(item, 0, 64));
    if (item >= (uint32_t)numItems) break;
    const int req = __builtin_amdgcn_readfirstlane(items[item].request);
    const int tile = __builtin_amdgcn_readfirstlane(items[item].tile);
    const int group = __builtin_amdgcn_readfirstlane(items[item].group);
    const RequestDev& R = reqs[req];
    if (R.sppThisLaunch <= 0) continue;  // request already complete
    const int groups = __builtin_amdgcn_readfirstlane(R.groups);
    LaunchParams P = S.scene[__builtin_amdgcn_readfirstlane(R.scene)];
    P.width = R.width;
    P.height = R.height;
    P.sppTotal = R.sppTotal;
    P.sampleOffset = R.sampleOffset;
    P.sppThisLaunch = R.sppThisLaunch;
    P.seed = R.seed;
    P.accum = R.accum;
    P.albedo = R.albedo;
    P.normal = R.normal;
    P.cam = R.cam;
    const int tilesX = (P.width + tileSize - 1) / tileSize;
    const int tx = (tile % tilesX) * tileSize, ty = (tile / tilesX) * tileSize;
    // this item's share of the pass's samples (all of them unless split)
    int gBegin = 0, gEnd = P.sppThisLaunch;
    if (groups > 1) {
      const int per = (P.sppThisLaunch + groups - 1) / groups;
      gBegin = min(P.sppThisLaunch, group * per);
      gEnd = min(P.sppThisLaunch, gBegin + per);
      if (gBegin >= gEnd) continue;
    }
    // each 64-pixel group of the tile drains one shared sample pool
    for (int base = 0; base < pxPerTile; base += 64) {
      auto slotPixel = [&](int j, int& px, int& py) -> bool {
        const int k = base + j;
        px = tx + (k % tileSize);
        py = ty + (k / tileSize);
        return k < pxPerTile && px < P.width && py < P.height;
      };
      if (groups > 1) {
        // samples [gBegin, gEnd) into the pool's region; k_fold_groups folds
        const size_t region = (size_t)R.regionBase + (size_t)tile * (pxPerTile / 64) + base / 64;
        renderWavePooled(P, slotPixel, &s_poolNext[wave], groupBuf + region * (kPoolPlanes * kPoolSamples * 64),
                         st, wc, rays, gBegin, gEnd, false);
      } else {
        renderWavePooled(P, slotPixel, &s_poolNext[wave], waveBuf, st, wc, rays, 0, P.sppThisLaunch, true);
      }
    }
  }
  flushCounters<false>(rayCounter != nullptr, rayCounter, nullptr, rays, wc);
}

// The fold of the passes that k_persistent split into sample groups: block
// (item, pool) folds the 64 pixels of one pool of a group-0 item's tile, each
// pixel's samples in sample order into its running means (the same
// operations as the in-kernel fold, so films stay bit-identical).
__global__ __launch_bounds__(64) void k_fold_groups(const RequestDev* __restrict__ reqs,
                                                    const WorkItem* __restrict__ items, int tileSize,
                                                    const float* __restrict__ groupBuf) {
  const int pools = tileSize * tileSize / 64;
  const int item = blockIdx.x / pools, pool = blockIdx.x % pools;
  const WorkItem& wi = items[item];
  if (wi.group != 0) return;
  const RequestDev& R = reqs[wi.request];
  if (R.groups <= 1 || R.sppThisLaunch <= 0) return;
  const int tilesX = (R.width + tileSize - 1) / tileSize;
  const int tx = (wi.tile % tilesX) * tileSize, ty = (wi.tile / tilesX) * tileSize;
  const int j = threadIdx.x, k = pool * 64 + j;
  const int px = tx + k % tileSize, py = ty + k / tileSize;
  if (px >= R.width || py >= R.height) return;
  const size_t region = (size_t)R.regionBase + (size_t)wi.tile * pools + pool;
  const float* buf = groupBuf + region * (kPoolPlanes * kPoolSamples * 64);
  const int pixel = py * R.width + px;
  float4 fAcc = R.accum[pixel], fAlb = R.albedo[pixel], fNrm = R.normal[pixel];
  for (int si = 0; si < R.sppThisLaunch; si++) {
    float x[kPoolPlanes];
#pragma unroll
    for (int c = 0; c < kPoolPlanes; c++) x[c] = buf[c * (kPoolSamples * 64) + si * 64 + j];
    const int s = R.sampleOffset + si;
    fAcc = runningMean(fAcc, vec3{x[0], x[1], x[2]}, s);
    fAlb = runningMean(fAlb, vec3{x[3], x[4], x[5]}, s);
    fNrm = runningMean(fNrm, vec3{x[6], x[7], x[8]}, s);
  }
  R.accum[pixel] = fAcc;
  R.albedo[pixel] = fAlb;
  R.normal[pixel] = fNrm;
}

// ---------------------------------------------------------------------------
// wavefront variant: the same path stages with the state in HBM planes, one
// kernel per stage (shade, closest-hit query, shadow query), compacted
// queues between them, and shading work ordered by material kind.
// ---------------------------------------------------------------------------

enum WfStage { WF_BEGIN = 0, WF_HIT = 1, WF_SHADOW = 2, WF_DONE = 3 };
enum WfPlane {
  WP_O = 0,   // o.xyz, prevPdf
  WP_D,       // d.xyz, s
  WP_L,       // L.xyz, si
  WP_BETA,    // beta.xyz, stage | specular << 2 | depth << 8 | medDepth << 16
  WP_PREVX,   // prevX.xyz, medLo
  WP_AOVA,    // aovA.xyz, medHi
  WP_AOVN,    // aovN.xyz, -
  WP_RNG,     // rng state lo, hi, -, -
  WP_RES0,    // closest: t, inst, prim, b1 | shadow: transmittance.xyz, -
  WP_RES1,    // closest: b2, -, -, -
  WP_SX,      // surface x.xyz, etaRel
  WP_SNS,     // ns.xyz, -
  WP_SALB,    // albedo.xyz, matId | frontface << 16 | isDelta << 17 | transmitNee << 18
  WP_SWI,     // wi.xyz, pdf
  WP_SLI,     // Li.xyz, dist
  WP_SMISC,   // cosS, k, -, -
};
constexpr int kKeyMiss = 0, kKeyHit = 1, kKeyShadow = 8;  // + material kind

DEV float4 wfLd(const WfParams& W, int plane, int slot) { return W.state[(size_t)plane * W.numSlots + slot]; }
DEV void wfSt(const WfParams& W, int plane, int slot, float4 v) { W.state[(size_t)plane * W.numSlots + slot] = v; }
DEV float4 f4(vec3 v, float w) { return make_float4(v.x, v.y, v.z, w); }
DEV float4 f4i(vec3 v, uint32_t w) { return make_float4(v.x, v.y, v.z, __uint_as_float(w)); }
DEV uint32_t wbits(float4 v) { return __float_as_uint(v.w); }

DEV int wfLoadPath(const WfParams& W, int slot, PathState& ps) {
  float4 a = wfLd(W, WP_O, slot), b = wfLd(W, WP_D, slot), c = wfLd(W, WP_L, slot), e = wfLd(W, WP_BETA, slot);
  float4 f = wfLd(W, WP_PREVX, slot), g = wfLd(W, WP_AOVA, slot), h = wfLd(W, WP_AOVN, slot);
  float4 r = wfLd(W, WP_RNG, slot);
  ps.o = f4xyz(a);
  ps.prevPdf = a.w;
  ps.d = f4xyz(b);
  ps.s = (int)wbits(b);
  ps.L = f4xyz(c);
  ps.si = (int)wbits(c);
  ps.beta = f4xyz(e);
  uint32_t fl = wbits(e);
  ps.specular = (fl >> 2) & 1u;
  ps.depth = (int)((fl >> 8) & 0xFFu);
  ps.medDepth = (int)((fl >> 16) & 0xFFu);
  ps.prevX = f4xyz(f);
  ps.medLo = wbits(f);
  ps.aovA = f4xyz(g);
  ps.medHi = wbits(g);
  ps.aovN = f4xyz(h);
  ps.rngState = (uint64_t)__float_as_uint(r.x) | ((uint64_t)__float_as_uint(r.y) << 32);
  return (int)(fl & 3u);
}

DEV void wfStorePath(const WfParams& W, int slot, const PathState& ps, int stage) {
  uint32_t fl = (uint32_t)stage | ((ps.specular ? 1u : 0u) << 2) | ((uint32_t)ps.depth << 8) |
                ((uint32_t)ps.medDepth << 16);
  wfSt(W, WP_O, slot, f4(ps.o, ps.prevPdf));
  wfSt(W, WP_D, slot, f4i(ps.d, (uint32_t)ps.s));
  wfSt(W, WP_L, slot, f4i(ps.L, (uint32_t)ps.si));
  wfSt(W, WP_BETA, slot, f4i(ps.beta, fl));
  wfSt(W, WP_PREVX, slot, f4i(ps.prevX, ps.medLo));
  wfSt(W, WP_AOVA, slot, f4i(ps.aovA, ps.medHi));
  wfSt(W, WP_AOVN, slot, f4(ps.aovN, 0.0f));
  wfSt(W, WP_RNG, slot,
       make_float4(__uint_as_float((uint32_t)ps.rngState), __uint_as_float((uint32_t)(ps.rngState >> 32)), 0.0f,
                   0.0f));
}

DEV void wfLoadSurf(const WfParams& W, int slot, SurfState& sf) {
  float4 a = wfLd(W, WP_SX, slot), b = wfLd(W, WP_SNS, slot), c = wfLd(W, WP_SALB, slot);
  float4 d = wfLd(W, WP_SWI, slot), e = wfLd(W, WP_SLI, slot), f = wfLd(W, WP_SMISC, slot);
  sf.x = f4xyz(a);
  sf.etaRel = a.w;
  sf.ns = f4xyz(b);
  sf.albedo = f4xyz(c);
  uint32_t m = wbits(c);
  sf.matId = m & 0xFFFFu;
  sf.frontface = (m >> 16) & 1u;
  sf.isDelta = (m >> 17) & 1u;
  sf.transmitNee = (m >> 18) & 1u;
  sf.wi = f4xyz(d);
  sf.pdf = d.w;
  sf.Li = f4xyz(e);
  sf.dist = e.w;
  sf.cosS = f.x;
  sf.k = (int)__float_as_uint(f.y);
}

DEV void wfStoreSurf(const WfParams& W, int slot, const SurfState& sf) {
  uint32_t m = sf.matId | ((sf.frontface ? 1u : 0u) << 16) | ((sf.isDelta ? 1u : 0u) << 17) |
               ((sf.transmitNee ? 1u : 0u) << 18);
  wfSt(W, WP_SX, slot, f4(sf.x, sf.etaRel));
  wfSt(W, WP_SNS, slot, f4(sf.ns, 0.0f));
  wfSt(W, WP_SALB, slot, f4i(sf.albedo, m));
  wfSt(W, WP_SWI, slot, f4(sf.wi, sf.pdf));
  wfSt(W, WP_SLI, slot, f4(sf.Li, sf.dist));
  wfSt(W, WP_SMISC, slot, make_float4(sf.cosS, __uint_as_float((uint32_t)sf.k), 0.0f, 0.0f));
}

// Appends `value` for every lane with `want` to list[*n ...] with one atomic
// per wave. Every lane of the wave must call it.
DEV void wavePush(uint32_t* list, uint32_t* n, bool want, uint32_t value) {
  const uint64_t m = __ballot(want);
  if (m == 0) return;
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((long long)m) - 1;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(n, (uint32_t)__popcll(m));
  base = __shfl(base, leader, 64);
  if (want) list[base + __popcll(m & ((1ull << lane) - 1ull))] = value;
}

// Appends `value` to bucket `key` (kWfKeys lists of `cap` entries), one
// atomic per distinct key in the wave. Every lane of the wave must call it.
DEV void wavePushKeyed(uint32_t* buckets, int cap, uint32_t* bucketN, bool want, int key, uint32_t value) {
  uint64_t pending = __ballot(want);
  const int lane = threadIdx.x & 63;
  while (pending) {
    const int leader = __ffsll((long long)pending) - 1;
    const int lk = __shfl(key, leader, 64);
    const uint64_t m = pending & __ballot(key == lk);
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(&bucketN[lk], (uint32_t)__popcll(m));
    base = __shfl(base, leader, 64);
    if ((m >> lane) & 1ull) buckets[(size_t)lk * cap + base + __popcll(m & ((1ull << lane) - 1ull))] = value;
    pending &= ~m;
  }
}

DEV void wfFold(const LaunchParams& P, int pixel, const PathState& ps) {
  P.accum[pixel] = runningMean(P.accum[pixel], sanitize(ps.L), ps.s);
  P.albedo[pixel] = runningMean(P.albedo[pixel], ps.aovA, ps.s);
  P.normal[pixel] = runningMean(P.normal[pixel], ps.aovN, ps.s);
}

// Shading pass: every slot of the current buckets (or every slot on the
// first pass of a launch) resumes its path from the query it waited for and
// runs until its next query, which it enqueues.
__global__ __launch_bounds__(kBlock) void k_wf_shade(LaunchParams P, WfParams W) {
  __shared__ uint32_t s_pre[kWfKeys + 1];
  if (threadIdx.x == 0) {
    uint32_t acc = 0;
    for (int k = 0; k < kWfKeys; k++) {
      s_pre[k] = acc;
      acc += W.first ? 0u : W.cnt->bucketN[k];
    }
    s_pre[kWfKeys] = W.first ? (uint32_t)W.numSlots : acc;
  }
  __syncthreads();
  const uint32_t total = s_pre[kWfKeys];
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t base = blockIdx.x * blockDim.x; base < total; base += stride) {
    const uint32_t i = base + threadIdx.x;
    const bool valid = i < total;
    int slot = 0, next = WF_DONE;
    if (valid) {
      if (W.first) {
        slot = (int)i;
      } else {
        int b = 0;
        while (b + 1 < kWfKeys && s_pre[b + 1] <= i) b++;
        slot = (int)W.buckets[(size_t)b * W.numSlots + (i - s_pre[b])];
      }
      const int px = slot % P.width, py = slot / P.width;
      PathState ps;
      SurfState sf;
      int r;
      if (W.first) {
        ps.si = 0;
        ps.medLo = ps.medHi = 0;
        r = pathBegin(P, ps, px, py) ? PN_TRACE : -1;
      } else {
        const int stage = wfLoadPath(W, slot, ps);
        if (stage == WF_HIT) {
          float4 r0 = wfLd(W, WP_RES0, slot), r1 = wfLd(W, WP_RES1, slot);
          Hit hit;
          hit.t = r0.x;
          hit.inst = (int)__float_as_uint(r0.y);
          hit.prim = (int)__float_as_uint(r0.z);
          hit.b1 = r0.w;
          hit.b2 = r1.x;
          const bool isHit = hit.inst >= 0;
          vec3 oObj = v3(0.0f), dObj = v3(0.0f);
          if (isHit) {
            const DevInstance& in = P.instances[hit.inst];
            oObj = xfPoint(in.inv, ps.o);
            dObj = xfVector(in.inv, ps.d);
          }
          r = pathOnHit(P, ps, isHit, hit, oObj, dObj, sf);
        } else {  // WF_SHADOW
          wfLoadSurf(W, slot, sf);
          r = pathOnShadow(P, ps, sf, f4xyz(wfLd(W, WP_RES0, slot)));
        }
      }
      while (r == PN_END) {
        wfFold(P, slot, ps);
        r = pathBegin(P, ps, px, py) ? PN_TRACE : -1;
      }
      if (r == PN_TRACE) {
        next = WF_HIT;
      } else if (r == PN_SHADOW) {
        next = WF_SHADOW;
        wfStoreSurf(W, slot, sf);
      }
      if (next != WF_DONE) wfStorePath(W, slot, ps, next);
    }
    wavePush(W.closestQ, &W.cnt->closestN, valid && next == WF_HIT, (uint32_t)slot);
    wavePush(W.shadowQ, &W.cnt->shadowN, valid && next == WF_SHADOW, (uint32_t)slot);
  }
}

// Closest-hit queries of the queue; each slot goes to the bucket of the
// material it will shade (its instance's front material kind) or of misses.
template <int kWaves>
__global__ __launch_bounds__(kBlock, kWaves) void k_wf_closest(LaunchParams P, WfParams W) {
  __shared__ uint64_t s_stack[kStackLds * kBlock];
  Stack st;
  st.lds = ldsColumn(s_stack);
  st.spill = reinterpret_cast<uint64_t*>(P.spill) + (size_t)blockIdx.x * kBlock + threadIdx.x;
  st.stride = P.spillStride;
  st.sp = 0;
  st.cap = kStackLds;
  WorkCount wc;
  uint32_t rays = 0;
  const uint32_t total = W.cnt->closestN;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t base = blockIdx.x * blockDim.x; base < total; base += stride) {
    const uint32_t i = base + threadIdx.x;
    const bool valid = i < total;
    int slot = 0, key = kKeyMiss;
    if (valid) {
      slot = (int)W.closestQ[i];
      vec3 o = f4xyz(wfLd(W, WP_O, slot)), d = f4xyz(wfLd(W, WP_D, slot));
      Hit hit;
      vec3 oo, dd;
      const bool isHit = traceClosest(P, st, o, d, 0.0f, 1e16f, hit, oo, dd, wc);
      rays++;
      const int inst = isHit ? hit.inst : -1;
      wfSt(W, WP_RES0, slot,
           make_float4(hit.t, __uint_as_float((uint32_t)inst), __uint_as_float((uint32_t)hit.prim), hit.b1));
      wfSt(W, WP_RES1, slot, make_float4(hit.b2, 0.0f, 0.0f, 0.0f));
      if (isHit) {
        // sort key: the kind of the instance's front material (a face
        // without a material, 0xFFFF, sorts with the first kind)
        const uint32_t front = P.instances[inst].mats & 0xFFFFu;
        key = kKeyHit + (front == 0xFFFFu ? 0 : P.materials[front].kind);
      }
    }
    wavePushKeyed(W.buckets, W.numSlots, W.cnt->bucketN, valid, key, (uint32_t)slot);
  }
  flushCounters<false>(W.rayCounter != nullptr, W.rayCounter, nullptr, rays, wc);
}

// Shadow queries of the queue; each slot goes to the bucket of the material
// its suspended bounce is shading.
template <int kWaves>
__global__ __launch_bounds__(kBlock, kWaves) void k_wf_shadow(LaunchParams P, WfParams W) {
  __shared__ uint64_t s_stack[kStackLds * kBlock];
  Stack st;
  st.lds = ldsColumn(s_stack);
  st.spill = reinterpret_cast<uint64_t*>(P.spill) + (size_t)blockIdx.x * kBlock + threadIdx.x;
  st.stride = P.spillStride;
  st.sp = 0;
  st.cap = kStackLds;
  WorkCount wc;
  uint32_t rays = 0;
  const uint32_t total = W.cnt->shadowN;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t base = blockIdx.x * blockDim.x; base < total; base += stride) {
    const uint32_t i = base + threadIdx.x;
    const bool valid = i < total;
    int slot = 0, key = kKeyShadow;
    if (valid) {
      slot = (int)W.shadowQ[i];
      SurfState sf;
      float4 a = wfLd(W, WP_SX, slot), b = wfLd(W, WP_SNS, slot), c = wfLd(W, WP_SALB, slot);
      float4 d = wfLd(W, WP_SWI, slot), e = wfLd(W, WP_SLI, slot);
      sf.x = f4xyz(a);
      sf.ns = f4xyz(b);
      sf.transmitNee = (wbits(c) >> 18) & 1u;
      sf.wi = f4xyz(d);
      sf.dist = e.w;
      vec3 so, wi;
      float tmax;
      shadowRay(sf, so, wi, tmax);
      vec3 tr = traceShadow(P, st, so, wi, tmax, wc);
      rays++;
      wfSt(W, WP_RES0, slot, f4(tr, 0.0f));
      key = kKeyShadow + P.materials[wbits(c) & 0xFFFFu].kind;
    }
    wavePushKeyed(W.buckets, W.numSlots, W.cnt->bucketN, valid, key, (uint32_t)slot);
  }
  flushCounters<false>(W.rayCounter != nullptr, W.rayCounter, nullptr, rays, wc);
}

// Clears the shading buckets (what = 0) or the two query queues (what = 1);
// a kernel rather than a memset so a pass can be graph-captured.
__global__ void k_wf_clear(WfCounters* c, int what) {
  int t = threadIdx.x;
  if (what == 0) {
    if (t < kWfKeys) c->bucketN[t] = 0;
  } else if (t == 0) {
    c->closestN = 0;
    c->shadowN = 0;
  }
}

// ---------------------------------------------------------------------------
// probe kernels (parity tests and measurement)
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(kBlock) void k_trace(LaunchParams P, const float* rays, int n, float* tOut,
                                                 int* instOut, int* primOut, int* counters) {
  __shared__ uint64_t s_stack[kStackLds * kBlock];
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  Stack st;
  st.lds = ldsColumn(s_stack);
  st.spill = reinterpret_cast<uint64_t*>(P.spill) + i;
  st.stride = P.spillStride;
  st.sp = 0;
  st.cap = kStackLds;
  const float* r = rays + 8 * (size_t)i;
  vec3 o = vec3{r[0], r[1], r[2]}, d = vec3{r[4], r[5], r[6]};
  Hit h;
  vec3 oo, dd;
  WorkCount wc;
  bool hit = traceClosest(P, st, o, d, r[3], r[7], h, oo, dd, wc);
  tOut[i] = hit ? h.t : __builtin_inff();
  instOut[i] = hit ? h.inst : -1;
  primOut[i] = hit ? h.prim : -1;
  if (counters) {
    counters[4 * i] = (int)wc.nodes;
    counters[4 * i + 1] = (int)wc.tris;
    counters[4 * i + 2] = (int)wc.quads;
    counters[4 * i + 3] = (int)wc.insts;
  }
}

// Brute force over every primitive of every instance (same tests, same tie
// rule), for the bit-exact traversal check.
__global__ __launch_bounds__(kBlock) void k_trace_brute(LaunchParams P, int numInst, const float* rays, int n,
                                                       float* tOut, int* instOut, int* primOut) {
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float* r = rays + 8 * (size_t)i;
  vec3 o = vec3{r[0], r[1], r[2]}, d = vec3{r[4], r[5], r[6]};
  float tmin = r[3];
  Hit h;
  h.t = r[7];
  h.inst = 0x7fffffff;
  h.prim = 0x7fffffff;
  bool found = false;
  for (int ii = 0; ii < numInst; ii++) {
    const DevInstance& in = P.instances[ii];
    vec3 oo = xfPoint(in.inv, o), dd = xfVector(in.inv, d);
    if (in.geomKind != GK_MESH) {
      float ts[2];
      int nr = quadricRoots(in.geomKind, oo, dd, tmin, candBound(h.t), ts);
      for (int k = 0; k < nr; k++) {
        if (closer(ts[k], ii, k, h) && maskAccepts(P, in, k, 0, 0, oo, dd, ts[k])) {
          h.t = ts[k];
          h.inst = ii;
          h.prim = k;
          found = true;
          break;
        }
      }
      continue;
    }
    const DevMesh& m = P.meshes[in.meshId];
    RayTri rt = rayTriSetup(dd);
    const float* pos = P.positions + 3 * m.posOffset;
    const uint32_t* idx = P.indices + m.idxOffset;
    for (int t = 0; t < m.numTris; t++) {
      vec3 v0 = ld3i(pos, idx[3 * t]), v1 = ld3i(pos, idx[3 * t + 1]), v2 = ld3i(pos, idx[3 * t + 2]);
      float tt, b1, b2;
      if (triHit(rt, oo, v0, v1, v2, tmin, candBound(h.t), tt, b1, b2) && closer(tt, ii, t, h) &&
          maskAccepts(P, in, t, b1, b2, oo, dd, tt)) {
        h.t = tt;
        h.inst = ii;
        h.prim = t;
        found = true;
      }
    }
  }
  tOut[i] = found ? h.t : __builtin_inff();
  instOut[i] = found ? h.inst : -1;
  primOut[i] = found ? h.prim : -1;
}

__global__ __launch_bounds__(kBlock) void k_shadow(LaunchParams P, const float* rays, int n, float* trOut) {
  __shared__ uint64_t s_stack[kStackLds * kBlock];
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  Stack st;
  st.lds = ldsColumn(s_stack);
  st.spill = reinterpret_cast<uint64_t*>(P.spill) + i;
  st.stride = P.spillStride;
  st.sp = 0;
  st.cap = kStackLds;
  const float* r = rays + 8 * (size_t)i;
  WorkCount wc;
  vec3 tr = traceShadow(P, st, vec3{r[0], r[1], r[2]}, vec3{r[4], r[5], r[6]}, r[7], wc);
  trOut[3 * i] = tr.x;
  trOut[3 * i + 1] = tr.y;
  trOut[3 * i + 2] = tr.z;
}

__global__ void k_env_bins(LaunchParams P, const float* u, int n, int* rowOut, int* colOut) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int row = envRowBin(P, u[2 * i]);
  const int col = envColBin(P, row, u[2 * i + 1]);
  rowOut[i] = row;
  colOut[i] = col;
}

// Wave-cooperative CDF search for comparison with the per-lane table the
// renderer uses: the wave serves its lanes' lookups one lane after another;
// for each, the 64 lanes load 64 evenly spaced keys strictly inside the
// current interval (lo, hi) and one ballot counts those <= u, which narrows
// the interval to the gap between two probes (a 65-ary step), keeping
// cdf[lo] <= u < cdf[hi] until hi = lo + 1. The answer, lo, is the largest
// i with cdf[i] <= u, bisection's bin. u is wave-uniform; every lane of the
// wave must call it.
DEV int cdfBinWave(const float* cdf, int nEntries, float u) {
  const int lane = threadIdx.x & 63;
  int lo = 0, hi = nEntries - 1;
  while (hi - lo > 1) {
    const int m = hi - lo - 1;  // keys strictly inside (lo, hi)
    const int cnt = m < 64 ? m : 64;
    auto probe = [&](int j) { return m <= 64 ? lo + 1 + j : lo + 1 + (int)(((long long)j * m) / 64); };
    const bool valid = lane < cnt;
    const float key = valid ? cdf[probe(lane)] : 0.0f;
    const int c = __popcll(__ballot(valid && key <= u));
    const int nlo = c > 0 ? probe(c - 1) : lo;
    const int nhi = c < cnt ? probe(c) : hi;
    lo = nlo;
    hi = nhi;
  }
  return lo;
}

__global__ __launch_bounds__(256) void k_env_bins_wave(LaunchParams P, const float* u, int n, int* rowOut,
                                                       int* colOut) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const bool mine = i < n;
  const float u0 = mine ? u[2 * i] : 0.0f, u1 = mine ? u[2 * i + 1] : 0.0f;
  const uint64_t active = __ballot(mine);
  int row = 0, col = 0;
  for (int l = 0; l < 64; l++) {
    if (!((active >> l) & 1ull)) continue;  // wave-uniform
    const float ur = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(u0), l));
    const float uc = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(u1), l));
    const int r = cdfBinWave(P.margCdf, P.envH + 1, ur);
    const int c = cdfBinWave(P.condCdf + (size_t)r * (P.envW + 1), P.envW + 1, uc);
    if (lane == l) {
      row = r;
      col = c;
    }
  }
  if (mine) {
    rowOut[i] = row;
    colOut[i] = col;
  }
}

// Light samples from n points: sample i draws from init(seed0 + i, seq).
// out: 12 floats per sample (wi, dist, Li, pdf, flags, draws, MIS-side pdf
// at the sampled point or NaN for mesh / delta lights, pad).
__global__ void k_light_sample(LaunchParams P, int light, const float* xs, int n, uint64_t seed0, uint64_t seq,
                               float* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Pcg rng = Pcg::make(seed0 + (uint64_t)i, seq);
  const uint64_t s0 = rng.state;
  const vec3 x = vec3{xs[3 * i], xs[3 * i + 1], xs[3 * i + 2]};
  const DevLight& lt = P.lights[light];
  const LightSample s = sampleLight(P, lt, x, rng);
  // draws taken: PCG steps between the two states (at most a handful)
  Pcg walk;
  walk.state = s0;
  walk.inc = rng.inc;
  int draws = 0;
  while (walk.state != rng.state && draws < 16) {
    walk.next();
    draws++;
  }
  float back = __builtin_nanf("");
  if (s.valid && !s.isDelta && lt.kind != LT_MESH) back = lightPdfSolidAngle(lt, x, x + s.wi * s.dist, v3(0.0f));
  float* o = out + 12 * (size_t)i;
  o[0] = s.wi.x;
  o[1] = s.wi.y;
  o[2] = s.wi.z;
  o[3] = s.dist;
  o[4] = s.Li.x;
  o[5] = s.Li.y;
  o[6] = s.Li.z;
  o[7] = s.valid ? s.pdf : 0.0f;
  o[8] = (float)((s.valid ? 1 : 0) | (s.isDelta ? 2 : 0));
  o[9] = (float)draws;
  o[10] = back;
  o[11] = 0.0f;
}

__global__ void k_flame_field(LaunchParams P, int flame, const float* pts, int n, float* sigmaOut, float* emOut) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float sg;
  vec3 em;
  flameField(P.flames[flame], vec3{pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]}, sg, em);
  sigmaOut[i] = sg;
  emOut[3 * i] = em.x;
  emOut[3 * i + 1] = em.y;
  emOut[3 * i + 2] = em.z;
}

__global__ void k_pcg(uint64_t seed, uint64_t seq, int n, uint32_t* out) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  Pcg r = Pcg::make(seed, seq);
  for (int i = 0; i < n; i++) out[i] = r.next();
}

// ---------------------------------------------------------------------------
// host-side launch helpers (called from session.hip)
// ---------------------------------------------------------------------------

// groups > 1 splits the launch's samples (at most kPoolSamples) over that
// many blocks per tile, then folds them with k_fold_samples: small frames
// fill the GPU. The spill region must hold groups x the grid's threads.
hipError_t launchRender(const LaunchParams& P, bool countWork, int waves, int groups, hipStream_t s) {
  dim3 grid((P.width + 15) / 16, (P.height + 15) / 16, groups);
  const size_t lds = (size_t)P.stackLds * kBlock * sizeof(uint64_t);
  // every wave of the grid owns one per-sample result region
  if (!P.sampleBuf || P.sampleBufWaves < (int64_t)grid.x * grid.y * (kBlock / 64)) return hipErrorInvalidValue;
  if (groups < 1 || groups > kPoolSamples || (groups > 1 && P.sppThisLaunch > kPoolSamples))
    return hipErrorInvalidValue;
  if (countWork)
    hipLaunchKernelGGL((k_render<true, 2>), grid, dim3(kBlock), lds, s, P);
  else if (waves >= 5)
    hipLaunchKernelGGL((k_render<false, 5>), grid, dim3(kBlock), lds, s, P);
  else if (waves == 4)
    hipLaunchKernelGGL((k_render<false, 4>), grid, dim3(kBlock), lds, s, P);
  else if (waves == 3)
    hipLaunchKernelGGL((k_render<false, 3>), grid, dim3(kBlock), lds, s, P);
  else
    hipLaunchKernelGGL((k_render<false, 2>), grid, dim3(kBlock), lds, s, P);
  if (groups > 1) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_fold_samples, dim3(grid.x, grid.y), dim3(kBlock), 0, s, P);
  }
  return hipGetLastError();
}

hipError_t launchTrace(const LaunchParams& P, const float* rays, int n, bool brute, int numInst, float* t,
                       int* inst, int* prim, int* counters, hipStream_t s) {
  dim3 grid((n + kBlock - 1) / kBlock);
  if (brute)
    hipLaunchKernelGGL(k_trace_brute, grid, dim3(kBlock), 0, s, P, numInst, rays, n, t, inst, prim);
  else
    hipLaunchKernelGGL(k_trace, grid, dim3(kBlock), 0, s, P, rays, n, t, inst, prim, counters);
  return hipGetLastError();
}

hipError_t launchShadow(const LaunchParams& P, const float* rays, int n, float* tr, hipStream_t s) {
  dim3 grid((n + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(k_shadow, grid, dim3(kBlock), 0, s, P, rays, n, tr);
  return hipGetLastError();
}

hipError_t launchEnvBins(const LaunchParams& P, const float* u, int n, int* row, int* col, hipStream_t s,
                         int mode) {
  if (mode == 1)
    hipLaunchKernelGGL(k_env_bins_wave, dim3((n + 255) / 256), dim3(256), 0, s, P, u, n, row, col);
  else
    hipLaunchKernelGGL(k_env_bins, dim3((n + 255) / 256), dim3(256), 0, s, P, u, n, row, col);
  return hipGetLastError();
}

hipError_t launchLightSample(const LaunchParams& P, int light, const float* xs, int n, uint64_t seed0, uint64_t seq,
                            float* out, hipStream_t s) {
  hipLaunchKernelGGL(k_light_sample, dim3((n + 255) / 256), dim3(256), 0, s, P, light, xs, n, seed0, seq, out);
  return hipGetLastError();
}

hipError_t launchFlameField(const LaunchParams& P, int flame, const float* pts, int n, float* sigma, float* em,
                            hipStream_t s) {
  hipLaunchKernelGGL(k_flame_field, dim3((n + 255) / 256), dim3(256), 0, s, P, flame, pts, n, sigma, em);
  return hipGetLastError();
}

hipError_t launchPcg(uint64_t seed, uint64_t seq, int n, uint32_t* out, hipStream_t s) {
  hipLaunchKernelGGL(k_pcg, dim3(1), dim3(64), 0, s, seed, seq, n, out);
  return hipGetLastError();
}

hipError_t launchQueueReset(QueueState* q, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_queue_reset, dim3((n + 255) / 256), dim3(256), 0, s, q, n);
  return hipGetLastError();
}

hipError_t launchQueuePrep(QueueState* q, RequestDev* reqs, const int32_t* firstSample, int numReq,
                           hipStream_t s) {
  hipLaunchKernelGGL(k_queue_prep, dim3(1), dim3(256), 0, s, q, reqs, firstSample, numReq);
  return hipGetLastError();
}

hipError_t launchPersistent(const SceneTable& S, const RequestDev* reqs, const WorkItem* items, int numItems,
                            QueueState* q, int tileSize, int waves, int blocks, uint32_t* spill, int spillStride,
                            int stackLds, float* sampleBuf, int64_t sampleBufWaves, float* groupBuf,
                            unsigned long long* rayCounter, hipStream_t s) {
  const size_t lds = (size_t)stackLds * kBlock * sizeof(uint64_t);
  // every wave of the grid owns one per-sample result region
  if (!sampleBuf || sampleBufWaves < (int64_t)blocks * (kBlock / 64)) return hipErrorInvalidValue;
  if (waves >= 5)
    hipLaunchKernelGGL((k_persistent<5>), dim3(blocks), dim3(kBlock), lds, s, S, reqs, items, numItems, q,
                       tileSize, spill, spillStride, stackLds, sampleBuf, groupBuf, rayCounter);
  else if (waves == 4)
    hipLaunchKernelGGL((k_persistent<4>), dim3(blocks), dim3(kBlock), lds, s, S, reqs, items, numItems, q,
                       tileSize, spill, spillStride, stackLds, sampleBuf, groupBuf, rayCounter);
  else if (waves == 3)
    hipLaunchKernelGGL((k_persistent<3>), dim3(blocks), dim3(kBlock), lds, s, S, reqs, items, numItems, q,
                       tileSize, spill, spillStride, stackLds, sampleBuf, groupBuf, rayCounter);
  else
    hipLaunchKernelGGL((k_persistent<2>), dim3(blocks), dim3(kBlock), lds, s, S, reqs, items, numItems, q,
                       tileSize, spill, spillStride, stackLds, sampleBuf, groupBuf, rayCounter);
  return hipGetLastError();
}

hipError_t launchFoldGroups(const RequestDev* reqs, const WorkItem* items, int numItems, int tileSize,
                            const float* groupBuf, hipStream_t s) {
  if (numItems <= 0) return hipSuccess;
  if (tileSize != 8 && tileSize != 16 && tileSize != 32) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_fold_groups, dim3((unsigned)numItems * (tileSize * tileSize / 64)), dim3(64), 0, s, reqs,
                     items, tileSize, groupBuf);
  return hipGetLastError();
}

// Resident blocks per CU of the persistent kernel at a given occupancy variant.
int persistentBlocksPerCU(int waves, int stackLds) {
  int n = 0;
  hipError_t e;
  const size_t lds = (size_t)stackLds * kBlock * sizeof(uint64_t);
  if (waves >= 5)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_persistent<5>, kBlock, lds);
  else if (waves == 4)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_persistent<4>, kBlock, lds);
  else if (waves == 3)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_persistent<3>, kBlock, lds);
  else
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_persistent<2>, kBlock, lds);
  return (e == hipSuccess && n > 0) ? n : 1;
}

int stackLdsEntries() { return kStackLds; }
int stackLdsDefault() { return kStackLdsMegakernel; }
int stackLdsMax() { return kStackLdsMax; }
int renderBlockThreads() { return kBlock; }

hipError_t launchWfShade(const LaunchParams& P, const WfParams& W, int blocks, hipStream_t s) {
  hipLaunchKernelGGL(k_wf_shade, dim3(blocks), dim3(kBlock), 0, s, P, W);
  return hipGetLastError();
}
hipError_t launchWfClosest(const LaunchParams& P, const WfParams& W, int blocks, int waves, hipStream_t s) {
  if (waves >= 5)
    hipLaunchKernelGGL(k_wf_closest<5>, dim3(blocks), dim3(kBlock), 0, s, P, W);
  else
    hipLaunchKernelGGL(k_wf_closest<4>, dim3(blocks), dim3(kBlock), 0, s, P, W);
  return hipGetLastError();
}
hipError_t launchWfShadow(const LaunchParams& P, const WfParams& W, int blocks, int waves, hipStream_t s) {
  if (waves >= 5)
    hipLaunchKernelGGL(k_wf_shadow<5>, dim3(blocks), dim3(kBlock), 0, s, P, W);
  else
    hipLaunchKernelGGL(k_wf_shadow<4>, dim3(blocks), dim3(kBlock), 0, s, P, W);
  return hipGetLastError();
}
hipError_t launchWfClear(WfCounters* c, int what, hipStream_t s) {
  hipLaunchKernelGGL(k_wf_clear, dim3(1), dim3(64), 0, s, c, what);
  return hipGetLastError();
}
// Resident blocks per CU of each wavefront kernel: 0 shade, 1 closest, 2
// shadow (closest and shadow at the given waves-per-SIMD variant).
int wfBlocksPerCU(int which, int waves) {
  int n = 0;
  hipError_t e;
  if (which == 0)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_wf_shade, kBlock, 0);
  else if (which == 1)
    e = waves >= 5 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_wf_closest<5>, kBlock, 0)
                   : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_wf_closest<4>, kBlock, 0);
  else
    e = waves >= 5 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_wf_shadow<5>, kBlock, 0)
                   : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_wf_shadow<4>, kBlock, 0);
  return (e == hipSuccess && n > 0) ? n : 1;
}

}  // namespace sdk
