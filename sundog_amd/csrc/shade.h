// Device shading library of the gfx950 path tracer: random streams and
// warps, textures, value noise, flame volumes, the environment light and its
// table search, surface closures (BSDF sample / evaluation), emitters, and
// the analytic and triangle intersectors.
//
// The mathematics follows the reference's device headers so that the GPU and
// the CPU oracle see the same PCG32 streams, draw order and formulas:
// device/rng.cuh:10-61, device/texture_eval.cuh:11-46, device/noise.cuh:13-79,
// device/volume.cuh:27-153, device/env_light.cuh:17-116, device/bsdf.cuh:27-424,
// device/light_sample.cuh:24-226, device/intersect.cuh:37-162 and :382-500.
// The structure is this library's own. A hit is turned once into a
// `Closure` (lobe kind, GGX width, relative IOR, Fresnel base and tint) so
// that per-material constants are derived once per vertex, not once per
// call; every lobe works in the shading frame; one fused evaluation returns
// f and pdf together (the NEE connection needs both, and the coated lobe's
// sample reuses it), sharing the half vector, D and the Smith terms. Area
// emitters share one area-to-solid-angle conversion, flames one marcher for
// emission and transmittance, and the environment's inverse-CDF lookups run
// on a 16-ary table (three 64-byte loads per lookup instead of eleven
// dependent loads) that returns exactly the binary search's bins.
#pragma once

#include <hip/hip_runtime.h>

#include "dev.h"

namespace sdk {

#define SDEV __device__ __forceinline__

SDEV vec3 ld3(const float* p) { return vec3{p[0], p[1], p[2]}; }
SDEV vec3 ld3i(const float* pool, int i) { return vec3{pool[3 * i], pool[3 * i + 1], pool[3 * i + 2]}; }
SDEV vec3 f4xyz(float4 a) { return vec3{a.x, a.y, a.z}; }
SDEV bool finite3(vec3 c) { return isfinite(c.x) && isfinite(c.y) && isfinite(c.z); }
SDEV vec3 sanitize(vec3 c) { return finite3(c) ? c : v3(0.0f); }

// ===========================================================================
// random streams: PCG-XSH-RR 64/32 (O'Neill 2014, pcg_basic's srandom
// seeding); one stream per path keyed (pixel << 32 ^ sample, render seed)
// ===========================================================================

struct Pcg {
  uint64_t state, inc;
  static constexpr uint64_t kMul = 6364136223846793005ULL;
  SDEV uint32_t next() {
    const uint64_t s = state;
    state = s * kMul + inc;
    const uint32_t x = (uint32_t)(((s >> 18u) ^ s) >> 27u);
    const uint32_t r = (uint32_t)(s >> 59u);
    return (x >> r) | (x << ((32u - r) & 31u));
  }
  SDEV static Pcg make(uint64_t seed, uint64_t seq) {
    Pcg g{0u, (seq << 1u) | 1u};
    g.next();
    g.state += seed;
    g.next();
    return g;
  }
  // 24 high bits scaled into [0, 1)
  SDEV float rnd() { return (float)(next() >> 8) * 0x1p-24f; }
  SDEV vec2 rnd2() {
    const float a = rnd();
    return vec2{a, rnd()};
  }
};

// Shirley-Chiu concentric square-to-disk map.
SDEV vec2 concentricDisk(vec2 u) {
  const float a = 2.0f * u.x - 1.0f, b = 2.0f * u.y - 1.0f;
  if (a == 0.0f && b == 0.0f) return vec2{0.0f, 0.0f};
  const bool xMajor = fabsf(a) > fabsf(b);
  const float r = xMajor ? a : b;
  const float phi = xMajor ? (kPi / 4.0f) * (b / a) : (kPi / 2.0f) - (kPi / 4.0f) * (a / b);
  return vec2{r * cosf(phi), r * sinf(phi)};
}
// Malley: project the concentric disk up to the hemisphere.
SDEV vec3 cosineHemisphere(vec2 u) {
  const vec2 d = concentricDisk(u);
  return vec3{d.x, d.y, sqrtf(fmaxf(0.0f, 1.0f - d.x * d.x - d.y * d.y))};
}
// Archimedes: uniform z, uniform azimuth.
SDEV vec3 uniformSphere(vec2 u) {
  const float z = 1.0f - 2.0f * u.x;
  const float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
  const float phi = 2.0f * kPi * u.y;
  return vec3{r * cosf(phi), r * sinf(phi), z};
}

// ===========================================================================
// textures: solid, two-colour squares, grid lines, bilinear images (linear
// float texels; wrap in u, wrap or clamp in v)
// ===========================================================================

SDEV int wrapIndex(int i, int n) {
  const int r = i % n;
  return r < 0 ? r + n : r;
}

SDEV float4 bilinear(const float4* texels, int w, int h, float u, float v, bool wrapV) {
  u -= floorf(u);
  if (wrapV) v -= floorf(v);
  // texel centres at (i + 0.5) / n; clamped so that NaN and far-out
  // coordinates still convert to in-range integers
  const float x = fminf(fmaxf(u * (float)w - 0.5f, -1.0f), (float)w);
  const float y = fminf(fmaxf(v * (float)h - 0.5f, -1.0f), (float)h);
  const float x0f = floorf(x), y0f = floorf(y);
  const float ax = x - x0f, ay = y - y0f;
  const int xi = (int)x0f, yi = (int)y0f;
  const int c0 = wrapIndex(xi, w), c1 = wrapIndex(xi + 1, w);
  const int r0 = wrapV ? wrapIndex(yi, h) : min(max(yi, 0), h - 1);
  const int r1 = wrapV ? wrapIndex(yi + 1, h) : min(max(yi + 1, 0), h - 1);
  const float4 p00 = texels[(size_t)r0 * w + c0], p10 = texels[(size_t)r0 * w + c1];
  const float4 p01 = texels[(size_t)r1 * w + c0], p11 = texels[(size_t)r1 * w + c1];
  const float w00 = (1.0f - ax) * (1.0f - ay), w10 = ax * (1.0f - ay);
  const float w01 = (1.0f - ax) * ay, w11 = ax * ay;
  return make_float4(p00.x * w00 + p10.x * w10 + p01.x * w01 + p11.x * w11,
                     p00.y * w00 + p10.y * w10 + p01.y * w01 + p11.y * w11,
                     p00.z * w00 + p10.z * w10 + p01.z * w01 + p11.z * w11,
                     p00.w * w00 + p10.w * w10 + p01.w * w01 + p11.w * w11);
}

SDEV float4 rgb1(const float* c) { return make_float4(c[0], c[1], c[2], 1.0f); }

SDEV float4 evalTexture(const LaunchParams& P, int id, float u, float v) {
  const DevTexture& t = P.textures[id];
  if (t.kind == TX_SOLID) return rgb1(t.a);
  if (t.kind == TX_IMAGE) return bilinear(P.texels + t.imgOffset, t.imgW, t.imgH, u, 1.0f - v, true);
  const float su = u * t.scale[0], sv = v * t.scale[1];
  if (t.kind == TX_CHECKER) {
    const int parity = ((int)floorf(su) + (int)floorf(sv)) & 1;
    return rgb1(parity == 0 ? t.a : t.b);
  }
  if (t.kind == TX_GRID) {
    const float fu = su - floorf(su), fv = sv - floorf(sv);
    const float lw = t.width;
    const bool onLine = fu < lw || fv < lw || fu > 1.0f - lw || fv > 1.0f - lw;
    return rgb1(onLine ? t.b : t.a);
  }
  return make_float4(1.0f, 0.0f, 1.0f, 1.0f);  // unknown kind: magenta
}

SDEV vec3 surfaceColor(const LaunchParams& P, const DevMaterial& m, float u, float v) {
  if (m.texId < 0) return ld3(m.color);
  const float4 c = evalTexture(P, m.texId, u, v);
  return vec3{c.x, c.y, c.z};
}

// ===========================================================================
// value noise: Teschner et al. 2003 spatial hash of the lattice point,
// scrambled by the PCG hash (Jarzynski & Olano 2020), smoothstep-weighted
// trilinear interpolation; fbm sums octaves at lacunarity 2.03, gain 1/2
// ===========================================================================

SDEV float smoothstepf(float e0, float e1, float x) {
  const float t = clampf((x - e0) / (e1 - e0), 0.0f, 1.0f);
  return t * t * (3.0f - 2.0f * t);
}
SDEV uint32_t pcgHash(uint32_t x) {
  x = x * 747796405u + 2891336453u;
  x = ((x >> ((x >> 28u) + 4u)) ^ x) * 277803737u;
  return (x >> 22u) ^ x;
}
SDEV float latticeValue(int x, int y, int z, uint32_t seed) {
  const uint32_t k = (uint32_t)x * 73856093u ^ (uint32_t)y * 19349663u ^ (uint32_t)z * 83492791u ^ seed;
  return (float)pcgHash(k) * (1.0f / 4294967296.0f);
}
SDEV float fade(float t) { return t * t * (3.0f - 2.0f * t); }

SDEV float valueNoise(vec3 p, uint32_t seed) {
  const float gx = floorf(p.x), gy = floorf(p.y), gz = floorf(p.z);
  const int ix = (int)gx, iy = (int)gy, iz = (int)gz;
  const float wx = fade(p.x - gx), wy = fade(p.y - gy), wz = fade(p.z - gz);
  // the 8 corners, interpolated along x, then y, then z
  float edge[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int dy = k & 1, dz = k >> 1;
    const float a = latticeValue(ix, iy + dy, iz + dz, seed);
    const float b = latticeValue(ix + 1, iy + dy, iz + dz, seed);
    edge[k] = a + (b - a) * wx;
  }
  const float f0 = edge[0] + (edge[1] - edge[0]) * wy;
  const float f1 = edge[2] + (edge[3] - edge[2]) * wy;
  return f0 + (f1 - f0) * wz;
}

constexpr float kLacunarity = 2.03f;
SDEV vec3 nextOctave(vec3 p) { return kLacunarity * p + vec3{11.7f, 5.3f, 7.9f}; }
SDEV uint32_t octaveSeed(uint32_t seed, int i) { return seed + (uint32_t)i * 101u; }

SDEV float fbm(vec3 p, uint32_t seed, int octaves) {
  float amp = 0.5f, sum = 0.0f;
  for (int i = 0; i < octaves; i++) {
    sum += amp * valueNoise(p, octaveSeed(seed, i));
    p = nextOctave(p);
    amp *= 0.5f;
  }
  return sum;
}

// Water ripples: the height field y = (amp / freq) fbm(freq x, 0, freq z),
// its slope by central differences, the normal kept on the incoming side and
// dropped when it would face away from the viewer.
SDEV vec3 rippleNormal(vec3 x, vec3 ns, vec3 wo, float amp, float freq, int octaves) {
  constexpr uint32_t kSeed = 1337u;
  constexpr float h = 0.05f;
  const vec3 q = vec3{x.x * freq, 0.0f, x.z * freq};
  const float sx = (fbm(q + vec3{h, 0.0f, 0.0f}, kSeed, octaves) - fbm(q - vec3{h, 0.0f, 0.0f}, kSeed, octaves)) /
                   (2.0f * h);
  const float sz = (fbm(q + vec3{0.0f, 0.0f, h}, kSeed, octaves) - fbm(q - vec3{0.0f, 0.0f, h}, kSeed, octaves)) /
                   (2.0f * h);
  const float side = ns.y >= 0.0f ? 1.0f : -1.0f;
  const vec3 n = normalize(vec3{-amp * sx, 1.0f, -amp * sz}) * side;
  return dot(n, wo) <= 1e-4f ? ns : n;
}

// ===========================================================================
// flame volumes: procedural density / emission, bounds clipping, and one
// marcher for both the radiance march and shadow transmittance
// ===========================================================================

// Density sigma and emission at p. The noise is a sum of non-negative
// octaves and adding a non-negative term never lowers a float sum, so every
// partial sum bounds the shaped distance dd from below: once
// a + (partial - 0.45) >= 1.01, dd >= 1 and the mask is 0 whatever the
// remaining octaves (0.01 of margin for rounding), and they are skipped.
//
// Returns whether the field is non-zero; when it returns false, sigma and
// emission are 0 (a march step there changes nothing). (1 - h)^1.55 is
// 2^(1.55 log2(1 - h)) on the hardware log2 / exp2 (1 - h >= 2^-24 is
// normal here; within 1e-6 relative of powf, far inside the parity
// tolerance, and the cut-offs below sit where the mask is already 0).
SDEV bool flameField(const DevFlame& fl, vec3 p, float& sigma, vec3& emission) {
  sigma = 0.0f;
  emission = v3(0.0f);
  const float h = (p.y - fl.base[1]) / fl.height;
  if (h <= 0.0f || h >= 1.0f) return false;
  const float rx = p.x - fl.base[0], rz = p.z - fl.base[2];
  const float r = sqrtf(rx * rx + rz * rz) / fl.radius;
  const float profile = 3.2f * sqrtf(h) * __builtin_amdgcn_exp2f(1.55f * __builtin_amdgcn_logf(1.0f - h));
  const float a = r / fmaxf(profile, 1e-3f);
  if (a >= 1.46f) return false;
  vec3 q = vec3{rx / fl.radius, (p.y - fl.base[1]) / fl.radius * 0.32f, rz / fl.radius} * fl.noiseScale;
  float amp = 0.5f, n = 0.0f;
  for (int i = 0; i < fl.octaves; i++) {
    n += amp * valueNoise(q, octaveSeed(fl.seed, i));
    if (a + (n - 0.45f) >= 1.01f) return false;
    q = nextOctave(q);
    amp *= 0.5f;
  }
  const float dd = a + (n - 0.45f) * 1.0f;
  const float mask = smoothstepf(1.0f, 0.62f, dd) * smoothstepf(0.0f, 0.12f, h);
  if (mask <= 0.0f) return false;
  sigma = fl.sigma * mask;
  const float heat = clampf((0.88f - dd) * (1.25f - h), 0.0f, 1.0f);
  const float h2 = heat * heat;
  const vec3 ramp = vec3{fminf(heat * 2.6f, 1.0f), fminf(h2 * 2.1f, 1.0f), fminf(h2 * heat * 1.7f, 1.0f)};
  emission = ramp * (fl.intensity * mask);
  return true;
}

// Narrows [t0, t1] to the ray's overlap with one axis slab [lo, hi]; false
// when a ray parallel to the slab lies outside it.
SDEV bool clipSlab(float o, float d, float lo, float hi, float& t0, float& t1) {
  if (fabsf(d) < 1e-8f) return o >= lo && o <= hi;
  float ta = (lo - o) / d, tb = (hi - o) / d;
  if (ta > tb) {
    const float s = ta;
    ta = tb;
    tb = s;
  }
  t0 = fmaxf(t0, ta);
  t1 = fminf(t1, tb);
  return true;
}

// The ray's overlap with the flame's bounds: an upright cylinder or an
// axis-aligned box of half-width `radius`, both `height` tall.
SDEV bool clipFlame(const DevFlame& fl, vec3 o, vec3 d, float& t0, float& t1) {
  if (fabsf(d.y) < 1e-8f) {
    if (o.y < fl.base[1] || o.y > fl.base[1] + fl.height) return false;
  } else {
    clipSlab(o.y, d.y, fl.base[1], fl.base[1] + fl.height, t0, t1);
  }
  const float ox = o.x - fl.base[0], oz = o.z - fl.base[2], r = fl.radius;
  if (fl.bounds == 1) {
    if (!clipSlab(ox, d.x, -r, r, t0, t1) || !clipSlab(oz, d.z, -r, r, t0, t1)) return false;
    return t0 < t1;
  }
  const float a = d.x * d.x + d.z * d.z;
  if (a < 1e-12f) {
    if (ox * ox + oz * oz > r * r) return false;
  } else {
    const float b = ox * d.x + oz * d.z;
    const float c = ox * ox + oz * oz - r * r;
    const float disc = b * b - a * c;
    if (disc <= 0.0f) return false;
    const float s = sqrtf(disc);
    t0 = fmaxf(t0, (-b - s) / a);
    t1 = fminf(t1, (-b + s) / a);
  }
  return t0 < t1;
}

SDEV int marchSteps(const DevFlame& fl, float len) {
  if (fl.step <= 0.0f) return 32;
  return (int)fminf(fmaxf(ceilf(len / fl.step), 1.0f), 65536.0f);
}

// Marches every flame the segment (o, d, [1e-4, tEnd]) enters, skipping
// flame `skip`, with one jitter draw per entered flame; returns the product
// of the flames' transmittances and, with kEmit, adds the emission seen
// through the flames already crossed to `radiance`.
template <bool kEmit>
SDEV float marchVolumes(const LaunchParams& P, vec3 o, vec3 d, float tEnd, Pcg& rng, int skip, vec3& radiance) {
  float T = 1.0f;
  for (int i = 0; i < P.numFlames; i++) {
    if (i == skip) continue;
    const DevFlame& fl = P.flames[i];
    float t0 = 1e-4f, t1 = tEnd;
    if (!clipFlame(fl, o, d, t0, t1)) continue;
    const int steps = marchSteps(fl, t1 - t0);
    const float dt = (t1 - t0) / steps;
    const float jitter = rng.rnd();
    float tl = 1.0f;
    vec3 acc = v3(0.0f);
    for (int k = 0; k < steps; k++) {
      float sigma;
      vec3 em;
      // a zero field adds +0 to acc (which starts at +0 and only grows) and
      // multiplies tl by exp(-0) = 1: skipping the step changes no bit
      if (!flameField(fl, o + (t0 + (k + jitter) * dt) * d, sigma, em)) continue;
      if (kEmit) acc += (tl * dt) * em;
      tl *= expf(-sigma * dt);
      if (tl < 1e-4f) break;
    }
    if (kEmit) radiance += v3(T) * acc;
    T *= tl;
  }
  return T;
}

// Transmittance of the flames along a shadow segment (the light's own flame
// excluded).
SDEV float volumeTransmittance(const LaunchParams& P, vec3 o, vec3 d, float tEnd, Pcg& rng, int skip) {
  vec3 none = v3(0.0f);
  return marchVolumes<false>(P, o, d, tEnd, rng, skip, none);
}

// ===========================================================================
// environment light: equirect lookup, importance sampling on
// luminance * sin(theta) through marginal / conditional CDFs
// ===========================================================================

SDEV vec2 envUv(const LaunchParams& P, vec3 dir) {
  float u = (atan2f(dir.z, dir.x) - P.envRotate + kPi) * (0.5f * kInvPi);
  u -= floorf(u);
  return vec2{u, acosf(clampf(dir.y, -1.0f, 1.0f)) * kInvPi};
}
SDEV vec3 envRadiance(const LaunchParams& P, vec3 dir) {
  const vec2 uv = envUv(P, dir);
  const float4 c = bilinear(P.envTex, P.envW, P.envH, uv.x, uv.y, false);
  return vec3{c.x, c.y, c.z} * P.envIntensity;
}

// Bin of u in a CDF of n entries: the largest i in [0, n-2] with
// cdf[i] <= u. The CDFs are non-decreasing, so that bin is the count of
// keys cdf[1..n-2] that are <= u. The search table holds those keys in
// 16-float blocks under two levels of 16-float splitter blocks (first key of
// every following block, +inf padding): one block per level, so a lookup
// issues three 64-byte loads whose 16 compares are independent.
SDEV int count16(const float4* blk, float u) {
  int c = 0;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const float4 v = blk[q];
    c += (v.x <= u) + (v.y <= u) + (v.z <= u) + (v.w <= u);
  }
  return c;
}
SDEV int cdfBin16(const float4* table, int midBlocks, float u) {
  const int c = count16(table, u);
  const int b = 16 * c + count16(table + 4 * (1 + c), u);
  return 16 * b + count16(table + 4 * (1 + midBlocks + b), u);
}
// The same bin by bisection, for tables too large for three levels.
SDEV int cdfBinBisect(const float* cdf, int n, float u) {
  int lo = 0, hi = n - 1;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (cdf[mid] <= u) lo = mid; else hi = mid;
  }
  return lo;
}
SDEV int envRowBin(const LaunchParams& P, float u) {
  if (P.envSearch) return cdfBin16(P.envSearch, P.envSearchMargMid, u);
  return cdfBinBisect(P.margCdf, P.envH + 1, u);
}
SDEV int envColBin(const LaunchParams& P, int row, float u) {
  if (P.envSearch)
    return cdfBin16(P.envSearch + P.envSearchMargBlocks * 4 + (size_t)row * P.envSearchRowBlocks * 4,
                    P.envSearchRowMid, u);
  return cdfBinBisect(P.condCdf + (size_t)row * (P.envW + 1), P.envW + 1, u);
}

struct LightSample {
  vec3 wi;
  float dist;
  vec3 Li;
  float pdf;
  bool isDelta;
  bool valid;
};

SDEV LightSample noLight(float dist) {
  LightSample s;
  s.wi = v3(0.0f);
  s.dist = dist;
  s.Li = v3(0.0f);
  s.pdf = 0.0f;
  s.isDelta = false;
  s.valid = false;
  return s;
}

// Solid-angle density of the importance sampler for direction (u, v):
// the product of the two table bins' probabilities over the Jacobian of the
// equirect map, 2 pi^2 sin(theta) / (W H).
SDEV float envDensity(const LaunchParams& P, float pm, float pc, float sinT) {
  return pm * pc * P.envW * P.envH / (2.0f * kPi * kPi * sinT);
}

// Two draws in either mode (uniform sphere, or inverse CDF).
SDEV LightSample sampleEnv(const LaunchParams& P, Pcg& rng) {
  LightSample s = noLight(1e16f);
  const vec2 r = rng.rnd2();
  if (!P.envImportance) {
    s.wi = uniformSphere(r);
    s.pdf = 0.25f * kInvPi;
    s.Li = envRadiance(P, s.wi);
    s.valid = true;
    return s;
  }
  const int W = P.envW, H = P.envH;
  const int row = envRowBin(P, r.x);
  const float m0 = P.margCdf[row], pm = P.margCdf[row + 1] - m0;
  const float* cond = P.condCdf + (size_t)row * (W + 1);
  const int col = envColBin(P, row, r.y);
  const float c0 = cond[col], pc = cond[col + 1] - c0;
  // continuous position inside the chosen texel
  const float v = (row + (pm > 0.0f ? (r.x - m0) / pm : 0.5f)) / H;
  const float u = (col + (pc > 0.0f ? (r.y - c0) / pc : 0.5f)) / W;
  const float theta = v * kPi;
  const float sinT = sinf(theta);
  if (sinT < 1e-6f || pm <= 0.0f || pc <= 0.0f) return s;
  const float phi = (2.0f * kPi * u - kPi) + P.envRotate;
  s.wi = vec3{sinT * cosf(phi), cosf(theta), sinT * sinf(phi)};
  s.pdf = envDensity(P, pm, pc, sinT);
  s.Li = envRadiance(P, s.wi);
  s.valid = s.pdf > 0.0f;
  return s;
}

// The sampler's density for a direction reached by BSDF sampling.
SDEV float envPdf(const LaunchParams& P, vec3 dir) {
  if (!P.envImportance) return 0.25f * kInvPi;
  const float sinT = sqrtf(fmaxf(0.0f, 1.0f - dir.y * dir.y));
  if (!(sinT >= 1e-6f)) return 0.0f;  // also rejects NaN directions
  const vec2 uv = envUv(P, dir);
  const int W = P.envW, H = P.envH;
  // fmaxf drops NaN, so the texel index stays inside the image
  const int col = (int)fminf(fmaxf(uv.x * W, 0.0f), (float)(W - 1));
  const int row = (int)fminf(fmaxf(uv.y * H, 0.0f), (float)(H - 1));
  const float* cond = P.condCdf + (size_t)row * (W + 1);
  return envDensity(P, P.margCdf[row + 1] - P.margCdf[row], cond[col + 1] - cond[col], sinT);
}

// ===========================================================================
// surface closures
// ===========================================================================

enum LobeKind : int32_t {
  LK_NONE = 0,       // emissive: the path ends at the hit
  LK_DIFFUSE = 1,    // lambert
  LK_MIRROR = 2,     // smooth metal (delta)
  LK_INTERFACE = 3,  // smooth dielectric or water (delta reflect / refract)
  LK_CONDUCTOR = 4,  // rough metal: GGX reflection, Schlick F0 = tint
  LK_GLASS = 5,      // rough dielectric: GGX reflection + transmission
  LK_COATED = 6,     // plastic: GGX dielectric coat over a Fresnel-coupled lambert base
};

// Below this roughness metals and dielectrics are smooth (delta) lobes.
constexpr float kSmoothRoughness = 1e-3f;

struct Closure {
  int32_t lobe;
  float alpha;  // GGX width
  float eta;    // relative IOR seen from the incoming side (interfaces)
  float f0;     // Fresnel reflectance at normal incidence (dielectrics, coat)
  vec3 tint;    // albedo / metal colour
};

// etaRel: outside / inside IOR as seen from the incoming ray (dielectrics
// and water), frontface ? etaExt / ior : ior / etaExt.
SDEV float dielectricF0(float eta) {
  const float r = (1.0f - eta) / (1.0f + eta);
  return r * r;
}
SDEV Closure makeClosure(const DevMaterial& m, vec3 tint, float etaRel) {
  Closure c;
  c.tint = tint;
  c.eta = etaRel;
  c.alpha = m.roughness * m.roughness;
  c.f0 = 0.0f;  // conductors take F0 from the tint, lambert has none
  switch (m.kind) {
    case MT_LAMBERT:
      c.lobe = LK_DIFFUSE;
      break;
    case MT_METAL:
      c.lobe = m.roughness < kSmoothRoughness ? LK_MIRROR : LK_CONDUCTOR;
      break;
    case MT_DIELECTRIC:
      c.lobe = m.roughness < kSmoothRoughness ? LK_INTERFACE : LK_GLASS;
      c.f0 = dielectricF0(etaRel);
      break;
    case MT_WATER:  // always smooth: its waves come from the ripple normal
      c.lobe = LK_INTERFACE;
      c.f0 = dielectricF0(etaRel);
      break;
    case MT_PLASTIC: {
      c.lobe = LK_COATED;
      // the coat is always glossy (never delta) and always seen from air:
      // its Fresnel base comes from the material's IOR, not the medium stack
      const float rc = fmaxf(m.roughness, kSmoothRoughness);
      c.alpha = rc * rc;
      const float q = (m.ior - 1.0f) / (m.ior + 1.0f);
      c.f0 = q * q;
      break;
    }
    default:
      c.lobe = LK_NONE;
  }
  return c;
}
SDEV bool closureIsDelta(const Closure& c) { return c.lobe == LK_MIRROR || c.lobe == LK_INTERFACE; }
SDEV bool materialIsDelta(const DevMaterial& m) {
  return m.kind == MT_WATER || ((m.kind == MT_DIELECTRIC || m.kind == MT_METAL) && m.roughness < kSmoothRoughness);
}

// GGX (Trowbridge-Reitz) with Smith masking and visible-normal sampling
// (Heitz 2018), all in the local frame (n = +z).
struct Ggx {
  float a, a2;
  SDEV explicit Ggx(float alpha) : a(alpha), a2(alpha * alpha) {}
  // written as a2 c^2 + (1 - c^2) so that k >= a2 > 0 even when c = 1
  SDEV float D(vec3 h) const {
    const float c2 = h.z * h.z;
    const float k = a2 * c2 + (1.0f - c2);
    return a2 / (kPi * k * k);
  }
  SDEV float lambda(vec3 w) const {
    const float c2 = w.z * w.z;
    if (c2 >= 1.0f) return 0.0f;
    const float t2 = (1.0f - c2) / c2;
    return 0.5f * (-1.0f + sqrtf(1.0f + a2 * t2));
  }
  SDEV float G1(vec3 w) const { return 1.0f / (1.0f + lambda(w)); }
  SDEV float G(vec3 wo, vec3 wi) const { return 1.0f / (1.0f + lambda(wo) + lambda(wi)); }
  SDEV vec3 sampleVisible(vec3 wo, vec2 u) const {
    const vec3 vh = normalize(vec3{a * wo.x, a * wo.y, wo.z});
    const float l2 = vh.x * vh.x + vh.y * vh.y;
    const vec3 t1v = l2 > 0.0f ? vec3{-vh.y, vh.x, 0.0f} * (1.0f / sqrtf(l2)) : vec3{1.0f, 0.0f, 0.0f};
    const vec3 t2v = cross(vh, t1v);
    const float r = sqrtf(u.x), phi = 2.0f * kPi * u.y;
    const float p1 = r * cosf(phi);
    const float s = 0.5f * (1.0f + vh.z);
    const float p2 = (1.0f - s) * sqrtf(fmaxf(0.0f, 1.0f - p1 * p1)) + s * (r * sinf(phi));
    const vec3 nh = p1 * t1v + p2 * t2v + sqrtf(fmaxf(0.0f, 1.0f - p1 * p1 - p2 * p2)) * vh;
    return normalize(vec3{a * nh.x, a * nh.y, fmaxf(1e-6f, nh.z)});
  }
  // density of a reflected direction under visible-normal sampling
  SDEV float pdfReflect(vec3 wo, vec3 wi) const {
    if (wo.z <= 0.0f || wi.z <= 0.0f) return 0.0f;
    return G1(wo) * D(normalize(wo + wi)) / (4.0f * wo.z);
  }
  // density of a refracted direction: visible normals times the refraction
  // Jacobian |wi.h| / (eta (wo.h) + wi.h)^2 (Walter et al. 2007)
  SDEV float pdfRefract(vec3 wo, vec3 wi, vec3 h, float eta) const {
    const float oh = dot(wo, h), ih = dot(wi, h);
    if (wo.z <= 0.0f || wi.z >= 0.0f || oh <= 0.0f || ih >= 0.0f) return 0.0f;
    const float den = eta * oh + ih;
    if (fabsf(den) < 1e-6f) return 0.0f;
    return G1(wo) * D(h) * oh / wo.z * fabsf(ih) / (den * den);
  }
};

// Dielectric Fresnel, Schlick at the cosine on the low-IOR side; 1 past the
// critical angle. cosI > 0 is the incident cosine against the facet.
SDEV float fresnelDielectric(float cosI, float eta, float f0) {
  const float k = 1.0f - eta * eta * (1.0f - cosI * cosI);
  if (k <= 0.0f) return 1.0f;
  return schlick(eta < 1.0f ? cosI : sqrtf(k), f0);
}

struct LobeEval {
  vec3 f;     // BSDF value
  float pdf;  // solid-angle density of the closure's own sampler
};
SDEV LobeEval noEval() { return LobeEval{v3(0.0f), 0.0f}; }

// Plastic: coat value, coupled diffuse value, and the 0.5 / 0.5 mixture of
// the coat's visible-normal density and the cosine density.
SDEV LobeEval evalCoated(const Closure& c, vec3 lo, vec3 li) {
  const Ggx g(c.alpha);
  const vec3 h = normalize(lo + li);
  const float F = schlick(dot(lo, h), c.f0);
  const float D = g.D(h);
  const float spec = F * D * g.G(lo, li) / (4.0f * lo.z * li.z);
  const vec3 diff = c.tint * (kInvPi * (1.0f - schlick(lo.z, c.f0)) * (1.0f - schlick(li.z, c.f0)));
  const float pSpec = (lo.z <= 0.0f || li.z <= 0.0f) ? 0.0f : g.G1(lo) * D / (4.0f * lo.z);
  return LobeEval{v3(spec) + diff, 0.5f * (pSpec + li.z * kInvPi)};
}

// f and pdf of a non-delta closure for local directions lo (outgoing, toward
// the previous vertex) and li (toward the light), together: the half
// vector, D and the Smith terms are shared by the value and the density.
SDEV LobeEval evalClosure(const Closure& c, vec3 lo, vec3 li) {
  if (lo.z <= 0.0f) return noEval();
  if (li.z <= 0.0f) {
    // only the rough dielectric reaches into the lower hemisphere
    if (c.lobe != LK_GLASS || li.z == 0.0f) return noEval();
    const Ggx g(c.alpha);
    vec3 h = normalize(c.eta * lo + li);
    if (h.z < 0.0f) h = -h;
    const float oh = dot(lo, h), ih = dot(li, h);
    if (oh <= 0.0f || ih >= 0.0f) return noEval();
    const float den = c.eta * oh + ih;
    if (fabsf(den) < 1e-6f) return noEval();
    const float F = fresnelDielectric(oh, c.eta, c.f0);
    const float D = g.D(h);
    // Walter et al. 2007 eq. 21 with the eta^2 radiance factor
    const float ft = c.eta * c.eta * (1.0f - F) * D * g.G(lo, li) * fabsf(oh * ih) /
                     (lo.z * fabsf(li.z) * den * den);
    const float pt = g.G1(lo) * D * oh / lo.z * fabsf(ih) / (den * den);
    return LobeEval{v3(ft), (1.0f - F) * pt};
  }
  switch (c.lobe) {
    case LK_DIFFUSE:
      return LobeEval{c.tint * kInvPi, li.z * kInvPi};
    case LK_COATED:
      return evalCoated(c, lo, li);
    case LK_CONDUCTOR:
    case LK_GLASS: {
      const Ggx g(c.alpha);
      const vec3 h = normalize(lo + li);
      const float D = g.D(h);
      const float refl = D * g.G(lo, li) / (4.0f * lo.z * li.z);
      const float pr = g.G1(lo) * D / (4.0f * lo.z);
      if (c.lobe == LK_CONDUCTOR) return LobeEval{schlick3(dot(lo, h), c.tint) * refl, pr};
      const float F = fresnelDielectric(dot(lo, h), c.eta, c.f0);
      return LobeEval{v3(F * D * g.G(lo, li) / (4.0f * lo.z * li.z)), F * pr};
    }
    default:
      return noEval();
  }
}

// The same from world directions wo, wi at shading normal n; the lambert
// lobe needs only the two cosines, every other lobe the shading frame.
SDEV LobeEval evalClosureAt(const Closure& c, vec3 n, vec3 wo, vec3 wi) {
  if (c.lobe == LK_DIFFUSE) {
    if (dot(wo, n) <= 0.0f) return noEval();
    const float ci = dot(wi, n);
    return ci <= 0.0f ? noEval() : LobeEval{c.tint * kInvPi, ci * kInvPi};
  }
  const Frame fr(n);
  return evalClosure(c, fr.toLocal(wo), fr.toLocal(wi));
}

struct BsdfSample {
  vec3 wi;
  vec3 weight;  // f |cos| / pdf, or the delta lobe's weight
  float pdf;    // 0 for delta lobes
  bool isDelta;
  bool valid;
};

SDEV BsdfSample noSample() {
  BsdfSample s;
  s.wi = v3(0.0f);
  s.weight = v3(0.0f);
  s.pdf = 0.0f;
  s.isDelta = false;
  s.valid = false;
  return s;
}

// Smooth interface: Fresnel-weighted choice (one draw) between mirror
// reflection and refraction, or reflection without a draw past the critical
// angle.
SDEV BsdfSample sampleInterface(const Closure& c, vec3 rayDir, vec3 n, Pcg& rng) {
  BsdfSample s = noSample();
  s.isDelta = true;
  s.valid = true;
  s.weight = v3(1.0f);
  vec3 t;
  if (!refract(rayDir, n, c.eta, t)) {
    s.wi = normalize(reflect(rayDir, n));
    return s;
  }
  const float R = schlick(c.eta < 1.0f ? -dot(rayDir, n) : -dot(t, n), c.f0);
  s.wi = rng.rnd() < R ? normalize(reflect(rayDir, n)) : normalize(t);
  return s;
}

// Rough dielectric: visible normal first (two draws), then the Fresnel
// reflect / refract choice (one draw, none on facet-level TIR); the weight
// is G / G1, times eta^2 for transmission.
SDEV BsdfSample sampleGlass(const Closure& c, vec3 lo, Pcg& rng, vec3& li) {
  BsdfSample s = noSample();
  const Ggx g(c.alpha);
  const vec3 h = g.sampleVisible(lo, rng.rnd2());
  const float oh = dot(lo, h);
  if (oh <= 0.0f) return s;
  vec3 t;
  if (!refract(-lo, h, c.eta, t)) {
    li = reflect(-lo, h);
    if (li.z <= 0.0f) return s;
    s.pdf = g.pdfReflect(lo, li);
  } else {
    const float R = schlick(c.eta < 1.0f ? oh : -dot(t, h), c.f0);
    if (rng.rnd() < R) {
      li = reflect(-lo, h);
      if (li.z <= 0.0f) return s;
      s.pdf = R * g.pdfReflect(lo, li);
    } else {
      li = normalize(t);
      if (li.z >= 0.0f) return s;
      s.pdf = (1.0f - R) * g.pdfRefract(lo, li, h, c.eta);
    }
  }
  float w = g.G(lo, li) / g.G1(lo);
  if (li.z < 0.0f) w *= c.eta * c.eta;
  s.weight = v3(w);
  s.valid = s.pdf > 0.0f;
  return s;
}

// Samples the closure for an incoming ray rayDir at shading normal n (on the
// incoming side). Draws: diffuse 2, mirror 0, interface 1 (0 on TIR),
// conductor 2, glass 2 + (0 or 1), coated exactly 3.
SDEV BsdfSample sampleClosure(const Closure& c, vec3 rayDir, vec3 n, Pcg& rng) {
  if (c.lobe == LK_INTERFACE) return sampleInterface(c, rayDir, n, rng);
  if (c.lobe == LK_MIRROR) {
    BsdfSample s = noSample();
    s.wi = normalize(reflect(rayDir, n));
    const float cw = dot(s.wi, n);
    if (cw <= 0.0f) return s;
    s.weight = schlick3(cw, c.tint);
    s.isDelta = true;
    s.valid = true;
    return s;
  }
  BsdfSample s = noSample();
  if (c.lobe == LK_NONE) return s;
  const Frame fr(n);
  if (c.lobe == LK_DIFFUSE) {
    const vec3 li = cosineHemisphere(rng.rnd2());
    s.wi = fr.toWorld(li);
    s.pdf = fmaxf(li.z, 1e-8f) * kInvPi;
    s.weight = c.tint;
    s.valid = li.z > 0.0f;
    return s;
  }
  const vec3 lo = fr.toLocal(-rayDir);
  if (lo.z <= 0.0f) return s;
  vec3 li;
  if (c.lobe == LK_CONDUCTOR) {
    const Ggx g(c.alpha);
    const vec3 h = g.sampleVisible(lo, rng.rnd2());
    li = reflect(-lo, h);
    if (li.z <= 0.0f) return s;
    s.weight = schlick3(dot(lo, h), c.tint) * (g.G(lo, li) / g.G1(lo));
    s.pdf = g.pdfReflect(lo, li);
    s.valid = s.pdf > 0.0f;
  } else if (c.lobe == LK_GLASS) {
    s = sampleGlass(c, lo, rng, li);
  } else {  // LK_COATED: lobe choice first, then that lobe's two draws
    const bool coat = rng.rnd() < 0.5f;
    const vec2 u = rng.rnd2();
    li = coat ? reflect(-lo, Ggx(c.alpha).sampleVisible(lo, u)) : cosineHemisphere(u);
    if (li.z <= 0.0f) return s;
    const LobeEval e = evalCoated(c, lo, li);
    if (e.pdf <= 0.0f) return s;
    s.pdf = e.pdf;
    s.weight = e.f * (li.z / e.pdf);
    s.valid = true;
  }
  if (s.valid) s.wi = fr.toWorld(li);
  return s;
}

// ===========================================================================
// emitters: area lights (rect, disk, mesh) through one area-to-solid-angle
// conversion, sphere lights by cone sampling, soft point and distant lights
// ===========================================================================

// Converts a point p (normal nL) sampled uniformly on an emitter of area
// `area` into a direction sample from x.
SDEV bool areaToSolidAngle(vec3 x, vec3 p, vec3 nL, float area, bool twoSided, LightSample& s) {
  const vec3 to = p - x;
  const float d2 = len2(to);
  if (d2 < 1e-10f) return false;
  const float dist = sqrtf(d2);
  const vec3 wi = to / dist;
  float cosL = -dot(wi, nL);
  if (!twoSided && cosL <= 0.0f) return false;  // the back of a one-sided emitter
  cosL = fabsf(cosL);
  if (cosL < 1e-6f) return false;
  s.wi = wi;
  s.dist = dist;
  s.pdf = d2 / (cosL * area);
  s.valid = true;
  return true;
}

SDEV vec3 emitterRadiance(const LaunchParams& P, const DevLight& lt, float u, float v) {
  if (lt.texId < 0) return ld3(lt.L);
  const float4 c = evalTexture(P, lt.texId, u, v);
  return ld3(lt.L) * vec3{c.x, c.y, c.z};
}

// Mesh emitter: triangle by world area (CDF bisection), then a uniform point
// by the square-root barycentric warp; the normal is the world winding
// normal, signed to the emitting side.
SDEV bool sampleMeshEmitter(const LaunchParams& P, const DevLight& lt, vec3 x, Pcg& rng, LightSample& s) {
  const int tri = cdfBinBisect(P.lightCdf + lt.mCdfOffset, lt.mNumTris + 1, rng.rnd());
  const uint32_t* ix = P.indices + lt.mIdxOffset + 3 * tri;
  const uint32_t i0 = ix[0], i1 = ix[1], i2 = ix[2];
  const float* wp = P.lightPos + 3 * lt.mPosOffset;
  const vec3 p0 = ld3i(wp, i0), p1 = ld3i(wp, i1), p2 = ld3i(wp, i2);
  const vec2 r = rng.rnd2();
  const float su = sqrtf(r.x);
  const float b1 = su * (1.0f - r.y), b2 = su * r.y, b0 = 1.0f - b1 - b2;
  const vec3 p = p0 * b0 + p1 * b1 + p2 * b2;
  const vec3 ng = cross(p1 - p0, p2 - p0);
  const float ngL2 = len2(ng);
  if (ngL2 < 1e-20f) return false;
  if (!areaToSolidAngle(x, p, ng * (lt.mNgSign / sqrtf(ngL2)), lt.area, lt.twoSided != 0, s)) return false;
  float tu = b1, tv = b2;  // barycentric uv without a uv channel
  if (lt.texId >= 0 && lt.mUvOffset >= 0) {
    const float* uv = P.uvs + 2 * lt.mUvOffset;
    tu = uv[2 * i0] * b0 + uv[2 * i1] * b1 + uv[2 * i2] * b2;
    tv = uv[2 * i0 + 1] * b0 + uv[2 * i1 + 1] * b1 + uv[2 * i2 + 1] * b2;
  }
  s.Li = emitterRadiance(P, lt, tu, tv);
  return true;
}

// Draws: rect 2, disk 2, sphere 2, mesh 3, point 2, distant 0.
SDEV LightSample sampleLight(const LaunchParams& P, const DevLight& lt, vec3 x, Pcg& rng) {
  LightSample s = noLight(0.0f);
  const vec3 c = ld3(lt.p);
  switch (lt.kind) {
    case LT_RECT: {
      const float a = 2.0f * rng.rnd() - 1.0f;
      const float b = 2.0f * rng.rnd() - 1.0f;
      if (areaToSolidAngle(x, c + a * ld3(lt.u) + b * ld3(lt.v), ld3(lt.n), lt.area, lt.twoSided != 0, s))
        s.Li = emitterRadiance(P, lt, (a + 1.0f) * 0.5f, (b + 1.0f) * 0.5f);
      return s;
    }
    case LT_DISK: {
      const vec2 d = concentricDisk(rng.rnd2());
      if (areaToSolidAngle(x, c + d.x * ld3(lt.u) + d.y * ld3(lt.v), ld3(lt.n), lt.area, lt.twoSided != 0, s))
        s.Li = emitterRadiance(P, lt, (atan2f(d.y, d.x) + kPi) / (2.0f * kPi), sqrtf(d.x * d.x + d.y * d.y));
      return s;
    }
    case LT_MESH:
      if (!sampleMeshEmitter(P, lt, x, rng, s)) s = noLight(0.0f);
      return s;
    case LT_SPHERE: {
      // uniform over the cone of directions the sphere subtends
      const vec3 to = c - x;
      const float d2 = len2(to), r2 = lt.radius * lt.radius;
      if (d2 <= r2 * 1.0001f) return s;  // inside the emitter
      const float dist = sqrtf(d2);
      const float cosMax = sqrtf(fmaxf(0.0f, 1.0f - r2 / d2));
      const vec2 u = rng.rnd2();
      const float cosT = 1.0f - u.x * (1.0f - cosMax);
      const float sinT = sqrtf(fmaxf(0.0f, 1.0f - cosT * cosT));
      const float phi = 2.0f * kPi * u.y;
      s.wi = Frame(to / dist).toWorld(vec3{sinT * cosf(phi), sinT * sinf(phi), cosT});
      // the near intersection with the sphere along wi
      const float b = dot(s.wi, to);
      s.dist = b - sqrtf(fmaxf(0.0f, b * b - d2 + r2));
      s.Li = ld3(lt.L);
      s.pdf = 1.0f / (2.0f * kPi * (1.0f - cosMax));
      s.valid = true;
      return s;
    }
    case LT_POINT: {
      // a point on a disk of the light's radius facing the shading point;
      // intensity falls off as 1 / d^2 from that point
      s.isDelta = true;
      const vec3 to = c - x;
      const float d2c = len2(to);
      if (d2c < 1e-10f) return s;
      const Frame fr(to / sqrtf(d2c));
      const vec2 d = concentricDisk(rng.rnd2()) * lt.radius;
      const vec3 q = (c + d.x * fr.t + d.y * fr.b) - x;
      const float d2 = len2(q), dist = sqrtf(d2);
      s.wi = q / dist;
      s.dist = fmaxf(dist - 1e-4f, 1e-4f);
      s.Li = ld3(lt.L) / d2;
      s.pdf = 1.0f;
      s.valid = true;
      return s;
    }
    case LT_DISTANT:
      s.isDelta = true;
      s.wi = -ld3(lt.dir);
      s.dist = 1e16f;
      s.Li = ld3(lt.L);
      s.pdf = 1.0f;
      s.valid = true;
      return s;
  }
  return s;
}

// The density sampleLight gives to a BSDF-sampled hit at hitP (normal nHit
// on mesh emitters), for the MIS weight of that hit.
SDEV float lightPdfSolidAngle(const DevLight& lt, vec3 from, vec3 hitP, vec3 nHit) {
  if (lt.kind == LT_SPHERE) {
    const vec3 to = ld3(lt.p) - from;
    const float d2 = len2(to), r2 = lt.radius * lt.radius;
    if (d2 <= r2 * 1.0001f) return 0.0f;
    return 1.0f / (2.0f * kPi * (1.0f - sqrtf(fmaxf(0.0f, 1.0f - r2 / d2))));
  }
  if (lt.kind != LT_RECT && lt.kind != LT_DISK && lt.kind != LT_MESH) return 0.0f;
  const vec3 to = hitP - from;
  const float d2 = len2(to);
  if (d2 < 1e-10f) return 0.0f;
  const float cosL = fabsf(dot(to / sqrtf(d2), lt.kind == LT_MESH ? nHit : ld3(lt.n)));
  if (cosL < 1e-6f) return 0.0f;
  return d2 / (cosL * lt.area);
}

// ===========================================================================
// intersectors. FP contraction is off in these functions so that the CPU
// oracle, compiled without contraction, reproduces every t bit.
// ===========================================================================

// Roots of the canonical quadrics in object space (unit sphere; rect and
// disk in y = 0; unit cylinder |y| <= 1; parabola y = (x^2 + z^2) / 2 over
// the unit disk) inside (tmin, tmax), ascending. Products are written out:
// the pragma governs only expressions parsed here, not inlined helpers.
SDEV int quadricRoots(int kind, vec3 o, vec3 d, float tmin, float tmax, float t[2]) {
#pragma clang fp contract(off)
  auto keep = [&](int n, float r) {
    if (r > tmin && r < tmax) t[n++] = r;
    return n;
  };
  if (kind == GK_RECT || kind == GK_DISK) {
    if (d.y == 0.0f) return 0;
    const float tt = -o.y / d.y;
    if (tt <= tmin || tt >= tmax) return 0;
    const float x = o.x + tt * d.x, z = o.z + tt * d.z;
    const bool inside = kind == GK_RECT ? (fabsf(x) <= 1.0f && fabsf(z) <= 1.0f) : sqrtf(x * x + z * z) <= 1.0f;
    if (!inside) return 0;
    t[0] = tt;
    return 1;
  }
  if (kind == GK_SPHERE) {
    const float a = d.x * d.x + d.y * d.y + d.z * d.z;
    const float b = o.x * d.x + o.y * d.y + o.z * d.z;
    const float c = (o.x * o.x + o.y * o.y + o.z * o.z) - 1.0f;
    const float disc = b * b - a * c;
    if (disc <= 0.0f) return 0;
    const float s = sqrtf(disc);
    return keep(keep(0, (-b - s) / a), (-b + s) / a);
  }
  if (kind == GK_CYLINDER) {
    const float a = d.x * d.x + d.z * d.z;
    if (a == 0.0f) return 0;
    const float b = o.x * d.x + o.z * d.z;
    const float c = o.x * o.x + o.z * o.z - 1.0f;
    const float disc = b * b - a * c;
    if (disc <= 0.0f) return 0;
    const float s = sqrtf(disc);
    const float r0 = (-b - s) / a, r1 = (-b + s) / a;
    int n = 0;
    if (fabsf(o.y + r0 * d.y) <= 1.0f) n = keep(n, r0);
    if (fabsf(o.y + r1 * d.y) <= 1.0f) n = keep(n, r1);
    return n;
  }
  if (kind == GK_PARABOLA) {
    const float A = 0.5f * (d.x * d.x + d.z * d.z);
    const float B = o.x * d.x + o.z * d.z - d.y;
    const float C = 0.5f * (o.x * o.x + o.z * o.z) - o.y;
    float r[2];
    int m;
    if (fabsf(A) < 1e-12f) {
      if (B == 0.0f) return 0;
      r[0] = -C / B;
      m = 1;
    } else {
      const float disc = B * B - 4.0f * A * C;
      if (disc <= 0.0f) return 0;
      const float s = sqrtf(disc);
      r[0] = (-B - s) / (2.0f * A);
      r[1] = (-B + s) / (2.0f * A);
      m = 2;
    }
    int n = 0;
    for (int i = 0; i < m; i++) {
      const float px = o.x + r[i] * d.x, pz = o.z + r[i] * d.z;
      if (px * px + pz * pz <= 1.0f) n = keep(n, r[i]);
    }
    return n;
  }
  return 0;
}

SDEV float azimuthU(float x, float z) { return (atan2f(z, x) + kPi) / (2.0f * kPi); }

// Object-space normal of the canonical front side and the surface uv at a
// quadric hit point p, one branch per kind (each with its own azimuth: a
// shared or selected one changes the traversal loop's register allocation
// for the worse).
SDEV void quadricSurface(int kind, vec3 p, vec3& n, float& u, float& v) {
  switch (kind) {
    case GK_RECT:
      n = vec3{0.0f, 1.0f, 0.0f};
      u = (p.x + 1.0f) * 0.5f;
      v = (p.z + 1.0f) * 0.5f;
      break;
    case GK_SPHERE:
      n = normalize(p);
      u = azimuthU(n.x, n.z);
      v = (asinf(clampf(n.y, -1.0f, 1.0f)) + kPi / 2.0f) / kPi;
      break;
    case GK_DISK:
      n = vec3{0.0f, 1.0f, 0.0f};
      u = azimuthU(p.x, p.z);
      v = sqrtf(p.x * p.x + p.z * p.z);
      break;
    case GK_CYLINDER:
      n = normalize(vec3{p.x, 0.0f, p.z});
      u = azimuthU(p.x, p.z);
      v = (p.y + 1.0f) * 0.5f;
      break;
    default:  // parabola
      n = normalize(vec3{p.x, -1.0f, p.z});
      u = azimuthU(p.x, p.z);
      v = sqrtf(p.x * p.x + p.z * p.z);
  }
}

// Watertight ray / triangle test (Woop, Benthin, Wald, JCGT 2013): the ray
// is sheared so it runs along +z; per-ray constants below. The shear
// factors are the reciprocal 1 / d_z times d_x and d_y (one division where
// the paper divides three times); every triangle a ray meets is sheared by
// the same factors, so shared edges still evaluate alike.
struct RayTri {
  int kx, ky, kz;
  float Sx, Sy, Sz;
};

SDEV RayTri rayTriSetup(vec3 d) {
#pragma clang fp contract(off)
  RayTri r;
  const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
  r.kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
  r.kx = r.kz == 2 ? 0 : r.kz + 1;
  r.ky = r.kx == 2 ? 0 : r.kx + 1;
  const float dz = comp(d, r.kz);
  if (dz < 0.0f) {
    const int k = r.kx;
    r.kx = r.ky;
    r.ky = k;
  }
  r.Sz = 1.0f / dz;
  r.Sx = comp(d, r.kx) * r.Sz;
  r.Sy = comp(d, r.ky) * r.Sz;
  return r;
}

// Edge function of the sheared 2D vertices; recomputed in double when the
// float result is exactly zero (an edge or vertex hit), as the paper does.
SDEV float edgeFn(float ax, float ay, float bx, float by) {
#pragma clang fp contract(off)
  const float e = ax * by - ay * bx;
  return e;
}

// True with (t, b1, b2) when the triangle is hit in (tmin, tmax); b1, b2 are
// the barycentric weights of v1 and v2.
SDEV bool triHit(const RayTri& R, vec3 o, vec3 v0, vec3 v1, vec3 v2, float tmin, float tmax, float& t, float& b1,
                 float& b2) {
#pragma clang fp contract(off)
  const vec3 A = v0 - o, B = v1 - o, C = v2 - o;
  const float Az0 = comp(A, R.kz), Bz0 = comp(B, R.kz), Cz0 = comp(C, R.kz);
  const float Ax = comp(A, R.kx) - R.Sx * Az0, Ay = comp(A, R.ky) - R.Sy * Az0;
  const float Bx = comp(B, R.kx) - R.Sx * Bz0, By = comp(B, R.ky) - R.Sy * Bz0;
  const float Cx = comp(C, R.kx) - R.Sx * Cz0, Cy = comp(C, R.ky) - R.Sy * Cz0;
  float U = edgeFn(Cx, Cy, Bx, By);
  float V = edgeFn(Ax, Ay, Cx, Cy);
  float W = edgeFn(Bx, By, Ax, Ay);
  if (U == 0.0f || V == 0.0f || W == 0.0f) {
    U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
    V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
    W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
  }
  if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
  const float det = U + V + W;
  if (det == 0.0f) return false;
  const float T = U * (R.Sz * Az0) + V * (R.Sz * Bz0) + W * (R.Sz * Cz0);
  const float inv = 1.0f / det;
  const float tt = T * inv;
  if (!(tt > tmin && tt < tmax)) return false;
  t = tt;
  b1 = V * inv;
  b2 = W * inv;
  return true;
}

// Row-major 3x4 transforms.
SDEV vec3 xfPoint(const float* m, vec3 p) {
#pragma clang fp contract(off)
  return vec3{m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3], m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7],
              m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11]};
}
SDEV vec3 xfVector(const float* m, vec3 v) {
#pragma clang fp contract(off)
  return vec3{m[0] * v.x + m[1] * v.y + m[2] * v.z, m[4] * v.x + m[5] * v.y + m[6] * v.z,
              m[8] * v.x + m[9] * v.y + m[10] * v.z};
}
// The inverse transpose applied to an object-space normal, given the inverse.
SDEV vec3 xfNormal(const float* inv, vec3 n) {
  return vec3{inv[0] * n.x + inv[4] * n.y + inv[8] * n.z, inv[1] * n.x + inv[5] * n.y + inv[9] * n.z,
              inv[2] * n.x + inv[6] * n.y + inv[10] * n.z};
}

#undef SDEV

}  // namespace sdk
