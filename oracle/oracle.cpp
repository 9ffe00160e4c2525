// CPU oracle for the sundog render path (TEST INFRASTRUCTURE ONLY; see
// oracle.h). Scalar C++17 + OpenMP. Every function cites the reference
// source it restates. Compiled with -ffp-contract=off so that the
// geometry functions reproduce the GPU's bits (the GPU disables contraction
// in the same functions).
//
// Parity unpinned (no check outside the two restatements fixes them):
// fbm's per-octave coordinate offset (11.7, 5.3, 7.9) and seed stride 101,
// the flame noise coordinates' y compression 0.32, and the flame mask's
// smoothstep edges (0.62 in d', 0.12 in h). The flame profile, heat and
// colour ramp constants are pinned against the reference report's closed
// forms, the hash, value noise and fbm gain/lacunarity against published
// definitions.
#include "oracle.h"

#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

constexpr float PI = 3.14159265358979323846f;
constexpr float INV_PI = 0.31830988618379067154f;
constexpr uint16_t NO_MAT = 0xFFFF;

// ---------------------------------------------------------------------------
// vector math (device/math.cuh:17-154)
// ---------------------------------------------------------------------------
struct V {
  float x, y, z;
};
inline V mk(float x, float y, float z) { return V{x, y, z}; }
inline V spl(float s) { return V{s, s, s}; }
inline V operator+(V a, V b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V operator-(V a, V b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V operator-(V a) { return mk(-a.x, -a.y, -a.z); }
inline V operator*(V a, V b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
inline V operator*(V a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
inline V operator*(float s, V a) { return a * s; }
inline V operator/(V a, float s) { return a * (1.0f / s); }
inline float dot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V cross(V a, V b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
inline float len2(V a) { return dot(a, a); }
inline float len(V a) { return sqrtf(dot(a, a)); }
inline V norm(V a) { return a * (1.0f / sqrtf(dot(a, a))); }
inline float clampf(float x, float lo, float hi) { return fminf(hi, fmaxf(lo, x)); }
inline V vminv(V a, V b) { return mk(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)); }
inline float maxComp(V v) { return fmaxf(v.x, fmaxf(v.y, v.z)); }
inline V expv(V a) { return mk(expf(a.x), expf(a.y), expf(a.z)); }
inline V reflect(V v, V n) { return v - 2.0f * dot(v, n) * n; }
inline float at(V v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
inline V ld(const float* p) { return mk(p[0], p[1], p[2]); }

inline bool refract(V v, V n, float eta, V& out) {
  float cosi = -dot(v, n);
  float k = 1.0f - eta * eta * (1.0f - cosi * cosi);
  if (k < 0.0f) return false;
  out = eta * v + (eta * cosi - sqrtf(k)) * n;
  return true;
}
inline float schlick(float c, float f0) {
  float m = clampf(1.0f - c, 0.0f, 1.0f);
  float m2 = m * m;
  return f0 + (1.0f - f0) * m2 * m2 * m;
}
inline V schlick3(float c, V f0) {
  float m = clampf(1.0f - c, 0.0f, 1.0f);
  float m2 = m * m;
  return f0 + (spl(1.0f) - f0) * (m2 * m2 * m);
}
struct Onb {
  V t, b, n;
  explicit Onb(V nn) : n(nn) {
    float s = copysignf(1.0f, nn.z);
    float a = -1.0f / (s + nn.z);
    float c = nn.x * nn.y * a;
    t = mk(1.0f + s * nn.x * nn.x * a, s * c, -s * nn.x);
    b = mk(c, s + nn.y * nn.y * a, -nn.y);
  }
  V toWorld(V v) const { return v.x * t + v.y * b + v.z * n; }
  V toLocal(V v) const { return mk(dot(v, t), dot(v, b), dot(v, n)); }
};
inline V offsetRay(V p, V n) {
  float s = 1e-4f * (1.0f + fmaxf(fabsf(p.x), fmaxf(fabsf(p.y), fabsf(p.z))));
  return p + n * s;
}
inline V sanitize(V c) {
  if (!std::isfinite(c.x) || !std::isfinite(c.y) || !std::isfinite(c.z)) return spl(0.0f);
  return c;
}

// ---------------------------------------------------------------------------
// PCG32 and warps (device/rng.cuh:10-61)
// ---------------------------------------------------------------------------
struct Rng {
  uint64_t state = 0, inc = 1;
  uint32_t draws = 0;
  uint32_t next() {
    uint64_t old = state;
    state = old * 6364136223846793005ULL + inc;
    uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    draws++;
    return (xorshifted >> rot) | (xorshifted << ((32u - rot) & 31u));
  }
  static Rng init(uint64_t seed, uint64_t seq) {
    Rng r;
    r.state = 0u;
    r.inc = (seq << 1u) | 1u;
    r.next();
    r.state += seed;
    r.next();
    r.draws = 0;
    return r;
  }
  float rnd() { return (float)(next() >> 8) * 0x1p-24f; }
};

V uniformOnSphere(float ux, float uy) {
  float z = 1.0f - 2.0f * ux;
  float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
  float phi = 2.0f * PI * uy;
  return mk(r * cosf(phi), r * sinf(phi), z);
}
void concentricDisk(float ux, float uy, float& dx, float& dy) {
  float ox = 2.0f * ux - 1.0f, oy = 2.0f * uy - 1.0f;
  if (ox == 0.0f && oy == 0.0f) {
    dx = dy = 0.0f;
    return;
  }
  float r, theta;
  if (fabsf(ox) > fabsf(oy)) {
    r = ox;
    theta = (PI / 4.0f) * (oy / ox);
  } else {
    r = oy;
    theta = (PI / 2.0f) - (PI / 4.0f) * (ox / oy);
  }
  dx = r * cosf(theta);
  dy = r * sinf(theta);
}
V cosineHemisphere(float ux, float uy) {
  float dx, dy;
  concentricDisk(ux, uy, dx, dy);
  return mk(dx, dy, sqrtf(fmaxf(0.0f, 1.0f - dx * dx - dy * dy)));
}

// ---------------------------------------------------------------------------
// scene wrapper
// ---------------------------------------------------------------------------
struct Node2 {
  float lo[3], hi[3];
  int left = -1, right = -1;  // inner when left >= 0
  int first = 0, count = 0;
};

struct Bvh2 {
  std::vector<Node2> nodes;
  std::vector<int> prims;
};

struct MeshLightData {
  std::vector<V> wp;
  std::vector<float> cdf;
  float area = 0.0f;
  float ngSign = 1.0f;
  int mesh = -1;
};

struct Scene {
  const sundog_flat_scene* f = nullptr;
  std::vector<Bvh2> blas;  // per mesh
  Bvh2 tlas;
  std::vector<float> marg, cond;
  bool envImportance = false;
  std::vector<MeshLightData> meshLight;  // per light (empty for non-mesh)
  std::vector<float> lightArea;          // effective area per light
};

// ---------------------------------------------------------------------------
// textures (device/texture_eval.cuh:11-46); bilinear image lookup is the
// textbook filter with wrap (textures) or clamp-v (environment) addressing.
// ---------------------------------------------------------------------------
int wrap(int i, int n) {
  int r = i % n;
  return r < 0 ? r + n : r;
}
void bilinear(const float* img, int w, int h, float u, float v, bool wrapV, float out[4]) {
  u = u - floorf(u);
  if (wrapV) v = v - floorf(v);
  float x = u * (float)w - 0.5f, y = v * (float)h - 0.5f;
  x = fminf(fmaxf(x, -1.0f), (float)w);
  y = fminf(fmaxf(y, -1.0f), (float)h);
  float fx = floorf(x), fy = floorf(y);
  float ax = x - fx, ay = y - fy;
  int x0 = wrap((int)fx, w), x1 = wrap((int)fx + 1, w);
  int y0, y1;
  if (wrapV) {
    y0 = wrap((int)fy, h);
    y1 = wrap((int)fy + 1, h);
  } else {
    y0 = std::min(std::max((int)fy, 0), h - 1);
    y1 = std::min(std::max((int)fy + 1, 0), h - 1);
  }
  const float* t00 = img + 4 * ((size_t)y0 * w + x0);
  const float* t10 = img + 4 * ((size_t)y0 * w + x1);
  const float* t01 = img + 4 * ((size_t)y1 * w + x0);
  const float* t11 = img + 4 * ((size_t)y1 * w + x1);
  float w00 = (1.0f - ax) * (1.0f - ay), w10 = ax * (1.0f - ay), w01 = (1.0f - ax) * ay, w11 = ax * ay;
  for (int c = 0; c < 4; c++) out[c] = t00[c] * w00 + t10[c] * w10 + t01[c] * w01 + t11[c] * w11;
}
void evalTexture(const sundog_flat_texture& t, float u, float v, float out[4]) {
  out[3] = 1.0f;
  const float* c = t.a;
  switch (t.kind) {
    case 0:
      break;
    case 2: {
      int iu = (int)floorf(u * t.scale[0]);
      int iv = (int)floorf(v * t.scale[1]);
      c = (((iu + iv) & 1) == 0) ? t.a : t.b;
      break;
    }
    case 3: {
      float fu = u * t.scale[0] - floorf(u * t.scale[0]);
      float fv = v * t.scale[1] - floorf(v * t.scale[1]);
      float w = t.width;
      bool line = fu < w || fv < w || fu > 1.0f - w || fv > 1.0f - w;
      c = line ? t.b : t.a;
      break;
    }
    case 1:
      bilinear(t.img, t.img_w, t.img_h, u, 1.0f - v, true, out);
      return;
    default: {
      out[0] = 1.0f;
      out[1] = 0.0f;
      out[2] = 1.0f;
      return;
    }
  }
  out[0] = c[0];
  out[1] = c[1];
  out[2] = c[2];
}
V materialAlbedo(const Scene& S, const sundog_flat_material& m, float u, float v) {
  if (m.tex_id < 0) return ld(m.color);
  float c[4];
  evalTexture(S.f->textures[m.tex_id], u, v, c);
  return mk(c[0], c[1], c[2]);
}

// ---------------------------------------------------------------------------
// noise (device/noise.cuh:13-79), octaves parameterised
// ---------------------------------------------------------------------------
float smoothstepf(float e0, float e1, float x) {
  float t = clampf((x - e0) / (e1 - e0), 0.0f, 1.0f);
  return t * t * (3.0f - 2.0f * t);
}
uint32_t hashU(uint32_t x) {
  x = x * 747796405u + 2891336453u;
  x = ((x >> ((x >> 28u) + 4u)) ^ x) * 277803737u;
  return (x >> 22u) ^ x;
}
float lattice(int x, int y, int z, uint32_t seed) {
  uint32_t h = hashU((uint32_t)x * 73856093u ^ (uint32_t)y * 19349663u ^ (uint32_t)z * 83492791u ^ seed);
  return (float)h * (1.0f / 4294967296.0f);
}
float valueNoise(V p, uint32_t seed) {
  float fx = floorf(p.x), fy = floorf(p.y), fz = floorf(p.z);
  int ix = (int)fx, iy = (int)fy, iz = (int)fz;
  float tx = p.x - fx, ty = p.y - fy, tz = p.z - fz;
  tx = tx * tx * (3.0f - 2.0f * tx);
  ty = ty * ty * (3.0f - 2.0f * ty);
  tz = tz * tz * (3.0f - 2.0f * tz);
  float c000 = lattice(ix, iy, iz, seed), c100 = lattice(ix + 1, iy, iz, seed);
  float c010 = lattice(ix, iy + 1, iz, seed), c110 = lattice(ix + 1, iy + 1, iz, seed);
  float c001 = lattice(ix, iy, iz + 1, seed), c101 = lattice(ix + 1, iy, iz + 1, seed);
  float c011 = lattice(ix, iy + 1, iz + 1, seed), c111 = lattice(ix + 1, iy + 1, iz + 1, seed);
  float x00 = c000 + (c100 - c000) * tx;
  float x10 = c010 + (c110 - c010) * tx;
  float x01 = c001 + (c101 - c001) * tx;
  float x11 = c011 + (c111 - c011) * tx;
  float y0 = x00 + (x10 - x00) * ty;
  float y1 = x01 + (x11 - x01) * ty;
  return y0 + (y1 - y0) * tz;
}
float fbmN(V p, uint32_t seed, int octaves) {
  float amp = 0.5f, sum = 0.0f;
  for (int i = 0; i < octaves; i++) {
    sum += amp * valueNoise(p, seed + (uint32_t)i * 101u);
    p = 2.03f * p + mk(11.7f, 5.3f, 7.9f);
    amp *= 0.5f;
  }
  return sum;
}
V waterNormal(V x, V ns, V wo, float amp, float freq, int octaves) {
  const uint32_t SEED = 1337u;
  const float e = 0.05f;
  V p = mk(x.x * freq, 0.0f, x.z * freq);
  float hx1 = fbmN(p + mk(e, 0, 0), SEED, octaves), hx0 = fbmN(p - mk(e, 0, 0), SEED, octaves);
  float hz1 = fbmN(p + mk(0, 0, e), SEED, octaves), hz0 = fbmN(p - mk(0, 0, e), SEED, octaves);
  float dhdx = (hx1 - hx0) / (2.0f * e);
  float dhdz = (hz1 - hz0) / (2.0f * e);
  float side = ns.y >= 0.0f ? 1.0f : -1.0f;
  V n = norm(mk(-amp * dhdx, 1.0f, -amp * dhdz)) * side;
  if (dot(n, wo) <= 1e-4f) return ns;
  return n;
}

// ---------------------------------------------------------------------------
// flames (device/volume.cuh:27-153)
// ---------------------------------------------------------------------------
void flameField(const sundog_flat_flame& fl, V p, float& sigma, V& emission) {
  sigma = 0.0f;
  emission = spl(0.0f);
  float h = (p.y - fl.base[1]) / fl.height;
  if (h <= 0.0f || h >= 1.0f) return;
  float rx = p.x - fl.base[0], rz = p.z - fl.base[2];
  float d = sqrtf(rx * rx + rz * rz) / fl.radius;
  float prof = 3.2f * sqrtf(h) * powf(1.0f - h, 1.55f);
  V q = mk(rx / fl.radius, (p.y - fl.base[1]) / fl.radius * 0.32f, rz / fl.radius);
  float n = fbmN(q * fl.noise_scale, fl.seed, fl.octaves);
  float dd = d / fmaxf(prof, 1e-3f) + (n - 0.45f) * 1.0f;
  float mask = smoothstepf(1.0f, 0.62f, dd) * smoothstepf(0.0f, 0.12f, h);
  if (mask <= 0.0f) return;
  sigma = fl.sigma * mask;
  float heat = clampf((0.88f - dd) * (1.25f - h), 0.0f, 1.0f);
  V col = mk(fminf(heat * 2.6f, 1.0f), fminf(heat * heat * 2.1f, 1.0f), fminf(heat * heat * heat * 1.7f, 1.0f));
  emission = col * (fl.intensity * mask);
}
bool slab(float o, float d, float lo, float hi, float& t0, float& t1) {
  if (fabsf(d) < 1e-8f) return !(o < lo || o > hi);
  float ta = (lo - o) / d, tb = (hi - o) / d;
  if (ta > tb) std::swap(ta, tb);
  t0 = fmaxf(t0, ta);
  t1 = fminf(t1, tb);
  return true;
}
bool clipFlameBounds(const sundog_flat_flame& fl, V o, V d, float& t0, float& t1) {
  if (!slab(o.y, d.y, fl.base[1], fl.base[1] + fl.height, t0, t1)) return false;
  float ox = o.x - fl.base[0], oz = o.z - fl.base[2];
  if (fl.bounds == 1) {
    float r = fl.radius;
    if (!slab(ox, d.x, -r, r, t0, t1)) return false;
    if (!slab(oz, d.z, -r, r, t0, t1)) return false;
    return t0 < t1;
  }
  float a = d.x * d.x + d.z * d.z;
  if (a < 1e-12f) {
    if (ox * ox + oz * oz > fl.radius * fl.radius) return false;
  } else {
    float b = ox * d.x + oz * d.z;
    float c = ox * ox + oz * oz - fl.radius * fl.radius;
    float disc = b * b - a * c;
    if (disc <= 0.0f) return false;
    float s = sqrtf(disc);
    t0 = fmaxf(t0, (-b - s) / a);
    t1 = fminf(t1, (-b + s) / a);
  }
  return t0 < t1;
}
int marchSteps(const sundog_flat_flame& fl, float length) {
  if (fl.step <= 0.0f) return 32;
  float n = ceilf(length / fl.step);
  return (int)fminf(fmaxf(n, 1.0f), 65536.0f);
}
V marchFlames(const Scene& S, V o, V d, float tEnd, Rng& rng, V& trans) {
  V Lv = spl(0.0f);
  for (int i = 0; i < S.f->num_flames; i++) {
    const sundog_flat_flame& fl = S.f->flames[i];
    float t0 = 1e-4f, t1 = tEnd;
    if (!clipFlameBounds(fl, o, d, t0, t1)) continue;
    int steps = marchSteps(fl, t1 - t0);
    float dt = (t1 - t0) / steps;
    float jitter = rng.rnd();
    float tLocal = 1.0f;
    V acc = spl(0.0f);
    for (int k = 0; k < steps; k++) {
      V p = o + (t0 + (k + jitter) * dt) * d;
      float sigma;
      V eps;
      flameField(fl, p, sigma, eps);
      acc = acc + (tLocal * dt) * eps;
      tLocal *= expf(-sigma * dt);
      if (tLocal < 1e-4f) break;
    }
    Lv = Lv + trans * acc;
    trans = trans * tLocal;
  }
  return Lv;
}
float flameTransmittance(const Scene& S, V o, V d, float tEnd, Rng& rng, int skip) {
  float tr = 1.0f;
  for (int i = 0; i < S.f->num_flames; i++) {
    if (i == skip) continue;
    const sundog_flat_flame& fl = S.f->flames[i];
    float t0 = 1e-4f, t1 = tEnd;
    if (!clipFlameBounds(fl, o, d, t0, t1)) continue;
    int steps = marchSteps(fl, t1 - t0);
    float dt = (t1 - t0) / steps;
    float jitter = rng.rnd();
    float tLocal = 1.0f;
    for (int k = 0; k < steps; k++) {
      V p = o + (t0 + (k + jitter) * dt) * d;
      float sigma;
      V eps;
      flameField(fl, p, sigma, eps);
      tLocal *= expf(-sigma * dt);
      if (tLocal < 1e-4f) break;
    }
    tr *= tLocal;
  }
  return tr;
}

// ---------------------------------------------------------------------------
// environment (src/env_light.cpp:27-59, device/env_light.cuh:17-116)
// ---------------------------------------------------------------------------
bool buildEnvCdf(const sundog_flat_scene* f, std::vector<float>& marg, std::vector<float>& cond) {
  int w = f->env_w, h = f->env_h;
  marg.assign(h + 1, 0.0f);
  cond.assign((size_t)h * (w + 1), 0.0f);
  double total = 0.0;
  for (int row = 0; row < h; row++) {
    float sinT = sinf(PI * (row + 0.5f) / h);
    float* cdf = &cond[(size_t)row * (w + 1)];
    double sum = 0.0;
    for (int col = 0; col < w; col++) {
      const float* p = f->env_rgba + 4 * ((size_t)row * w + col);
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
  bool ok = total > 0.0;
  if (ok) {
    for (int r = 1; r <= h; r++) marg[r] = (float)(marg[r] / total);
  } else {
    for (int r = 0; r <= h; r++) marg[r] = (float)r / h;
  }
  marg[h] = 1.0f;
  return ok;
}
void envDirToUv(const Scene& S, V dir, float& u, float& v) {
  u = (atan2f(dir.z, dir.x) - S.f->env_rotate + PI) * (0.5f * INV_PI);
  u -= floorf(u);
  v = acosf(clampf(dir.y, -1.0f, 1.0f)) * INV_PI;
}
V envEval(const Scene& S, V dir) {
  float u, v, c[4];
  envDirToUv(S, dir, u, v);
  bilinear(S.f->env_rgba, S.f->env_w, S.f->env_h, u, v, false, c);
  return mk(c[0], c[1], c[2]) * S.f->env_intensity;
}
int findInterval(const float* cdf, int n, float u) {
  int lo = 0, hi = n - 1;
  while (lo + 1 < hi) {
    int mid = (lo + hi) >> 1;
    if (cdf[mid] <= u) lo = mid;
    else hi = mid;
  }
  return lo;
}

struct LightSample {
  V wi = spl(0.0f);
  float dist = 0.0f;
  V Li = spl(0.0f);
  float pdf = 0.0f;
  bool isDelta = false;
  bool valid = false;
};

LightSample sampleEnvU(const Scene& S, float r0, float r1) {
  LightSample s;
  s.dist = 1e16f;
  if (!S.envImportance) {
    s.wi = uniformOnSphere(r0, r1);
    s.pdf = 0.25f * INV_PI;
    s.Li = envEval(S, s.wi);
    s.valid = true;
    return s;
  }
  int W = S.f->env_w, H = S.f->env_h;
  int row = findInterval(S.marg.data(), H + 1, r0);
  float dm = S.marg[row + 1] - S.marg[row];
  float dv = dm > 0.0f ? (r0 - S.marg[row]) / dm : 0.5f;
  float v = (row + dv) / H;
  const float* cdf = S.cond.data() + (size_t)row * (W + 1);
  int col = findInterval(cdf, W + 1, r1);
  float dc = cdf[col + 1] - cdf[col];
  float du = dc > 0.0f ? (r1 - cdf[col]) / dc : 0.5f;
  float u = (col + du) / W;
  float theta = v * PI;
  float sinT = sinf(theta);
  if (sinT < 1e-6f || dm <= 0.0f || dc <= 0.0f) return s;
  float phi = (2.0f * PI * u - PI) + S.f->env_rotate;
  s.wi = mk(sinT * cosf(phi), cosf(theta), sinT * sinf(phi));
  s.pdf = dm * dc * W * H / (2.0f * PI * PI * sinT);
  s.Li = envEval(S, s.wi);
  s.valid = s.pdf > 0.0f;
  return s;
}
// Same draw count in both modes: two rnd() (device/env_light.cuh:60).
LightSample sampleEnv(const Scene& S, Rng& rng) {
  float r0 = rng.rnd();
  float r1 = rng.rnd();
  return sampleEnvU(S, r0, r1);
}
float envPdfSolidAngle(const Scene& S, V dir) {
  if (!S.envImportance) return 0.25f * INV_PI;
  float sinT = sqrtf(fmaxf(0.0f, 1.0f - dir.y * dir.y));
  if (!(sinT >= 1e-6f)) return 0.0f;  // also rejects NaN directions
  float u, v;
  envDirToUv(S, dir, u, v);
  int W = S.f->env_w, H = S.f->env_h;
  int col = (int)fminf(fmaxf(u * W, 0.0f), (float)(W - 1));
  int row = (int)fminf(fmaxf(v * H, 0.0f), (float)(H - 1));
  float dm = S.marg[row + 1] - S.marg[row];
  const float* cdf = S.cond.data() + (size_t)row * (W + 1);
  float dc = cdf[col + 1] - cdf[col];
  return dm * dc * W * H / (2.0f * PI * PI * sinT);
}

// ---------------------------------------------------------------------------
// BSDFs (device/bsdf.cuh:27-424)
// ---------------------------------------------------------------------------
bool bsdfIsDelta(const sundog_flat_material& m) {
  if (m.kind == 4) return true;
  if (m.kind == 2 || m.kind == 1) return m.roughness < 1e-3f;
  return false;
}
float ggxD(V h, float alpha) {
  float c2 = h.z * h.z;
  float k = alpha * alpha * c2 + (1.0f - c2);
  return alpha * alpha / (PI * k * k);
}
float ggxLambda(V w, float alpha) {
  float c2 = w.z * w.z;
  if (c2 >= 1.0f) return 0.0f;
  float t2 = (1.0f - c2) / c2;
  return 0.5f * (-1.0f + sqrtf(1.0f + alpha * alpha * t2));
}
float ggxG1(V w, float alpha) { return 1.0f / (1.0f + ggxLambda(w, alpha)); }
float ggxG(V wo, V wi, float alpha) { return 1.0f / (1.0f + ggxLambda(wo, alpha) + ggxLambda(wi, alpha)); }
V ggxSampleVndf(V wo, float alpha, float ux, float uy) {
  V vh = norm(mk(alpha * wo.x, alpha * wo.y, wo.z));
  float l2 = vh.x * vh.x + vh.y * vh.y;
  V T1 = l2 > 0.0f ? mk(-vh.y, vh.x, 0.0f) * (1.0f / sqrtf(l2)) : mk(1.0f, 0.0f, 0.0f);
  V T2 = cross(vh, T1);
  float r = sqrtf(ux);
  float phi = 2.0f * PI * uy;
  float t1 = r * cosf(phi);
  float t2 = r * sinf(phi);
  float s = 0.5f * (1.0f + vh.z);
  t2 = (1.0f - s) * sqrtf(fmaxf(0.0f, 1.0f - t1 * t1)) + s * t2;
  V nh = t1 * T1 + t2 * T2 + sqrtf(fmaxf(0.0f, 1.0f - t1 * t1 - t2 * t2)) * vh;
  return norm(mk(alpha * nh.x, alpha * nh.y, fmaxf(1e-6f, nh.z)));
}
float ggxPdf(V wo, V wi, float alpha) {
  if (wo.z <= 0.0f || wi.z <= 0.0f) return 0.0f;
  V h = norm(wo + wi);
  return ggxG1(wo, alpha) * ggxD(h, alpha) / (4.0f * wo.z);
}
float fresnelDielectric(float cosI, float eta, float f0) {
  float k = 1.0f - eta * eta * (1.0f - cosI * cosI);
  if (k <= 0.0f) return 1.0f;
  return schlick(eta < 1.0f ? cosI : sqrtf(k), f0);
}
float ggxPdfTransmit(V lo, V li, V h, float eta, float alpha) {
  float loh = dot(lo, h), lih = dot(li, h);
  if (lo.z <= 0.0f || li.z >= 0.0f || loh <= 0.0f || lih >= 0.0f) return 0.0f;
  float denom = eta * loh + lih;
  if (fabsf(denom) < 1e-6f) return 0.0f;
  return ggxG1(lo, alpha) * ggxD(h, alpha) * loh / lo.z * fabsf(lih) / (denom * denom);
}
struct Plastic {
  float fSpec;
  V fDiff;
  float pSpec, pDiff;
};
Plastic plasticTerms(const sundog_flat_material& m, V albedo, V lo, V li) {
  Plastic t;
  float rough = fmaxf(m.roughness, 1e-3f);
  float alpha = rough * rough;
  float f0 = (m.ior - 1.0f) / (m.ior + 1.0f);
  f0 = f0 * f0;
  V h = norm(lo + li);
  float F = schlick(dot(lo, h), f0);
  t.fSpec = F * ggxD(h, alpha) * ggxG(lo, li, alpha) / (4.0f * lo.z * li.z);
  t.fDiff = albedo * (INV_PI * (1.0f - schlick(lo.z, f0)) * (1.0f - schlick(li.z, f0)));
  t.pSpec = ggxPdf(lo, li, alpha);
  t.pDiff = li.z * INV_PI;
  return t;
}
V bsdfEval(const sundog_flat_material& m, V albedo, V wo, V wi, V n, float eta) {
  float ci = dot(wi, n), co = dot(wo, n);
  if (co <= 0.0f) return spl(0.0f);
  bool roughGlass = m.kind == 2 && m.roughness >= 1e-3f;
  if (ci <= 0.0f) {
    if (!roughGlass || ci == 0.0f) return spl(0.0f);
    float alpha = m.roughness * m.roughness;
    Onb onb(n);
    V lo = onb.toLocal(wo), li = onb.toLocal(wi);
    V h = norm(eta * lo + li);
    if (h.z < 0.0f) h = -h;
    float loh = dot(lo, h), lih = dot(li, h);
    if (loh <= 0.0f || lih >= 0.0f) return spl(0.0f);
    float denom = eta * loh + lih;
    if (fabsf(denom) < 1e-6f) return spl(0.0f);
    float f0 = (1.0f - eta) / (1.0f + eta);
    f0 = f0 * f0;
    float F = fresnelDielectric(loh, eta, f0);
    float ft = eta * eta * (1.0f - F) * ggxD(h, alpha) * ggxG(lo, li, alpha) * fabsf(loh * lih) /
               (lo.z * fabsf(li.z) * denom * denom);
    return spl(ft);
  }
  if (m.kind == 0) return albedo * INV_PI;
  if (m.kind == 5) {
    Onb onb(n);
    Plastic t = plasticTerms(m, albedo, onb.toLocal(wo), onb.toLocal(wi));
    return spl(t.fSpec) + t.fDiff;
  }
  if (m.kind == 1 && m.roughness >= 1e-3f) {
    float alpha = m.roughness * m.roughness;
    Onb onb(n);
    V lo = onb.toLocal(wo), li = onb.toLocal(wi);
    V h = norm(lo + li);
    V F = schlick3(dot(lo, h), albedo);
    return F * (ggxD(h, alpha) * ggxG(lo, li, alpha) / (4.0f * lo.z * li.z));
  }
  if (roughGlass) {
    float alpha = m.roughness * m.roughness;
    Onb onb(n);
    V lo = onb.toLocal(wo), li = onb.toLocal(wi);
    V h = norm(lo + li);
    float f0 = (1.0f - eta) / (1.0f + eta);
    f0 = f0 * f0;
    float F = fresnelDielectric(dot(lo, h), eta, f0);
    return spl(F * ggxD(h, alpha) * ggxG(lo, li, alpha) / (4.0f * lo.z * li.z));
  }
  return spl(0.0f);
}
float bsdfPdf(const sundog_flat_material& m, V wo, V wi, V n, float eta) {
  float ci = dot(wi, n), co = dot(wo, n);
  if (co <= 0.0f) return 0.0f;
  bool roughGlass = m.kind == 2 && m.roughness >= 1e-3f;
  if (ci <= 0.0f) {
    if (!roughGlass || ci == 0.0f) return 0.0f;
    float alpha = m.roughness * m.roughness;
    Onb onb(n);
    V lo = onb.toLocal(wo), li = onb.toLocal(wi);
    V h = norm(eta * lo + li);
    if (h.z < 0.0f) h = -h;
    float loh = dot(lo, h);
    if (loh <= 0.0f) return 0.0f;
    float f0 = (1.0f - eta) / (1.0f + eta);
    f0 = f0 * f0;
    float F = fresnelDielectric(loh, eta, f0);
    return (1.0f - F) * ggxPdfTransmit(lo, li, h, eta, alpha);
  }
  if (m.kind == 0) return ci * INV_PI;
  if (m.kind == 5) {
    Onb onb(n);
    Plastic t = plasticTerms(m, spl(0.0f), onb.toLocal(wo), onb.toLocal(wi));
    return 0.5f * (t.pSpec + t.pDiff);
  }
  if (m.kind == 1 && m.roughness >= 1e-3f) {
    Onb onb(n);
    return ggxPdf(onb.toLocal(wo), onb.toLocal(wi), m.roughness * m.roughness);
  }
  if (roughGlass) {
    float alpha = m.roughness * m.roughness;
    Onb onb(n);
    V lo = onb.toLocal(wo), li = onb.toLocal(wi);
    V h = norm(lo + li);
    float f0 = (1.0f - eta) / (1.0f + eta);
    f0 = f0 * f0;
    float F = fresnelDielectric(dot(lo, h), eta, f0);
    return F * ggxPdf(lo, li, alpha);
  }
  return 0.0f;
}
struct BsdfSample {
  V wi = spl(0.0f);
  V weight = spl(0.0f);
  float pdf = 0.0f;
  bool isDelta = false;
  bool valid = false;
};
BsdfSample bsdfSample(const sundog_flat_material& m, V albedo, V rayDir, V n, bool frontface, float etaExt,
                      Rng& rng) {
  BsdfSample s;
  V wo = -rayDir;
  switch (m.kind) {
    case 0: {
      Onb onb(n);
      float a = rng.rnd();
      float b = rng.rnd();
      V li = cosineHemisphere(a, b);
      s.wi = onb.toWorld(li);
      s.pdf = fmaxf(li.z, 1e-8f) * INV_PI;
      s.weight = albedo;
      s.valid = li.z > 0.0f;
      return s;
    }
    case 1: {
      if (m.roughness < 1e-3f) {
        s.wi = norm(reflect(rayDir, n));
        if (dot(s.wi, n) <= 0.0f) return s;
        s.weight = schlick3(dot(s.wi, n), albedo);
        s.isDelta = true;
        s.valid = true;
        return s;
      }
      float alpha = m.roughness * m.roughness;
      Onb onb(n);
      V lo = onb.toLocal(wo);
      if (lo.z <= 0.0f) return s;
      float a = rng.rnd();
      float b = rng.rnd();
      V h = ggxSampleVndf(lo, alpha, a, b);
      V li = reflect(-lo, h);
      if (li.z <= 0.0f) return s;
      V F = schlick3(dot(lo, h), albedo);
      s.wi = onb.toWorld(li);
      s.weight = F * (ggxG(lo, li, alpha) / ggxG1(lo, alpha));
      s.pdf = ggxPdf(lo, li, alpha);
      s.valid = s.pdf > 0.0f;
      return s;
    }
    case 4:
    case 2: {
      V refr;
      float eta = frontface ? etaExt / m.ior : m.ior / etaExt;
      float f0 = (1.0f - eta) / (1.0f + eta);
      f0 = f0 * f0;
      if (m.kind == 4 || m.roughness < 1e-3f) {
        s.isDelta = true;
        if (refract(rayDir, n, eta, refr)) {
          float cosine = eta < 1.0f ? -dot(rayDir, n) : -dot(refr, n);
          float R = schlick(cosine, f0);
          if (rng.rnd() < R) s.wi = norm(reflect(rayDir, n));
          else s.wi = norm(refr);
        } else {
          s.wi = norm(reflect(rayDir, n));
        }
        s.weight = spl(1.0f);
        s.valid = true;
        return s;
      }
      float alpha = m.roughness * m.roughness;
      Onb onb(n);
      V lo = onb.toLocal(wo);
      if (lo.z <= 0.0f) return s;
      float a = rng.rnd();
      float b = rng.rnd();
      V h = ggxSampleVndf(lo, alpha, a, b);
      float loh = dot(lo, h);
      if (loh <= 0.0f) return s;
      V li;
      if (!refract(-lo, h, eta, refr)) {
        li = reflect(-lo, h);
        if (li.z <= 0.0f) return s;
        s.pdf = ggxPdf(lo, li, alpha);
      } else {
        float cosine = eta < 1.0f ? loh : -dot(refr, h);
        float R = schlick(cosine, f0);
        if (rng.rnd() < R) {
          li = reflect(-lo, h);
          if (li.z <= 0.0f) return s;
          s.pdf = R * ggxPdf(lo, li, alpha);
        } else {
          li = norm(refr);
          if (li.z >= 0.0f) return s;
          s.pdf = (1.0f - R) * ggxPdfTransmit(lo, li, h, eta, alpha);
        }
      }
      s.wi = onb.toWorld(li);
      float wgt = ggxG(lo, li, alpha) / ggxG1(lo, alpha);
      if (li.z < 0.0f) wgt *= eta * eta;
      s.weight = spl(wgt);
      s.valid = s.pdf > 0.0f;
      return s;
    }
    case 5: {
      Onb onb(n);
      V lo = onb.toLocal(wo);
      if (lo.z <= 0.0f) return s;
      bool coat = rng.rnd() < 0.5f;
      float a = rng.rnd();
      float b = rng.rnd();
      V li;
      if (coat) {
        float rough = fmaxf(m.roughness, 1e-3f);
        V h = ggxSampleVndf(lo, rough * rough, a, b);
        li = reflect(-lo, h);
      } else {
        li = cosineHemisphere(a, b);
      }
      if (li.z <= 0.0f) return s;
      Plastic t = plasticTerms(m, albedo, lo, li);
      s.pdf = 0.5f * (t.pSpec + t.pDiff);
      if (s.pdf <= 0.0f) return s;
      s.wi = onb.toWorld(li);
      s.weight = (spl(t.fSpec) + t.fDiff) * (li.z / s.pdf);
      s.valid = true;
      return s;
    }
    default:
      return s;
  }
}

// ---------------------------------------------------------------------------
// lights (device/light_sample.cuh:24-226)
// ---------------------------------------------------------------------------
// Mesh-light sample for fixed uniforms (src/capi_render.cpp:119-172 builds
// the area CDF; device/light_sample.cuh:94-154 samples it): uc picks a
// triangle with probability area/total by the CDF, (r0, r1) a uniform point in
// it by the square-root warp; the solid-angle pdf is d^2 / (|cos_l| * total).
LightSample sampleMeshLightU(const Scene& S, int li_, V x, float uc, float r0, float r1, int* triOut) {
  const sundog_flat_light& lt = S.f->lights[li_];
  LightSample s;
  const MeshLightData& ml = S.meshLight[li_];
  const sundog_flat_mesh& mesh = S.f->meshes[ml.mesh];
  int lo = 0, hi = mesh.num_tris;
  while (lo + 1 < hi) {
    int mid = (lo + hi) >> 1;
    if (ml.cdf[mid] <= uc) lo = mid;
    else hi = mid;
  }
  if (triOut) *triOut = lo;
  uint32_t i0 = mesh.indices[3 * lo], i1 = mesh.indices[3 * lo + 1], i2 = mesh.indices[3 * lo + 2];
  V p0 = ml.wp[i0], p1 = ml.wp[i1], p2 = ml.wp[i2];
  float su = sqrtf(r0);
  float bx = su * (1.0f - r1);
  float by = su * r1;
  V p = p0 * (1.0f - bx - by) + p1 * bx + p2 * by;
  V ngU = cross(p1 - p0, p2 - p0);
  float ngL2 = len2(ngU);
  if (ngL2 < 1e-20f) return s;
  V ng = ngU * (ml.ngSign / sqrtf(ngL2));
  V to = p - x;
  float d2 = len2(to);
  if (d2 < 1e-10f) return s;
  float dist = sqrtf(d2);
  V wi = to / dist;
  float cosL = -dot(wi, ng);
  if (!lt.two_sided && cosL <= 0.0f) return s;
  cosL = fabsf(cosL);
  if (cosL < 1e-6f) return s;
  s.wi = wi;
  s.dist = dist;
  s.Li = ld(lt.L);
  if (lt.tex_id >= 0) {
    float tu, tv;
    if (mesh.uvs) {
      tu = mesh.uvs[2 * i0] * (1.0f - bx - by) + mesh.uvs[2 * i1] * bx + mesh.uvs[2 * i2] * by;
      tv = mesh.uvs[2 * i0 + 1] * (1.0f - bx - by) + mesh.uvs[2 * i1 + 1] * bx + mesh.uvs[2 * i2 + 1] * by;
    } else {
      tu = bx;
      tv = by;
    }
    float c[4];
    evalTexture(S.f->textures[lt.tex_id], tu, tv, c);
    s.Li = ld(lt.L) * mk(c[0], c[1], c[2]);
  }
  s.pdf = d2 / (cosL * ml.area);
  s.valid = true;
  return s;
}

LightSample sampleLight(const Scene& S, int li_, V x, Rng& rng) {
  const sundog_flat_light& lt = S.f->lights[li_];
  LightSample s;
  switch (lt.kind) {
    case 0:
    case 1: {
      V p;
      float tu, tv;
      if (lt.kind == 0) {
        float a = 2.0f * rng.rnd() - 1.0f;
        float b = 2.0f * rng.rnd() - 1.0f;
        p = ld(lt.p) + a * ld(lt.u) + b * ld(lt.v);
        tu = (a + 1.0f) * 0.5f;
        tv = (b + 1.0f) * 0.5f;
      } else {
        float r0 = rng.rnd();
        float r1 = rng.rnd();
        float dx, dy;
        concentricDisk(r0, r1, dx, dy);
        p = ld(lt.p) + dx * ld(lt.u) + dy * ld(lt.v);
        tu = (atan2f(dy, dx) + PI) / (2.0f * PI);
        tv = sqrtf(dx * dx + dy * dy);
      }
      V to = p - x;
      float d2 = len2(to);
      if (d2 < 1e-10f) return s;
      float dist = sqrtf(d2);
      V wi = to / dist;
      float cosL = -dot(wi, ld(lt.n));
      if (!lt.two_sided && cosL <= 0.0f) return s;
      cosL = fabsf(cosL);
      if (cosL < 1e-6f) return s;
      s.wi = wi;
      s.dist = dist;
      s.Li = ld(lt.L);
      if (lt.tex_id >= 0) {
        float c[4];
        evalTexture(S.f->textures[lt.tex_id], tu, tv, c);
        s.Li = ld(lt.L) * mk(c[0], c[1], c[2]);
      }
      s.pdf = d2 / (cosL * S.lightArea[li_]);
      s.valid = true;
      return s;
    }
    case 2: {
      V to = ld(lt.p) - x;
      float d2 = len2(to);
      float r2 = lt.radius * lt.radius;
      if (d2 <= r2 * 1.0001f) return s;
      float dist = sqrtf(d2);
      V wc = to / dist;
      float sin2Max = r2 / d2;
      float cosMax = sqrtf(fmaxf(0.0f, 1.0f - sin2Max));
      float u0 = rng.rnd();
      float u1 = rng.rnd();
      float cosT = 1.0f - u0 * (1.0f - cosMax);
      float sinT = sqrtf(fmaxf(0.0f, 1.0f - cosT * cosT));
      float phi = 2.0f * PI * u1;
      Onb onb(wc);
      s.wi = onb.toWorld(mk(sinT * cosf(phi), sinT * sinf(phi), cosT));
      float b = dot(s.wi, to);
      float det = fmaxf(0.0f, b * b - d2 + r2);
      s.dist = b - sqrtf(det);
      s.Li = ld(lt.L);
      s.pdf = 1.0f / (2.0f * PI * (1.0f - cosMax));
      s.valid = true;
      return s;
    }
    case 5: {
      // three draws in this order: triangle pick, then the two barycentric
      float uc = rng.rnd();
      float r0 = rng.rnd();
      float r1 = rng.rnd();
      return sampleMeshLightU(S, li_, x, uc, r0, r1, nullptr);
    }
    case 3: {
      s.isDelta = true;
      V to = ld(lt.p) - x;
      float d2c = len2(to);
      if (d2c < 1e-10f) return s;
      V wc = to / sqrtf(d2c);
      float r0 = rng.rnd();
      float r1 = rng.rnd();
      float dx, dy;
      concentricDisk(r0, r1, dx, dy);
      dx *= lt.radius;
      dy *= lt.radius;
      Onb onb(wc);
      V p = ld(lt.p) + dx * onb.t + dy * onb.b;
      V top = p - x;
      float d2 = len2(top);
      float dist = sqrtf(d2);
      s.wi = top / dist;
      s.dist = fmaxf(dist - 1e-4f, 1e-4f);
      s.Li = ld(lt.L) / d2;
      s.pdf = 1.0f;
      s.valid = true;
      return s;
    }
    case 4: {
      s.isDelta = true;
      s.wi = -ld(lt.dir);
      s.dist = 1e16f;
      s.Li = ld(lt.L);
      s.pdf = 1.0f;
      s.valid = true;
      return s;
    }
  }
  return s;
}
float lightPdfSolidAngle(const Scene& S, int li_, V from, V hitP, V nHit) {
  const sundog_flat_light& lt = S.f->lights[li_];
  switch (lt.kind) {
    case 0:
    case 1:
    case 5: {
      V to = hitP - from;
      float d2 = len2(to);
      if (d2 < 1e-10f) return 0.0f;
      float dist = sqrtf(d2);
      float cosL = fabsf(dot(to / dist, lt.kind == 5 ? nHit : ld(lt.n)));
      if (cosL < 1e-6f) return 0.0f;
      return d2 / (cosL * S.lightArea[li_]);
    }
    case 2: {
      V to = ld(lt.p) - from;
      float d2 = len2(to);
      float r2 = lt.radius * lt.radius;
      if (d2 <= r2 * 1.0001f) return 0.0f;
      float cosMax = sqrtf(fmaxf(0.0f, 1.0f - r2 / d2));
      return 1.0f / (2.0f * PI * (1.0f - cosMax));
    }
    default:
      return 0.0f;
  }
}

// ---------------------------------------------------------------------------
// geometry: quadrics (device/intersect.cuh:37-162) and the watertight
// triangle test (Woop, Benthin, Wald, JCGT 2013), same operation order as
// the GPU; ray transforms by the provided inverse.
// ---------------------------------------------------------------------------
V xfP(const float* m, V p) {
  return mk(m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3], m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7],
            m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11]);
}
V xfV(const float* m, V v) {
  return mk(m[0] * v.x + m[1] * v.y + m[2] * v.z, m[4] * v.x + m[5] * v.y + m[6] * v.z,
            m[8] * v.x + m[9] * v.y + m[10] * v.z);
}
V xfN(const float* inv, V n) {  // inverse transpose
  return mk(inv[0] * n.x + inv[4] * n.y + inv[8] * n.z, inv[1] * n.x + inv[5] * n.y + inv[9] * n.z,
            inv[2] * n.x + inv[6] * n.y + inv[10] * n.z);
}

int quadricRoots(int kind, V o, V d, float tmin, float tmax, float t[2]) {
  int n = 0;
  if (kind == 0) {
    float a = dot(d, d), b = dot(o, d), c = dot(o, o) - 1.0f;
    float disc = b * b - a * c;
    if (disc <= 0.0f) return 0;
    float s = sqrtf(disc);
    float r[2] = {(-b - s) / a, (-b + s) / a};
    for (int i = 0; i < 2; i++)
      if (r[i] > tmin && r[i] < tmax) t[n++] = r[i];
    return n;
  }
  if (kind == 1 || kind == 2) {
    if (d.y == 0.0f) return 0;
    float tt = -o.y / d.y;
    if (tt <= tmin || tt >= tmax) return 0;
    float x = o.x + tt * d.x, z = o.z + tt * d.z;
    if (kind == 1 ? (fabsf(x) > 1.0f || fabsf(z) > 1.0f) : (sqrtf(x * x + z * z) > 1.0f)) return 0;
    t[0] = tt;
    return 1;
  }
  if (kind == 3) {
    float a = d.x * d.x + d.z * d.z;
    if (a == 0.0f) return 0;
    float b = o.x * d.x + o.z * d.z, c = o.x * o.x + o.z * o.z - 1.0f;
    float disc = b * b - a * c;
    if (disc <= 0.0f) return 0;
    float s = sqrtf(disc);
    float r[2] = {(-b - s) / a, (-b + s) / a};
    for (int i = 0; i < 2; i++)
      if (r[i] > tmin && r[i] < tmax && fabsf(o.y + r[i] * d.y) <= 1.0f) t[n++] = r[i];
    return n;
  }
  if (kind == 4) {
    float A = 0.5f * (d.x * d.x + d.z * d.z);
    float B = o.x * d.x + o.z * d.z - d.y;
    float C = 0.5f * (o.x * o.x + o.z * o.z) - o.y;
    float r[2];
    int m;
    if (fabsf(A) < 1e-12f) {
      if (B == 0.0f) return 0;
      r[0] = -C / B;
      m = 1;
    } else {
      float disc = B * B - 4.0f * A * C;
      if (disc <= 0.0f) return 0;
      float s = sqrtf(disc);
      r[0] = (-B - s) / (2.0f * A);
      r[1] = (-B + s) / (2.0f * A);
      m = 2;
    }
    for (int i = 0; i < m; i++) {
      if (r[i] > tmin && r[i] < tmax) {
        float px = o.x + r[i] * d.x, pz = o.z + r[i] * d.z;
        if (px * px + pz * pz <= 1.0f) t[n++] = r[i];
      }
    }
    return n;
  }
  return 0;
}
void quadricAttr(int kind, V p, V& n, float& u, float& v) {
  switch (kind) {
    case 0:
      n = norm(p);
      u = (atan2f(n.z, n.x) + PI) / (2.0f * PI);
      v = (asinf(clampf(n.y, -1.0f, 1.0f)) + PI / 2.0f) / PI;
      return;
    case 1:
      n = mk(0, 1, 0);
      u = (p.x + 1.0f) * 0.5f;
      v = (p.z + 1.0f) * 0.5f;
      return;
    case 2:
      n = mk(0, 1, 0);
      u = (atan2f(p.z, p.x) + PI) / (2.0f * PI);
      v = sqrtf(p.x * p.x + p.z * p.z);
      return;
    case 3:
      n = norm(mk(p.x, 0.0f, p.z));
      u = (atan2f(p.z, p.x) + PI) / (2.0f * PI);
      v = (p.y + 1.0f) * 0.5f;
      return;
    default:
      n = norm(mk(p.x, -1.0f, p.z));
      u = (atan2f(p.z, p.x) + PI) / (2.0f * PI);
      v = sqrtf(p.x * p.x + p.z * p.z);
  }
}

struct TriRay {
  int kx, ky, kz;
  float Sx, Sy, Sz;
};
TriRay triSetup(V d) {
  TriRay r;
  float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
  r.kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
  r.kx = (r.kz + 1) % 3;
  r.ky = (r.kx + 1) % 3;
  float dz = at(d, r.kz);
  if (dz < 0.0f) std::swap(r.kx, r.ky);
  r.Sz = 1.0f / dz;  // shear factors as products with 1 / d_z, as the GPU computes them
  r.Sx = at(d, r.kx) * r.Sz;
  r.Sy = at(d, r.ky) * r.Sz;
  return r;
}
bool triIntersect(const TriRay& R, V o, V v0, V v1, V v2, float tmin, float tmax, float& t, float& b1,
                  float& b2) {
  V A = v0 - o, B = v1 - o, C = v2 - o;
  float Az0 = at(A, R.kz), Bz0 = at(B, R.kz), Cz0 = at(C, R.kz);
  float Ax = at(A, R.kx) - R.Sx * Az0, Ay = at(A, R.ky) - R.Sy * Az0;
  float Bx = at(B, R.kx) - R.Sx * Bz0, By = at(B, R.ky) - R.Sy * Bz0;
  float Cx = at(C, R.kx) - R.Sx * Cz0, Cy = at(C, R.ky) - R.Sy * Cz0;
  float U = Cx * By - Cy * Bx;
  float Vv = Ax * Cy - Ay * Cx;
  float W = Bx * Ay - By * Ax;
  if (U == 0.0f || Vv == 0.0f || W == 0.0f) {
    U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
    Vv = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
    W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
  }
  if ((U < 0.0f || Vv < 0.0f || W < 0.0f) && (U > 0.0f || Vv > 0.0f || W > 0.0f)) return false;
  float det = U + Vv + W;
  if (det == 0.0f) return false;
  float T = U * (R.Sz * Az0) + Vv * (R.Sz * Bz0) + W * (R.Sz * Cz0);
  float rcp = 1.0f / det;
  float tt = T * rcp;
  if (!(tt > tmin && tt < tmax)) return false;
  t = tt;
  b1 = Vv * rcp;
  b2 = W * rcp;
  return true;
}

// ---------------------------------------------------------------------------
// BVH (own binary median-split build; the result of a closest-hit query is
// the lexicographic minimum (t, instance, primitive), so it is independent
// of the tree shape)
// ---------------------------------------------------------------------------
struct Box {
  float lo[3] = {3e38f, 3e38f, 3e38f}, hi[3] = {-3e38f, -3e38f, -3e38f};
  void add(V p) {
    lo[0] = fminf(lo[0], p.x); lo[1] = fminf(lo[1], p.y); lo[2] = fminf(lo[2], p.z);
    hi[0] = fmaxf(hi[0], p.x); hi[1] = fmaxf(hi[1], p.y); hi[2] = fmaxf(hi[2], p.z);
  }
  void add(const Box& b) {
    for (int i = 0; i < 3; i++) {
      lo[i] = fminf(lo[i], b.lo[i]);
      hi[i] = fmaxf(hi[i], b.hi[i]);
    }
  }
};
int buildNode(Bvh2& b, std::vector<Box>& boxes, int first, int count) {
  int id = (int)b.nodes.size();
  b.nodes.emplace_back();
  Box bb, cb;
  for (int i = first; i < first + count; i++) {
    const Box& x = boxes[b.prims[i]];
    bb.add(x);
    cb.add(mk(0.5f * (x.lo[0] + x.hi[0]), 0.5f * (x.lo[1] + x.hi[1]), 0.5f * (x.lo[2] + x.hi[2])));
  }
  memcpy(b.nodes[id].lo, bb.lo, sizeof bb.lo);
  memcpy(b.nodes[id].hi, bb.hi, sizeof bb.hi);
  if (count <= 4) {
    b.nodes[id].first = first;
    b.nodes[id].count = count;
    return id;
  }
  int axis = 0;
  float ext = cb.hi[0] - cb.lo[0];
  for (int a = 1; a < 3; a++)
    if (cb.hi[a] - cb.lo[a] > ext) {
      ext = cb.hi[a] - cb.lo[a];
      axis = a;
    }
  int mid = first + count / 2;
  std::nth_element(b.prims.begin() + first, b.prims.begin() + mid, b.prims.begin() + first + count,
                   [&](int p, int q) {
                     float cp = boxes[p].lo[axis] + boxes[p].hi[axis];
                     float cq = boxes[q].lo[axis] + boxes[q].hi[axis];
                     return cp < cq || (cp == cq && p < q);
                   });
  int l = buildNode(b, boxes, first, mid - first);
  int r = buildNode(b, boxes, mid, first + count - mid);
  b.nodes[id].left = l;
  b.nodes[id].right = r;
  return id;
}
Bvh2 buildBvh(std::vector<Box>& boxes) {
  Bvh2 b;
  b.prims.resize(boxes.size());
  for (size_t i = 0; i < boxes.size(); i++) b.prims[i] = (int)i;
  if (!boxes.empty()) buildNode(b, boxes, 0, (int)boxes.size());
  return b;
}
// Slab test with the far distance widened by 2*gamma(3) (Ize 2013).
bool hitBox(const Node2& n, V o, V inv, float tmin, float tmax, float& tnear) {
  float t0x = (n.lo[0] - o.x) * inv.x, t1x = (n.hi[0] - o.x) * inv.x;
  float t0y = (n.lo[1] - o.y) * inv.y, t1y = (n.hi[1] - o.y) * inv.y;
  float t0z = (n.lo[2] - o.z) * inv.z, t1z = (n.hi[2] - o.z) * inv.z;
  float tn = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fmaxf(fminf(t0z, t1z), tmin));
  float tf = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fminf(fmaxf(t0z, t1z), tmax));
  tnear = tn;
  return tn <= tf * 1.0000004f;
}

struct Hit {
  float t;
  int inst, prim;
  float b1, b2;
  V oObj, dObj;
};
bool better(float t, int inst, int prim, const Hit& h) {
  return t < h.t || (t == h.t && (inst < h.inst || (inst == h.inst && prim < h.prim)));
}

struct Shade {
  V nFront;
  float u, v;
  bool front;
};
// Hit resolve (device/programs.cu:135-190); the front side is the sign of
// the object-space cosine, which equals the sign of the world-space cosine
// under the inverse-transpose normal map.
Shade shade(const Scene& S, int inst, int prim, float b1, float b2, V oObj, V dObj, float t, bool forShading) {
  const sundog_flat_object& ob = S.f->objects[inst];
  Shade sp;
  if (ob.geom_kind != 5) {
    V p = oObj + t * dObj;
    V n;
    quadricAttr(ob.geom_kind, p, n, sp.u, sp.v);
    sp.front = dot(n, dObj) < 0.0f;
    sp.nFront = norm(xfN(ob.inv, n));
    return sp;
  }
  const sundog_flat_mesh& m = S.f->meshes[ob.mesh_id];
  uint32_t i0 = m.indices[3 * prim], i1 = m.indices[3 * prim + 1], i2 = m.indices[3 * prim + 2];
  V p0 = ld(m.positions + 3 * i0), p1 = ld(m.positions + 3 * i1), p2 = ld(m.positions + 3 * i2);
  V ngObj = cross(p1 - p0, p2 - p0);
  sp.front = dot(ngObj, dObj) < 0.0f;
  V ng = norm(xfN(ob.inv, ngObj));
  V ns = ng;
  float b0 = 1.0f - b1 - b2;
  if (m.normals) {
    V n = ld(m.normals + 3 * i0) * b0 + ld(m.normals + 3 * i1) * b1 + ld(m.normals + 3 * i2) * b2;
    ns = norm(xfN(ob.inv, n));
    if (dot(ns, ng) < 0.0f) ns = -ns;
  }
  sp.nFront = ns;
  if (m.uvs) {
    sp.u = m.uvs[2 * i0] * b0 + m.uvs[2 * i1] * b1 + m.uvs[2 * i2] * b2;
    sp.v = m.uvs[2 * i0 + 1] * b0 + m.uvs[2 * i1 + 1] * b1 + m.uvs[2 * i2 + 1] * b2;
  } else {
    sp.u = b1;
    sp.v = b2;
  }
  if (forShading && ob.light_id >= 0) {
    int side = sp.front ? ob.mat_front : ob.mat_back;
    if (side != NO_MAT && S.f->materials[side].kind == 3) sp.nFront = ng;
  }
  return sp;
}
bool masked(const sundog_flat_object& ob) {
  return ob.mat_front == NO_MAT || ob.mat_back == NO_MAT || ob.cutout_tex >= 0;
}
bool accepts(const Scene& S, int inst, int prim, float b1, float b2, V oObj, V dObj, float t) {
  const sundog_flat_object& ob = S.f->objects[inst];
  if (!masked(ob)) return true;
  Shade sp = shade(S, inst, prim, b1, b2, oObj, dObj, t, false);
  int side = sp.front ? ob.mat_front : ob.mat_back;
  if (side == NO_MAT) return false;
  if (ob.cutout_tex >= 0) {
    float c[4];
    evalTexture(S.f->textures[ob.cutout_tex], sp.u, sp.v, c);
    if (c[3] < 0.5f) return false;
  }
  return true;
}

// Current far bound of a query: the best hit so far (closest-hit queries,
// widened by a hair so equal-t candidates reach the tie rule) or the fixed
// segment end (shadow queries).
struct Bound {
  const float* t;
  bool widen;
  float get() const { return widen ? *t * 1.0000004f + 1e-30f : *t; }
};

// Near-first traversal of a binary tree; visit(i) handles leaf prim i.
template <class F>
void traverse(const Bvh2& b, V o, V d, float tmin, const Bound& bound, F&& visit) {
  if (b.nodes.empty()) return;
  V inv = mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
  int stack[128];
  int sp = 0;
  float tn;
  if (!hitBox(b.nodes[0], o, inv, tmin, *bound.t, tn)) return;
  stack[sp++] = 0;
  while (sp) {
    const Node2& n = b.nodes[stack[--sp]];
    if (n.left < 0) {
      for (int i = n.first; i < n.first + n.count; i++) visit(b.prims[i]);
      continue;
    }
    float tl, tr;
    bool hl = hitBox(b.nodes[n.left], o, inv, tmin, *bound.t, tl);
    bool hr = hitBox(b.nodes[n.right], o, inv, tmin, *bound.t, tr);
    if (hl && hr) {
      if (tl <= tr) {
        stack[sp++] = n.right;
        stack[sp++] = n.left;
      } else {
        stack[sp++] = n.left;
        stack[sp++] = n.right;
      }
    } else if (hl) {
      stack[sp++] = n.left;
    } else if (hr) {
      stack[sp++] = n.right;
    }
  }
}

// Candidate visitor for one instance: all primitives whose t lies in
// (tmin, bound) are passed to fn(t, prim, b1, b2, oObj, dObj).
template <class F>
void forInstance(const Scene& S, int inst, V o, V d, float tmin, const Bound& bound, bool brute, F&& fn) {
  const sundog_flat_object& ob = S.f->objects[inst];
  V oo = xfP(ob.inv, o), dd = xfV(ob.inv, d);
  if (ob.geom_kind != 5) {
    float ts[2];
    int n = quadricRoots(ob.geom_kind, oo, dd, tmin, bound.get(), ts);
    for (int k = 0; k < n; k++) fn(ts[k], k, 0.0f, 0.0f, oo, dd);
    return;
  }
  const sundog_flat_mesh& m = S.f->meshes[ob.mesh_id];
  TriRay tr = triSetup(dd);
  auto testTri = [&](int t) {
    V v0 = ld(m.positions + 3 * m.indices[3 * t]), v1 = ld(m.positions + 3 * m.indices[3 * t + 1]),
      v2 = ld(m.positions + 3 * m.indices[3 * t + 2]);
    float tt, b1, b2;
    if (triIntersect(tr, oo, v0, v1, v2, tmin, bound.get(), tt, b1, b2)) fn(tt, t, b1, b2, oo, dd);
  };
  if (brute) {
    for (int t = 0; t < m.num_tris; t++) testTri(t);
    return;
  }
  traverse(S.blas[ob.mesh_id], oo, dd, tmin, bound, testTri);
}

template <class F>
void forAllInstances(const Scene& S, V o, V d, float tmin, const Bound& bound, bool brute, F&& fn) {
  int n = S.f->num_objects;
  if (n == 0) return;
  if (brute) {
    for (int i = 0; i < n; i++) fn(i);
    return;
  }
  traverse(S.tlas, o, d, tmin, bound, fn);
}

bool closestHit(const Scene& S, V o, V d, float tmin, float tmax, bool brute, Hit& h) {
  h.t = tmax;
  h.inst = h.prim = 0x7fffffff;
  h.b1 = h.b2 = 0.0f;
  bool found = false;
  // Candidates within a hair of the current best still reach the tie rule.
  Bound bound{&h.t, true};
  forAllInstances(S, o, d, tmin, bound, brute, [&](int inst) {
    forInstance(S, inst, o, d, tmin, bound, brute,
                [&](float t, int prim, float b1, float b2, V oo, V dd) {
                  if (better(t, inst, prim, h) && accepts(S, inst, prim, b1, b2, oo, dd, t)) {
                    h.t = t;
                    h.inst = inst;
                    h.prim = prim;
                    h.b1 = b1;
                    h.b2 = b2;
                    h.oObj = oo;
                    h.dObj = dd;
                    found = true;
                  }
                });
  });
  return found;
}

// Shadow transmittance with transparent crossings (device/programs.cu:70-85,
// :247-292).
V shadowTransmittance(const Scene& S, V o, V d, float tmax) {
  bool occluded = false;
  float trF = 1.0f;
  V tau = spl(0.0f);
  Bound bound{&tmax, false};
  forAllInstances(S, o, d, 0.0f, bound, false, [&](int inst) {
    if (occluded) return;
    const sundog_flat_object& ob = S.f->objects[inst];
    forInstance(S, inst, o, d, 0.0f, bound, false, [&](float t, int prim, float b1, float b2, V oo, V dd) {
      if (occluded) return;
      Shade sp = shade(S, inst, prim, b1, b2, oo, dd, t, false);
      int side = sp.front ? ob.mat_front : ob.mat_back;
      if (side == NO_MAT) return;
      if (ob.cutout_tex >= 0) {
        float c[4];
        evalTexture(S.f->textures[ob.cutout_tex], sp.u, sp.v, c);
        if (c[3] < 0.5f) return;
      }
      const sundog_flat_material& mat = S.f->materials[side];
      if ((mat.kind != 2 && mat.kind != 4) || !S.f->transparent_shadows) {
        occluded = true;
        return;
      }
      V n = sp.front ? sp.nFront : -sp.nFront;
      float eta = sp.front ? 1.0f / mat.ior : mat.ior;
      V refr;
      if (!refract(d, n, eta, refr)) return;
      float f0 = (1.0f - eta) / (1.0f + eta);
      f0 = f0 * f0;
      float cosine = sp.front ? -dot(d, n) : -dot(refr, n);
      float F = schlick(cosine, f0);
      trF = trF * (1.0f - F);
      float s = sp.front ? -t : t;
      tau = tau + ld(mat.absorb) * s;
    });
  });
  if (occluded) return spl(0.0f);
  tau = mk(fmaxf(0.0f, tau.x), fmaxf(0.0f, tau.y), fmaxf(0.0f, tau.z));
  V tr = spl(trF);
  if (maxComp(tau) > 0.0f) tr = tr * expv(-tau);
  return tr;
}

// ---------------------------------------------------------------------------
// scene setup
// ---------------------------------------------------------------------------
void setupScene(const sundog_flat_scene* f, Scene& S) {
  S.f = f;
  S.blas.assign(f->num_meshes, Bvh2{});
  for (int mi = 0; mi < f->num_meshes; mi++) {
    const sundog_flat_mesh& m = f->meshes[mi];
    std::vector<Box> boxes(m.num_tris);
    for (int t = 0; t < m.num_tris; t++)
      for (int k = 0; k < 3; k++) boxes[t].add(ld(m.positions + 3 * m.indices[3 * t + k]));
    S.blas[mi] = buildBvh(boxes);
  }
  std::vector<Box> inst(f->num_objects);
  for (int i = 0; i < f->num_objects; i++) {
    const sundog_flat_object& ob = f->objects[i];
    Box ob0;
    if (ob.geom_kind == 5) {
      const Node2& r = S.blas[ob.mesh_id].nodes[0];
      memcpy(ob0.lo, r.lo, sizeof ob0.lo);
      memcpy(ob0.hi, r.hi, sizeof ob0.hi);
    } else {
      float e = 1e-4f;
      float lo[3] = {-1, -1, -1}, hi[3] = {1, 1, 1};
      if (ob.geom_kind == 1 || ob.geom_kind == 2) { lo[1] = -e; hi[1] = e; }
      if (ob.geom_kind == 4) { lo[1] = -e; hi[1] = 0.5f + e; }
      memcpy(ob0.lo, lo, sizeof lo);
      memcpy(ob0.hi, hi, sizeof hi);
    }
    for (int c = 0; c < 8; c++) {
      V p = mk((c & 1) ? ob0.hi[0] : ob0.lo[0], (c & 2) ? ob0.hi[1] : ob0.lo[1], (c & 4) ? ob0.hi[2] : ob0.lo[2]);
      inst[i].add(xfP(ob.xform, p));
    }
    for (int a = 0; a < 3; a++) {  // widen against transform rounding
      float mag = fmaxf(fabsf(inst[i].lo[a]), fabsf(inst[i].hi[a]));
      inst[i].lo[a] -= mag * 1e-5f + 1e-6f;
      inst[i].hi[a] += mag * 1e-5f + 1e-6f;
    }
  }
  S.tlas = buildBvh(inst);
  S.envImportance = false;
  if (f->bg_kind == 2 && f->env_rgba) {
    bool ok = buildEnvCdf(f, S.marg, S.cond);
    S.envImportance = ok && f->env_importance;
  }
  // mesh-light CDFs (src/capi_render.cpp:119-172)
  S.meshLight.assign(f->num_lights, MeshLightData{});
  S.lightArea.assign(f->num_lights, 0.0f);
  for (int li = 0; li < f->num_lights; li++) S.lightArea[li] = f->lights[li].area;
  for (int oi = 0; oi < f->num_objects; oi++) {
    const sundog_flat_object& ob = f->objects[oi];
    if (ob.geom_kind != 5 || ob.light_id < 0) continue;
    const sundog_flat_mesh& m = f->meshes[ob.mesh_id];
    MeshLightData& ml = S.meshLight[ob.light_id];
    ml.mesh = ob.mesh_id;
    ml.wp.resize(m.num_verts);
    for (int i = 0; i < m.num_verts; i++) ml.wp[i] = xfP(ob.xform, ld(m.positions + 3 * i));
    ml.cdf.assign(m.num_tris + 1, 0.0f);
    double total = 0.0;
    for (int t = 0; t < m.num_tris; t++) {
      V a = ml.wp[m.indices[3 * t]], b = ml.wp[m.indices[3 * t + 1]], c = ml.wp[m.indices[3 * t + 2]];
      total += 0.5 * (double)len(cross(b - a, c - a));
      ml.cdf[t + 1] = (float)total;
    }
    if (!(total > 0.0)) throw std::runtime_error("mesh NEE light has zero surface area");
    for (int t = 1; t < m.num_tris; t++) ml.cdf[t] = (float)(ml.cdf[t] / total);
    ml.cdf[m.num_tris] = 1.0f;
    ml.area = (float)total;
    S.lightArea[ob.light_id] = ml.area;
    V ux = xfV(ob.xform, mk(1, 0, 0)), uy = xfV(ob.xform, mk(0, 1, 0)), uz = xfV(ob.xform, mk(0, 0, 1));
    ml.ngSign = dot(uy, cross(uz, ux)) < 0.0f ? -1.0f : 1.0f;
  }
}

// ---------------------------------------------------------------------------
// the path loop (device/programs.cu:311-572)
// ---------------------------------------------------------------------------
V background(const Scene& S, V dir) {
  const sundog_flat_scene* f = S.f;
  if (f->bg_kind == 2) return envEval(S, dir);
  if (f->bg_kind == 1) {
    float t = 0.5f * (dir.y + 1.0f);
    V a = ld(f->bg_a), b = ld(f->bg_b);
    return a + (b - a) * t;
  }
  return ld(f->bg_a);
}

// Strata per axis over the total spp: the largest n with n * n <= spp
// (rounded square root, stepped down when it overshoots).
int strataPerAxis(int spp) {
  int n = (int)floorf(sqrtf((float)spp) + 0.5f);
  if (n * n > spp) n--;
  return n < 1 ? 1 : n;
}

// Jitter of sample s in its pixel (device/programs.cu:316-340): sample s <
// n^2 lies in stratum (s mod n, s div n) of the n x n grid, later samples are
// plain uniforms; two draws either way.
void cameraJitter(int s, int nStrata, Rng& rng, float& jx, float& jy) {
  if (s < nStrata * nStrata) {
    jx = ((s % nStrata) + rng.rnd()) / nStrata;
    jy = ((s / nStrata) + rng.rnd()) / nStrata;
  } else {
    jx = rng.rnd();
    jy = rng.rnd();
  }
}

void renderPixel(const Scene& S, int px, int py, int s0, int s1, bool brute, float* acc, float* alb, float* nrm,
                 uint64_t& radiance, uint64_t& shadow) {
  const sundog_flat_scene* f = S.f;
  const int W = f->width, H = f->height;
  const int pixel = py * W + px;
  const float* cam = f->cam;
  V camOrigin = ld(cam), lowerLeft = ld(cam + 3), horizontal = ld(cam + 6), vertical = ld(cam + 9);
  V cu = ld(cam + 12), cv = ld(cam + 15), cw = ld(cam + 18);
  float lens = cam[21];
  const int nStrata = strataPerAxis(f->spp);
  const bool hasEnv = f->bg_kind == 2;
  const int nStrat = f->num_lights + (hasEnv ? 1 : 0);
  float* A = acc + 4 * (size_t)pixel;
  float* Al = alb + 4 * (size_t)pixel;
  float* N = nrm + 4 * (size_t)pixel;
  for (int s = s0; s < s1; s++) {
    Rng rng = Rng::init(((uint64_t)pixel << 32) ^ (uint64_t)s, (uint64_t)f->seed);
    float jx, jy;
    cameraJitter(s, nStrata, rng, jx, jy);
    float u = (px + jx) / W;
    float v = ((H - 1 - py) + jy) / H;
    V org = camOrigin;
    if (lens > 0.0f) {
      float r0 = rng.rnd();
      float r1 = rng.rnd();
      float dx, dy;
      concentricDisk(r0, r1, dx, dy);
      org = org + cu * (dx * lens) + cv * (dy * lens);
    }
    V dir = norm(lowerLeft + u * horizontal + v * vertical - org);
    V L = spl(0.0f), beta = spl(1.0f);
    float prevPdf = 0.0f;
    bool specular = true;
    V prevX = org, o = org, d = dir;
    bool aovDone = false;
    const int MED = 4;
    V medSigma[MED];
    float medIor[MED];
    int medDepth = 0;
    for (int depth = 0; depth < f->max_depth; depth++) {
      Hit hit;
      bool isHit = closestHit(S, o, d, 0.0f, 1e16f, brute, hit);
      radiance++;
      if (medDepth > 0) {
        V sigma = medSigma[std::min(medDepth, MED) - 1];
        if (maxComp(sigma) > 0.0f) beta = beta * expv(-sigma * (isHit ? hit.t : 1e4f));
      }
      if (f->num_flames > 0) {
        V trans = spl(1.0f);
        V Lv = marchFlames(S, o, d, isHit ? hit.t : 1e16f, rng, trans);
        V c = beta * Lv;
        if (depth >= 1 && f->clamp > 0.0f) c = vminv(c, spl(f->clamp));
        L = L + c;
        beta = beta * trans;
      }
      if (!isHit) {
        V bg = background(S, d);
        float w = 1.0f;
        if (hasEnv && !specular) {
          float pdfE = envPdfSolidAngle(S, d) / nStrat;
          w = prevPdf / (prevPdf + pdfE);
        }
        V c = beta * bg * w;
        if (depth >= 1 && f->clamp > 0.0f) c = vminv(c, spl(f->clamp));
        L = L + c;
        if (!aovDone) {
          float b[3] = {fminf(bg.x, 1.0f), fminf(bg.y, 1.0f), fminf(bg.z, 1.0f)};
          for (int k = 0; k < 3; k++) {
            Al[k] = Al[k] + (b[k] - Al[k]) / (s + 1);
            N[k] = N[k] - N[k] / (s + 1);
          }
          Al[3] = N[3] = 1.0f;
          aovDone = true;
        }
        break;
      }
      const sundog_flat_object& ob = f->objects[hit.inst];
      Shade sp = shade(S, hit.inst, hit.prim, hit.b1, hit.b2, hit.oObj, hit.dObj, hit.t, true);
      int matId = sp.front ? ob.mat_front : ob.mat_back;
      V ns = sp.front ? sp.nFront : -sp.nFront;
      V x = o + hit.t * d;
      const sundog_flat_material& mat = f->materials[matId];
      V albedo = materialAlbedo(S, mat, sp.u, sp.v);
      if (mat.kind == 4 && mat.wave_amp > 0.0f)
        ns = waterNormal(x, ns, -d, mat.wave_amp, mat.wave_freq, mat.wave_octaves);
      if (!aovDone) {
        float nc[3] = {dot(ns, cu), dot(ns, cv), dot(ns, cw)};
        float ab[3] = {albedo.x, albedo.y, albedo.z};
        for (int k = 0; k < 3; k++) {
          Al[k] = Al[k] + (ab[k] - Al[k]) / (s + 1);
          N[k] = N[k] + (nc[k] - N[k]) / (s + 1);
        }
        Al[3] = N[3] = 1.0f;
        aovDone = true;
      }
      if (mat.kind == 3) {
        if (sp.front || mat.two_sided) {
          V Le = albedo * mat.intensity;
          float w = 1.0f;
          if (!specular && ob.light_id >= 0 && f->num_lights > 0) {
            float pdfL = lightPdfSolidAngle(S, ob.light_id, prevX, x, ns) / nStrat;
            w = prevPdf / (prevPdf + pdfL);
          }
          V c = beta * Le * w;
          if (depth >= 1 && f->clamp > 0.0f) c = vminv(c, spl(f->clamp));
          L = L + c;
        }
        break;
      }
      float etaExt = 1.0f;
      if (mat.kind == 2 || mat.kind == 4) {
        int outer = std::min(sp.front ? medDepth : medDepth - 1, MED);
        if (outer > 0) etaExt = medIor[outer - 1];
      }
      float etaRel = sp.front ? etaExt / mat.ior : mat.ior / etaExt;
      if (nStrat > 0 && !bsdfIsDelta(mat)) {
        int k = std::min((int)(rng.rnd() * nStrat), nStrat - 1);
        LightSample ls = (k < f->num_lights) ? sampleLight(S, k, x, rng) : sampleEnv(S, rng);
        float cosS = ls.valid ? dot(ls.wi, ns) : 0.0f;
        bool transmitNee = cosS < 0.0f && mat.kind == 2;
        if (ls.valid && (cosS > 0.0f || transmitNee)) {
          shadow++;
          V so = offsetRay(x, transmitNee ? -ns : ns);
          float tmaxS = ls.dist * 0.999f;
          V tr = shadowTransmittance(S, so, ls.wi, tmaxS);
          if (maxComp(tr) > 0.0f) {
            if (f->num_flames > 0 && f->transparent_shadows) {
              int skip = (k < f->num_lights) ? f->lights[k].flame_id : -1;
              tr = tr * flameTransmittance(S, so, ls.wi, tmaxS, rng, skip);
            }
            float pdfLe = (ls.isDelta ? 1.0f : ls.pdf) / nStrat;
            V fv = bsdfEval(mat, albedo, -d, ls.wi, ns, etaRel);
            float w = 1.0f;
            if (!ls.isDelta) {
              float pdfB = bsdfPdf(mat, -d, ls.wi, ns, etaRel);
              w = pdfLe / (pdfLe + pdfB);
            }
            V c = beta * fv * fabsf(cosS) * ls.Li * w / pdfLe;
            c = c * tr;
            if (depth >= 1 && f->clamp > 0.0f) c = vminv(c, spl(f->clamp));
            L = L + sanitize(c);
          }
        }
      }
      BsdfSample bs = bsdfSample(mat, albedo, d, ns, sp.front, etaExt, rng);
      if (!bs.valid) break;
      if ((mat.kind == 4 || mat.kind == 2) && dot(bs.wi, ns) < 0.0f) {
        if (sp.front) {
          if (medDepth < MED) {
            medSigma[medDepth] = ld(mat.absorb);
            medIor[medDepth] = mat.ior;
          }
          medDepth++;
        } else if (medDepth > 0) {
          medDepth--;
        }
      }
      beta = beta * bs.weight;
      specular = bs.isDelta;
      prevPdf = bs.isDelta ? 0.0f : bs.pdf;
      prevX = x;
      V offN = dot(bs.wi, ns) >= 0.0f ? ns : -ns;
      o = offsetRay(x, offN);
      d = bs.wi;
      if (depth >= 4) {
        float q = clampf(maxComp(beta), 0.05f, 0.95f);
        if (rng.rnd() >= q) break;
        beta = beta * (1.0f / q);
      }
      if (maxComp(beta) <= 0.0f) break;
    }
    L = sanitize(L);
    float l3[3] = {L.x, L.y, L.z};
    for (int k = 0; k < 3; k++) A[k] = A[k] + (l3[k] - A[k]) / (s + 1);
    A[3] = 1.0f;
  }
}

template <class F>
int guard(F&& fn) {
  try {
    fn();
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

}  // namespace

extern "C" const char* oracle_last_error(void) { return g_err.c_str(); }

extern "C" void oracle_pcg32(uint64_t seed, uint64_t seq, int n, uint32_t* out) {
  Rng r = Rng::init(seed, seq);
  for (int i = 0; i < n; i++) out[i] = r.next();
}

extern "C" int oracle_render(const sundog_flat_scene* f, const oracle_params* p, float* accum, float* albedo,
                             float* normal, oracle_stats* st) {
  return guard([&] {
    if (!f || !p || !accum || !albedo || !normal) throw std::runtime_error("oracle_render: null argument");
    Scene S;
    setupScene(f, S);
    const int W = f->width, H = f->height;
    const int tx = (W + 15) / 16, ty = (H + 15) / 16;
    const int stride = p->tile_stride > 0 ? p->tile_stride : 1;
    std::vector<int> tiles;
    for (int t = 0; t < tx * ty; t += stride) tiles.push_back(t);
    uint64_t rad = 0, sh = 0, paths = 0;
    int nthreads = p->threads > 0 ? p->threads : omp_get_max_threads();
    auto t0 = std::chrono::steady_clock::now();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads) reduction(+ : rad, sh, paths)
    for (int i = 0; i < (int)tiles.size(); i++) {
      int tile = tiles[i];
      int bx = (tile % tx) * 16, by = (tile / tx) * 16;
      for (int y = by; y < std::min(by + 16, H); y++)
        for (int x = bx; x < std::min(bx + 16, W); x++) {
          uint64_t r = 0, s = 0;
          renderPixel(S, x, y, p->sample_offset, p->sample_offset + p->sample_count, p->brute != 0, accum, albedo,
                      normal, r, s);
          rad += r;
          sh += s;
          paths += (uint64_t)p->sample_count;
        }
    }
    double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (st) {
      st->radiance_rays = rad;
      st->shadow_rays = sh;
      st->paths = paths;
      st->seconds = secs;
      st->threads = nthreads;
    }
  });
}

extern "C" int oracle_trace(const sundog_flat_scene* f, const float* rays, int n, int brute, float* t_out,
                            int32_t* inst_out, int32_t* prim_out) {
  return guard([&] {
    Scene S;
    setupScene(f, S);
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; i++) {
      const float* r = rays + 8 * (size_t)i;
      Hit h;
      bool hit = closestHit(S, mk(r[0], r[1], r[2]), mk(r[4], r[5], r[6]), r[3], r[7], brute != 0, h);
      t_out[i] = hit ? h.t : INFINITY;
      inst_out[i] = hit ? h.inst : -1;
      prim_out[i] = hit ? h.prim : -1;
    }
  });
}

extern "C" int oracle_shadow(const sundog_flat_scene* f, const float* rays, int n, float* tr_out) {
  return guard([&] {
    Scene S;
    setupScene(f, S);
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; i++) {
      const float* r = rays + 8 * (size_t)i;
      V tr = shadowTransmittance(S, mk(r[0], r[1], r[2]), mk(r[4], r[5], r[6]), r[7]);
      tr_out[3 * i] = tr.x;
      tr_out[3 * i + 1] = tr.y;
      tr_out[3 * i + 2] = tr.z;
    }
  });
}

extern "C" int oracle_env_cdf(const sundog_flat_scene* f, float* marg, float* cond) {
  return guard([&] {
    if (f->bg_kind != 2 || !f->env_rgba) throw std::runtime_error("scene has no environment map");
    std::vector<float> m, c;
    buildEnvCdf(f, m, c);
    memcpy(marg, m.data(), m.size() * 4);
    memcpy(cond, c.data(), c.size() * 4);
  });
}

extern "C" int oracle_env_bins(const sundog_flat_scene* f, const float* u, int n, int32_t* row, int32_t* col) {
  return guard([&] {
    if (f->bg_kind != 2 || !f->env_rgba) throw std::runtime_error("scene has no environment map");
    std::vector<float> m, c;
    buildEnvCdf(f, m, c);
    int W = f->env_w, H = f->env_h;
    for (int i = 0; i < n; i++) {
      row[i] = findInterval(m.data(), H + 1, u[2 * i]);
      col[i] = findInterval(c.data() + (size_t)row[i] * (W + 1), W + 1, u[2 * i + 1]);
    }
  });
}

extern "C" int oracle_bsdf(const sundog_flat_material* m, const float albedo[3], const float wo[3],
                           const float wi[3], float eta, float f_out[3], float* pdf_out) {
  V n = mk(0, 0, 1);
  V fv = bsdfEval(*m, ld(albedo), ld(wo), ld(wi), n, eta);
  f_out[0] = fv.x;
  f_out[1] = fv.y;
  f_out[2] = fv.z;
  *pdf_out = bsdfPdf(*m, ld(wo), ld(wi), n, eta);
  return 0;
}

extern "C" int oracle_bsdf_sample(const sundog_flat_material* m, const float albedo[3], const float raydir[3],
                                  int frontface, float eta_ext, uint64_t seed, uint64_t seq, float wi_out[3],
                                  float weight_out[3], float* pdf_out, int* delta_out, int* valid_out,
                                  int* draws_out) {
  Rng rng = Rng::init(seed, seq);
  BsdfSample s = bsdfSample(*m, ld(albedo), ld(raydir), mk(0, 0, 1), frontface != 0, eta_ext, rng);
  wi_out[0] = s.wi.x;
  wi_out[1] = s.wi.y;
  wi_out[2] = s.wi.z;
  weight_out[0] = s.weight.x;
  weight_out[1] = s.weight.y;
  weight_out[2] = s.weight.z;
  *pdf_out = s.pdf;
  *delta_out = s.isDelta;
  *valid_out = s.valid;
  *draws_out = (int)rng.draws;
  return 0;
}

extern "C" float oracle_flame_transmittance(const sundog_flat_scene* f, const float o[3], const float d[3],
                                            float tend, uint64_t seed, uint64_t seq, int skip, int* draws_out) {
  Scene S;
  S.f = f;
  Rng rng = Rng::init(seed, seq);
  float tr = flameTransmittance(S, ld(o), ld(d), tend, rng, skip);
  if (draws_out) *draws_out = (int)rng.draws;
  return tr;
}

// ---------------------------------------------------------------------------
// probes used by the oracle's own tests
// ---------------------------------------------------------------------------

extern "C" int oracle_env_sample(const sundog_flat_scene* f, float u0, float u1, float wi_out[3],
                                 float* pdf_out, float* pdf_back_out) {
  return guard([&] {
    if (f->bg_kind != 2 || !f->env_rgba) throw std::runtime_error("scene has no environment map");
    Scene S;
    S.f = f;
    S.envImportance = buildEnvCdf(f, S.marg, S.cond) && f->env_importance;
    LightSample s = sampleEnvU(S, u0, u1);
    wi_out[0] = s.wi.x;
    wi_out[1] = s.wi.y;
    wi_out[2] = s.wi.z;
    *pdf_out = s.valid ? s.pdf : 0.0f;
    *pdf_back_out = s.valid ? envPdfSolidAngle(S, s.wi) : 0.0f;
  });
}

extern "C" float oracle_march_flames(const sundog_flat_scene* f, const float o[3], const float d[3], float tend,
                                     uint64_t seed, uint64_t seq, float trans_out[3], float L_out[3],
                                     int* draws_out) {
  Scene S;
  S.f = f;
  Rng rng = Rng::init(seed, seq);
  V trans = spl(1.0f);
  V L = marchFlames(S, ld(o), ld(d), tend, rng, trans);
  trans_out[0] = trans.x;
  trans_out[1] = trans.y;
  trans_out[2] = trans.z;
  L_out[0] = L.x;
  L_out[1] = L.y;
  L_out[2] = L.z;
  if (draws_out) *draws_out = (int)rng.draws;
  return trans.x;
}

extern "C" int oracle_texture(const sundog_flat_scene* f, int tex, const float* uv, int n, float* out) {
  return guard([&] {
    if (tex < 0 || tex >= f->num_textures) throw std::runtime_error("oracle_texture: texture id out of range");
    for (int i = 0; i < n; i++) evalTexture(f->textures[tex], uv[2 * i], uv[2 * i + 1], out + 4 * (size_t)i);
  });
}

extern "C" void oracle_bilinear(const float* img, int w, int h, const float* uv, int n, int wrap_v, float* out) {
  for (int i = 0; i < n; i++) bilinear(img, w, h, uv[2 * i], uv[2 * i + 1], wrap_v != 0, out + 4 * (size_t)i);
}

extern "C" void oracle_noise(const float* p, int n, uint32_t seed, int octaves, float* out) {
  for (int i = 0; i < n; i++) {
    V q = ld(p + 3 * (size_t)i);
    out[i] = octaves > 0 ? fbmN(q, seed, octaves) : valueNoise(q, seed);
  }
}

extern "C" void oracle_water_normal(const float x[3], const float ns[3], const float wo[3], float amp, float freq,
                                    int octaves, float out[3]) {
  V n = waterNormal(ld(x), ld(ns), ld(wo), amp, freq, octaves);
  out[0] = n.x;
  out[1] = n.y;
  out[2] = n.z;
}

extern "C" int oracle_flame_field(const sundog_flat_scene* f, int flame, const float* p, int n, float* sigma_out,
                                  float* emission_out) {
  return guard([&] {
    if (flame < 0 || flame >= f->num_flames) throw std::runtime_error("oracle_flame_field: flame id out of range");
    for (int i = 0; i < n; i++) {
      float sg;
      V em;
      flameField(f->flames[flame], ld(p + 3 * (size_t)i), sg, em);
      sigma_out[i] = sg;
      emission_out[3 * i] = em.x;
      emission_out[3 * i + 1] = em.y;
      emission_out[3 * i + 2] = em.z;
    }
  });
}

extern "C" int oracle_mesh_light_sample(const sundog_flat_scene* f, int light, const float x[3], const float* u,
                                        int n, float* wi_out, float* dist_out, float* pdf_out, int32_t* tri_out,
                                        float* area_out) {
  return guard([&] {
    if (light < 0 || light >= f->num_lights) throw std::runtime_error("oracle_mesh_light_sample: light id out of range");
    Scene S;
    setupScene(f, S);
    if (S.meshLight[light].mesh < 0) throw std::runtime_error("oracle_mesh_light_sample: not a mesh light");
    *area_out = S.meshLight[light].area;
    for (int i = 0; i < n; i++) {
      int tri = -1;
      LightSample s = sampleMeshLightU(S, light, ld(x), u[3 * i], u[3 * i + 1], u[3 * i + 2], &tri);
      wi_out[3 * i] = s.wi.x;
      wi_out[3 * i + 1] = s.wi.y;
      wi_out[3 * i + 2] = s.wi.z;
      dist_out[i] = s.dist;
      pdf_out[i] = s.valid ? s.pdf : 0.0f;
      tri_out[i] = tri;
    }
  });
}

// Light samples of light `light` from n shading points xs (n x 3), sample i
// drawing from the stream init(seed0 + i, seq) (device/light_sample.cuh:24-188):
// direction, distance, radiance, pdf (0 when invalid), flags (bit 0 valid,
// bit 1 delta), the draws taken, and the MIS-side pdf that
// lightPdfSolidAngle (device/light_sample.cuh:197-226) returns for the
// sampled point x + wi * dist (NaN for mesh and delta lights, whose MIS side
// needs the hit normal or does not exist).
extern "C" int oracle_light_sample(const sundog_flat_scene* f, int light, const float* xs, int n, uint64_t seed0,
                                   uint64_t seq, float* wi_out, float* dist_out, float* li_out, float* pdf_out,
                                   int32_t* flags_out, int32_t* draws_out, float* pdf_back_out) {
  return guard([&] {
    if (light < 0 || light >= f->num_lights) throw std::runtime_error("oracle_light_sample: light id out of range");
    Scene S;
    setupScene(f, S);
    const sundog_flat_light& lt = f->lights[light];
    for (int i = 0; i < n; i++) {
      Rng rng = Rng::init(seed0 + (uint64_t)i, seq);
      V x = ld(xs + 3 * i);
      LightSample s = sampleLight(S, light, x, rng);
      wi_out[3 * i] = s.wi.x;
      wi_out[3 * i + 1] = s.wi.y;
      wi_out[3 * i + 2] = s.wi.z;
      dist_out[i] = s.dist;
      li_out[3 * i] = s.Li.x;
      li_out[3 * i + 1] = s.Li.y;
      li_out[3 * i + 2] = s.Li.z;
      pdf_out[i] = s.valid ? s.pdf : 0.0f;
      flags_out[i] = (s.valid ? 1 : 0) | (s.isDelta ? 2 : 0);
      draws_out[i] = (int32_t)rng.draws;
      float back = NAN;
      if (s.valid && !s.isDelta && lt.kind != 5) back = lightPdfSolidAngle(S, light, x, x + s.wi * s.dist, spl(0.0f));
      pdf_back_out[i] = back;
    }
  });
}

// lightPdfSolidAngle (device/light_sample.cuh:197-226) for n (from, hit
// point, hit normal) triples (n x 9 floats).
extern "C" int oracle_light_pdf(const sundog_flat_scene* f, int light, const float* q, int n, float* pdf_out) {
  return guard([&] {
    if (light < 0 || light >= f->num_lights) throw std::runtime_error("oracle_light_pdf: light id out of range");
    Scene S;
    setupScene(f, S);
    for (int i = 0; i < n; i++)
      pdf_out[i] = lightPdfSolidAngle(S, light, ld(q + 9 * i), ld(q + 9 * i + 3), ld(q + 9 * i + 6));
  });
}

// The camera-ray jitter of sample s of pixel (px, py), by the path loop's
// own cameraJitter.
extern "C" int oracle_camera_jitter(const sundog_flat_scene* f, int px, int py, int s, float* jx, float* jy) {
  return guard([&] {
    if (px < 0 || py < 0 || px >= f->width || py >= f->height || s < 0)
      throw std::runtime_error("oracle_camera_jitter: pixel or sample out of range");
    const int pixel = py * f->width + px;
    Rng rng = Rng::init(((uint64_t)pixel << 32) ^ (uint64_t)s, (uint64_t)f->seed);
    cameraJitter(s, strataPerAxis(f->spp), rng, *jx, *jy);
  });
}
