// Small vector / affine math shared by the host scene builder and the gfx950
// kernels. Plain structs (no HIP vector types) so host-only translation units
// compile with g++ as well as hipcc.
//
// Formulas follow the reference's conventions where results feed the image:
// offsetRay (device/math.cuh:151-154), refract / schlick (:52-70), the Duff
// et al. 2017 orthonormal basis (:72-83), the row-major 3x4 affine and its
// composition order (:86-116), detSign3 (:123-125).
#pragma once

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define SD_FN __host__ __device__ __forceinline__
#else
#define SD_FN inline
#endif

namespace sdk {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 0.31830988618379067154f;

struct vec3 {
  float x, y, z;
};
struct vec2 {
  float x, y;
};

SD_FN vec3 v3(float s) { return vec3{s, s, s}; }
SD_FN vec3 v3(float x, float y, float z) { return vec3{x, y, z}; }
SD_FN vec3 operator+(vec3 a, vec3 b) { return vec3{a.x + b.x, a.y + b.y, a.z + b.z}; }
SD_FN vec3 operator-(vec3 a, vec3 b) { return vec3{a.x - b.x, a.y - b.y, a.z - b.z}; }
SD_FN vec3 operator-(vec3 a) { return vec3{-a.x, -a.y, -a.z}; }
SD_FN vec3 operator*(vec3 a, vec3 b) { return vec3{a.x * b.x, a.y * b.y, a.z * b.z}; }
SD_FN vec3 operator*(vec3 a, float s) { return vec3{a.x * s, a.y * s, a.z * s}; }
SD_FN vec3 operator*(float s, vec3 a) { return vec3{a.x * s, a.y * s, a.z * s}; }
SD_FN vec3 operator/(vec3 a, float s) { return a * (1.0f / s); }
SD_FN vec3& operator+=(vec3& a, vec3 b) { a = a + b; return a; }
SD_FN vec3& operator*=(vec3& a, vec3 b) { a = a * b; return a; }
SD_FN vec3& operator*=(vec3& a, float s) { a = a * s; return a; }
SD_FN vec2 operator*(vec2 a, float s) { return vec2{a.x * s, a.y * s}; }
SD_FN vec2 operator+(vec2 a, vec2 b) { return vec2{a.x + b.x, a.y + b.y}; }

SD_FN float dot(vec3 a, vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
SD_FN vec3 cross(vec3 a, vec3 b) {
  return vec3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
SD_FN float len2(vec3 a) { return dot(a, a); }
SD_FN float length(vec3 a) { return sqrtf(dot(a, a)); }
SD_FN vec3 normalize(vec3 a) { return a * (1.0f / sqrtf(dot(a, a))); }
SD_FN float clampf(float x, float lo, float hi) { return fminf(hi, fmaxf(lo, x)); }
SD_FN vec3 vmin(vec3 a, vec3 b) { return vec3{fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)}; }
SD_FN vec3 vmax(vec3 a, vec3 b) { return vec3{fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)}; }
SD_FN float maxc(vec3 v) { return fmaxf(v.x, fmaxf(v.y, v.z)); }
SD_FN vec3 lerp(vec3 a, vec3 b, float t) { return a + (b - a) * t; }
SD_FN vec3 vexp(vec3 a) { return vec3{expf(a.x), expf(a.y), expf(a.z)}; }
SD_FN vec3 reflect(vec3 v, vec3 n) { return v - 2.0f * dot(v, n) * n; }
SD_FN float comp(vec3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }

// Snell refraction; v toward the surface, n against v.
SD_FN bool refract(vec3 v, vec3 n, float eta, vec3& out) {
  float cosi = -dot(v, n);
  float k = 1.0f - eta * eta * (1.0f - cosi * cosi);
  if (k < 0.0f) return false;
  out = eta * v + (eta * cosi - sqrtf(k)) * n;
  return true;
}

SD_FN float schlick(float c, float f0) {
  float m = clampf(1.0f - c, 0.0f, 1.0f);
  float m2 = m * m;
  return f0 + (1.0f - f0) * m2 * m2 * m;
}
SD_FN vec3 schlick3(float c, vec3 f0) {
  float m = clampf(1.0f - c, 0.0f, 1.0f);
  float m2 = m * m;
  return f0 + (v3(1.0f) - f0) * (m2 * m2 * m);
}

// Orthonormal basis around a unit normal (Duff et al. 2017, "Building an
// Orthonormal Basis, Revisited").
struct Frame {
  vec3 t, b, n;
  SD_FN explicit Frame(vec3 nn) : n(nn) {
    float s = copysignf(1.0f, nn.z);
    float a = -1.0f / (s + nn.z);
    float c = nn.x * nn.y * a;
    t = vec3{1.0f + s * nn.x * nn.x * a, s * c, -s * nn.x};
    b = vec3{c, s + nn.y * nn.y * a, -nn.y};
  }
  SD_FN vec3 toWorld(vec3 v) const { return v.x * t + v.y * b + v.z * n; }
  SD_FN vec3 toLocal(vec3 v) const { return vec3{dot(v, t), dot(v, b), dot(v, n)}; }
};

// Row-major 3x4 affine (object -> world when stored per instance).
struct Affine {
  float m[12];
  SD_FN static Affine identity() { return Affine{{1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0}}; }
  SD_FN vec3 point(vec3 p) const {
    return vec3{m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3],
                m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7],
                m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11]};
  }
  SD_FN vec3 vector(vec3 v) const {
    return vec3{m[0] * v.x + m[1] * v.y + m[2] * v.z,
                m[4] * v.x + m[5] * v.y + m[6] * v.z,
                m[8] * v.x + m[9] * v.y + m[10] * v.z};
  }
};

// a∘b (b first), same summation order as the reference composition.
SD_FN Affine compose(const Affine& a, const Affine& b) {
  Affine r{};
  for (int i = 0; i < 3; i++) {
    for (int j = 0; j < 3; j++)
      r.m[i * 4 + j] = a.m[i * 4 + 0] * b.m[j] + a.m[i * 4 + 1] * b.m[4 + j] +
                       a.m[i * 4 + 2] * b.m[8 + j];
    r.m[i * 4 + 3] = a.m[i * 4 + 0] * b.m[3] + a.m[i * 4 + 1] * b.m[7] +
                     a.m[i * 4 + 2] * b.m[11] + a.m[i * 4 + 3];
  }
  return r;
}

SD_FN float detSign3(vec3 ux, vec3 uy, vec3 uz) {
  return dot(uy, cross(uz, ux)) < 0.0f ? -1.0f : 1.0f;
}

// Self-intersection-safe origin (n faces the outgoing hemisphere).
SD_FN vec3 offsetRay(vec3 p, vec3 n) {
  float s = 1e-4f * (1.0f + fmaxf(fabsf(p.x), fmaxf(fabsf(p.y), fabsf(p.z))));
  return p + n * s;
}

}  // namespace sdk
