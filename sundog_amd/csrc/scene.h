// Host-side scene model filled by the scene-construction ABI. GPU-free.
//
// Field meanings and defaults restate the reference's scene model
// (src/scene.h:11-119) and descriptor layouts (device/params.h:12-114); the
// structures themselves are this library's own and are lowered to the
// device tables in render_host.cpp.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "kinds.h"
#include "sd_math.h"

namespace sdk {

// reference src/scene.h:11-21
struct RenderSettings {
  int width = 1280, height = 720;
  int spp = 64;
  int maxDepth = 16;
  float clampVal = 50.0f;
  unsigned seed = 7;
  float exposure = 0.0f;
  bool transparentShadows = true;
  int chunk = 16;
  bool denoise = false;
};

// reference src/scene.h:23-30
struct CameraSettings {
  vec3 lookfrom{0, 2, 8}, lookat{0, 0, 0}, up{0, 1, 0};
  float vfov = 40.0f;
  float aperture = 0.0f;
  float focusDist = 0.0f;
};

struct Texture {
  int kind = TX_SOLID;
  vec3 a{0, 0, 0}, b{0, 0, 0};
  vec2 scale{8.0f, 8.0f};
  float width = 0.05f;
  std::string file;
  bool srgb = true;
};

struct Material {
  int kind = MT_LAMBERT;
  int texId = -1;
  vec3 color{0.8f, 0.8f, 0.8f};
  float roughness = 0.0f;
  float ior = 1.5f;
  float intensity = 1.0f;
  int twoSided = 0;
  vec3 absorb{0, 0, 0};
  float waveAmp = 0.0f;
  float waveFreq = 0.0f;
  int waveOctaves = 3;
};

struct Light {
  int kind = LT_POINT;
  vec3 p{0, 0, 0}, u{0, 0, 0}, v{0, 0, 0}, n{0, 0, 0};
  vec3 L{0, 0, 0};
  float radius = 0.0f;
  float area = 0.0f;
  vec3 dir{0, 0, 0};
  int twoSided = 0;
  int texId = -1;
  int flameId = -1;
  int objectId = -1;
};

struct Flame {
  vec3 base{0, 0, 0};
  float height = 0, radius = 0, intensity = 20.0f, sigma = 4.0f, noiseScale = 3.0f;
  unsigned seed = 0;
  int octaves = 3;
  float step = 0.0f;   // <= 0: 32 fixed steps
  int bounds = 0;      // 0 cylinder, 1 box
};

struct MeshRef {
  std::string objFile;
  std::string usemtl;
  bool smoothNormals = true;
};

struct PhysicsBody {
  bool enabled = false, dynamic = false;
  float density = 250.0f;
  vec3 velocity{0, 0, 0}, angularVelocity{0, 0, 0};
  float friction = -1.0f, restitution = -1.0f;
  float thickness = 0.2f;
};

struct PhysicsSettings {
  bool enabled = false;
  vec3 gravity{0.0f, -9.81f, 0.0f};
  float timestep = 1.0f / 240.0f;
  float maxTime = 15.0f;
  float stopTime = 0.0f;
  float friction = 0.6f, restitution = 0.1f;
  int posIters = 8, velIters = 2;
  float sleepThreshold = -1.0f;
};

struct Object {
  int geomKind = GK_SPHERE;
  int meshId = -1;
  uint16_t matFront = kMatNone, matBack = kMatNone;
  int cutoutTexId = -1;
  Affine xform = Affine::identity();
  int lightId = -1;
  bool nee = true;
  PhysicsBody physics;
};

struct EnvSettings {
  std::string file;
  float rotateDeg = 0.0f;
  float intensity = 1.0f;
  bool importance = true;
};

struct Scene {
  RenderSettings render;
  CameraSettings camera;
  int bgKind = BG_SOLID;
  vec3 bgA{0, 0, 0}, bgB{0, 0, 0};
  EnvSettings env;
  std::vector<Texture> textures;
  std::vector<Material> materials;
  std::vector<MeshRef> meshes;
  std::vector<Object> objects;
  std::vector<Light> lights;
  std::vector<Flame> flames;
  PhysicsSettings physics;
  std::string baseDir;
};

// Device-facing pinhole / thin-lens camera (reference src/scene_build.cpp:13-29).
struct Camera {
  vec3 origin, lowerLeft, horizontal, vertical, u, v, w;
  float lensRadius;
};

Camera makeCamera(const CameraSettings& cs, int width, int height);

struct XformStep {
  int kind;
  float a, b, c;
};
Affine composeTransform(const XformStep* steps, int n);

[[noreturn]] void sceneFail(const std::string& msg);
void addObjectDerived(Scene& s, Object& o);
void addFlameDerived(Scene& s, const Flame& f, float lightIntensity);

inline bool materialTransmissive(const Scene& s, uint16_t m) {
  return m != kMatNone &&
         (s.materials[m].kind == MT_DIELECTRIC || s.materials[m].kind == MT_WATER);
}

}  // namespace sdk

// Opaque ABI handle: the scene plus the call-phase state machine
// (reference src/capi_internal.h:36-42).
struct sundog_scene {
  sdk::Scene scene;
  int phase = 0;
  bool hasCamera = false;
};
