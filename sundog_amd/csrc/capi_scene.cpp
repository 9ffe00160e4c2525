// Scene-construction half of the C ABI. GPU-free: no HIP call is made in
// this translation unit, so the CPU test suite exercises all of it.
//
// The observable behaviour is the reference's (src/sundog_api.h:30-179):
// defaults and sentinels (src/capi_scene.cpp:33-458, src/scene.h:11-28), the
// config -> objects -> flames -> lights call order (src/capi_internal.h:36-42),
// the thread-local error message (src/capi_scene.cpp:16-21), and the
// derived state — composed transforms, NEE light registration and light
// frames, the two flame lights, the camera (src/scene_build.cpp:13-181).
// The organisation is this library's own: every entry point runs through
// one `entry` runner (handle check, call order, error capture), and what
// varies between calls is data — per-kind material rules (defaults and the
// value checks each kind applies), collider rules, one emitter rule per
// shape (the light frame it derives), and the two flame-light rules.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/sundog_api.h"
#include "scene.h"

using namespace sdk;

namespace {

thread_local std::string t_lastError;

// ---- sentinels at the boundary: NaN or NULL = unset; doubles are narrowed
// to float exactly once, here ------------------------------------------------
bool given(double v) { return !std::isnan(v); }
float num(double v, float fallback) { return given(v) ? (float)v : fallback; }
vec3 vec(const double* p, vec3 fallback) { return p ? vec3{(float)p[0], (float)p[1], (float)p[2]} : fallback; }
bool nonNegative(vec3 v) { return v.x >= 0.0f && v.y >= 0.0f && v.z >= 0.0f; }

// ---- call order ---------------------------------------------------------------
enum Phase : int { kConfig = 0, kObjects = 1, kFlames = 2, kLights = 3 };

// Runs one entry point on handle h: the call may not come after a later
// phase's calls; the phase then advances to `phase`. Any failure becomes
// the entry point's error value and the thread's last-error message.
template <class R, class F>
R entry(R onError, sundog_scene* h, Phase phase, const char* what, F&& body) {
  try {
    if (!h) throw std::runtime_error("scene: null scene handle");
    if (h->phase > phase)
      throw std::runtime_error(std::string("scene: ") + what +
                               " must come before objects/flames/lights "
                               "(call order: config, objects, flames, lights)");
    h->phase = std::max(h->phase, (int)phase);
    return body(h->scene);
  } catch (const std::exception& e) {
    t_lastError = e.what();
  } catch (...) {
    t_lastError = "unknown error";
  }
  return onError;
}

// The same for calls that take no handle or ignore the call order.
template <class R, class F>
R guarded(R onError, F&& body) {
  try {
    return body();
  } catch (const std::exception& e) {
    t_lastError = e.what();
  } catch (...) {
    t_lastError = "unknown error";
  }
  return onError;
}

int registerId(size_t count) { return (int)count; }

// ---- textures -------------------------------------------------------------------
int addTexture(Scene& s, const Texture& t) {
  s.textures.push_back(t);
  return registerId(s.textures.size() - 1);
}
vec2 pair(const double* p, vec2 fallback) { return p ? vec2{(float)p[0], (float)p[1]} : fallback; }

// ---- materials ----------------------------------------------------------------
// Optional inputs of one add_material call (unset = the kind's default).
struct MaterialArgs {
  const double* color = nullptr;
  const double* absorb = nullptr;
  int texId = -1;
  double roughness = NAN, ior = NAN, intensity = NAN, waveAmp = NAN, waveFreq = NAN;
  int twoSided = -1;
};

enum MaterialCheck : unsigned {
  kIntensityNonNeg = 1u,
  kAbsorbNonNeg = 2u,
  kWavesNonNeg = 4u,
};

struct MaterialRule {
  int kind;
  const char* name;
  bool textured;      // accepts a texture id
  float roughness, ior;
  vec3 color, absorb;
  float waveAmp, waveFreq;
  unsigned checks;
};

// Per-kind defaults and value checks (reference src/capi_scene.cpp:197-314).
const MaterialRule kMaterialRules[] = {
    {MT_LAMBERT, "lambert", true, 0.0f, 1.5f, {0.8f, 0.8f, 0.8f}, {0, 0, 0}, 0.0f, 0.0f, 0u},
    {MT_METAL, "metal", true, 0.0f, 1.5f, {0.8f, 0.8f, 0.8f}, {0, 0, 0}, 0.0f, 0.0f, 0u},
    {MT_DIELECTRIC, "dielectric", true, 0.0f, 1.5f, {0.8f, 0.8f, 0.8f}, {0, 0, 0}, 0.0f, 0.0f, kAbsorbNonNeg},
    {MT_EMISSIVE, "emissive", true, 0.0f, 1.5f, {0.8f, 0.8f, 0.8f}, {0, 0, 0}, 0.0f, 0.0f, kIntensityNonNeg},
    {MT_WATER, "water", false, 0.0f, 1.33f, {1.0f, 1.0f, 1.0f}, {0.45f, 0.08f, 0.035f}, 0.05f, 2.0f,
     kAbsorbNonNeg | kWavesNonNeg},
    {MT_PLASTIC, "plastic", true, 0.15f, 1.5f, {0.8f, 0.8f, 0.8f}, {0, 0, 0}, 0.0f, 0.0f, 0u},
};

const MaterialRule& materialRule(int kind) {
  for (const MaterialRule& r : kMaterialRules)
    if (r.kind == kind) return r;
  throw std::runtime_error("scene: unknown material kind");
}

[[noreturn]] void failFor(const char* who, const char* what) {
  throw std::runtime_error(std::string("scene: ") + who + ": " + what);
}

int addMaterial(Scene& s, int kind, const MaterialArgs& a) {
  const MaterialRule& r = materialRule(kind);
  Material m;
  m.kind = kind;
  m.texId = r.textured ? a.texId : -1;
  m.color = vec(a.color, r.color);
  m.roughness = num(a.roughness, r.roughness);
  m.ior = num(a.ior, r.ior);
  m.intensity = num(a.intensity, 1.0f);
  m.twoSided = a.twoSided == 1 ? 1 : 0;
  m.absorb = vec(a.absorb, r.absorb);
  m.waveAmp = num(a.waveAmp, r.waveAmp);
  m.waveFreq = num(a.waveFreq, r.waveFreq);
  if ((r.checks & kIntensityNonNeg) && m.intensity < 0.0f) failFor(r.name, "intensity must be >= 0");
  if ((r.checks & kWavesNonNeg) && (m.waveAmp < 0.0f || m.waveFreq < 0.0f))
    failFor(r.name, "wave_amp and wave_freq must be >= 0");
  if ((r.checks & kAbsorbNonNeg) && !nonNegative(m.absorb)) failFor(r.name, "absorb components must be >= 0");
  if (m.texId != -1 && (m.texId < 0 || m.texId >= (int)s.textures.size()))
    sceneFail("material: texture id out of range");
  if (s.materials.size() >= kMatNone) sceneFail("too many materials");
  s.materials.push_back(m);
  return registerId(s.materials.size() - 1);
}

// ---- objects: colliders ----------------------------------------------------------
const char* const kShapeNames[] = {"sphere", "rect", "disk", "cylinder", "parabola", "mesh"};

// Physics opt-in checks, in order; the first that fails names the problem.
struct ColliderRule {
  bool (*breaks)(const Scene&, const Object&);
  const char* message;  // empty: the message names the shape
};
const ColliderRule kColliderRules[] = {
    {[](const Scene& s, const Object&) { return !s.physics.enabled; },
     "physics: object opts in but the scene has no top-level physics block"},
    {[](const Scene&, const Object& o) {
       return o.geomKind == GK_DISK || o.geomKind == GK_CYLINDER || o.geomKind == GK_PARABOLA;
     },
     ""},
    {[](const Scene&, const Object& o) { return o.geomKind == GK_RECT && o.physics.dynamic; },
     "physics: rect colliders are static-only (zero-thickness plates tunnel)"},
    {[](const Scene&, const Object& o) { return o.physics.dynamic && o.physics.density <= 0.0f; },
     "physics: dynamic body needs positive density"},
    {[](const Scene&, const Object& o) { return o.physics.thickness <= 0.0f; },
     "physics: thickness must be positive"},
};

void checkCollider(const Scene& s, const Object& o) {
  if (!o.physics.enabled) return;
  for (const ColliderRule& r : kColliderRules) {
    if (!r.breaks(s, o)) continue;
    if (*r.message) sceneFail(r.message);
    throw std::runtime_error(std::string("scene: physics: shape '") + kShapeNames[o.geomKind] +
                             "' is not supported as a collider (use sphere, rect, or mesh)");
  }
}

// ---- objects: emitters registered as NEE lights -----------------------------
// The object's transformed unit axes and the sign of their determinant
// (mirroring transforms flip the emitting side back to the canonical front).
struct Axes {
  vec3 x, y, z;
  float handed;
};

bool sameLength(float a, float b, float tol) { return fabsf(a - b) <= tol; }

// One rule per emitter shape: the light kind, whether a texture may tint
// it, and how the light frame follows from the object's axes.
struct EmitterRule {
  int geom;
  int lightKind;
  bool texturable;
  void (*frame)(const Axes&, Light&);
};

void planarFrame(const Axes& a, Light& l) {
  l.u = a.x;
  l.v = a.z;
  l.n = normalize(cross(a.z, a.x)) * a.handed;
}

const EmitterRule kEmitterRules[] = {
    {GK_RECT, LT_RECT, true,
     [](const Axes& a, Light& l) {
       planarFrame(a, l);
       l.area = 4.0f * length(cross(a.x, a.z));  // the canonical rect spans [-1, 1]^2
     }},
    {GK_DISK, LT_DISK, true,
     [](const Axes& a, Light& l) {
       const float ru = length(a.x), rz = length(a.z);
       if (!sameLength(ru, rz, 1e-3f * fmaxf(ru, rz)))
         sceneFail("disk NEE light requires uniform XZ scale (pdf would be wrong); set nee:false");
       planarFrame(a, l);
       l.area = kPi * ru * rz;
     }},
    {GK_SPHERE, LT_SPHERE, false,
     [](const Axes& a, Light& l) {
       const float rx = length(a.x), ry = length(a.y), rz = length(a.z);
       if (!sameLength(rx, ry, 1e-3f * rx) || !sameLength(rx, rz, 1e-3f * rx))
         sceneFail("sphere NEE light requires uniform scale; set nee:false");
       l.radius = rx;
     }},
    {GK_MESH, LT_MESH, true, [](const Axes&, Light&) {}},  // geometry patched when the mesh loads
};

const EmitterRule* emitterRule(int geom) {
  for (const EmitterRule& r : kEmitterRules)
    if (r.geom == geom) return &r;
  return nullptr;
}

// The light an emissive object contributes, or none. Every check runs
// before anything is added, so a rejected object leaves no light behind.
bool emitterLight(const Scene& s, const Object& o, Light& l) {
  if (!o.nee || o.matFront == kMatNone) return false;
  const Material& m = s.materials[o.matFront];
  if (m.kind != MT_EMISSIVE) return false;
  const EmitterRule* rule = emitterRule(o.geomKind);
  if (!rule) return false;
  if (o.physics.dynamic)
    sceneFail("physics: a dynamic object cannot be an NEE area light (the light "
              "frame is baked at parse time); set nee:false or keep it static");
  if (o.cutoutTexId >= 0)
    sceneFail("an alpha-cutout object cannot be an NEE area light (NEE would "
              "sample emission inside the holes); set nee:false");
  if (m.texId >= 0 && !rule->texturable)
    sceneFail("textured emissive sphere cannot be an NEE light (uv frame is "
              "object-space); use a solid color or set nee:false");
  Axes a{o.xform.vector(v3(1, 0, 0)), o.xform.vector(v3(0, 1, 0)), o.xform.vector(v3(0, 0, 1)), 0.0f};
  a.handed = detSign3(a.x, a.y, a.z);
  l = Light{};
  l.kind = rule->lightKind;
  l.p = o.xform.point(v3(0, 0, 0));
  l.twoSided = m.twoSided;
  // a textured emitter's radiance is intensity x texel, untextured colour x intensity
  l.texId = m.texId;
  l.L = m.texId >= 0 ? v3(m.intensity) : m.color * m.intensity;
  l.objectId = (int)s.objects.size();
  rule->frame(a, l);
  if (l.kind != LT_MESH) l.objectId = -1;
  return true;
}

// ---- flames: two derived soft point lights ----------------------------------------
struct FlameLightRule {
  float share;       // fraction of the flame's light intensity
  float heightFrac;  // height above the base, as a fraction of the flame's height
  vec3 warmth;       // colour
};
const FlameLightRule kFlameLights[2] = {
    {0.65f, 0.35f, {1.0f, 0.50f, 0.15f}},
    {0.35f, 0.70f, {1.0f, 0.62f, 0.25f}},
};

// ---- the transform steps ----------------------------------------------------------
Affine stepMatrix(const XformStep& st) {
  if (st.kind == SUNDOG_XF_SCALE) return Affine{{st.a, 0, 0, 0, 0, st.b, 0, 0, 0, 0, st.c, 0}};
  if (st.kind == SUNDOG_XF_TRANSLATE) return Affine{{1, 0, 0, st.a, 0, 1, 0, st.b, 0, 0, 1, st.c}};
  if (st.kind < SUNDOG_XF_ROTATE_X || st.kind > SUNDOG_XF_ROTATE_Z) sceneFail("unknown transform step kind");
  const float rad = st.a * (kPi / 180.0f);
  const float c = cosf(rad), s = sinf(rad);
  // rotation by rad about the named axis: the other two axes turn in their plane
  const int axis = st.kind - SUNDOG_XF_ROTATE_X;  // 0 x, 1 y, 2 z
  const int i = (axis + 1) % 3, j = (axis + 2) % 3;
  Affine m = Affine::identity();
  m.m[i * 4 + i] = c;
  m.m[j * 4 + j] = c;
  m.m[i * 4 + j] = -s;
  m.m[j * 4 + i] = s;
  return m;
}

}  // namespace

namespace sdk {

void setLastError(const std::string& msg) { t_lastError = msg; }

void sceneFail(const std::string& msg) { throw std::runtime_error("scene: " + msg); }

// Pinhole / thin-lens camera: the image plane sits at the focus distance
// (the look distance unless set), spanned by horizontal and vertical from
// its lower-left corner.
Camera makeCamera(const CameraSettings& cs, int width, int height) {
  Camera c{};
  const float halfH = tanf(cs.vfov * kPi / 180.0f / 2.0f);
  const float halfW = ((float)width / (float)height) * halfH;
  const float focus = cs.focusDist > 0.0f ? cs.focusDist : length(cs.lookfrom - cs.lookat);
  c.w = normalize(cs.lookfrom - cs.lookat);
  c.u = normalize(cross(cs.up, c.w));
  c.v = cross(c.w, c.u);
  c.origin = cs.lookfrom;
  c.lowerLeft = c.origin - halfW * focus * c.u - halfH * focus * c.v - focus * c.w;
  c.horizontal = 2.0f * halfW * focus * c.u;
  c.vertical = 2.0f * halfH * focus * c.v;
  c.lensRadius = cs.aperture / 2.0f;
  return c;
}

// Steps apply in order: each later step wraps the earlier ones.
Affine composeTransform(const XformStep* steps, int n) {
  Affine m = Affine::identity();
  for (int i = 0; i < n; i++) m = compose(stepMatrix(steps[i]), m);
  return m;
}

void addObjectDerived(Scene& s, Object& o) {
  if (o.matFront == kMatNone && o.matBack == kMatNone) sceneFail("object has no material on either face");
  checkCollider(s, o);
  Light l;
  if (emitterLight(s, o, l)) {
    o.lightId = (int)s.lights.size();
    s.lights.push_back(l);
  }
  s.objects.push_back(o);
}

void addFlameDerived(Scene& s, const Flame& f, float lightIntensity) {
  if (f.height <= 0.0f || f.radius <= 0.0f) sceneFail("flame: height and radius must be positive");
  if (f.intensity < 0.0f || f.sigma < 0.0f) sceneFail("flame: intensity and sigma must be >= 0");
  for (const FlameLightRule& r : kFlameLights) {
    Light l;
    l.kind = LT_POINT;
    l.flameId = (int)s.flames.size();
    l.p = f.base + v3(0.0f, r.heightFrac * f.height, 0.0f);
    l.radius = 0.3f * f.radius;
    l.L = r.warmth * (lightIntensity * r.share);
    s.lights.push_back(l);
  }
  s.flames.push_back(f);
}

}  // namespace sdk

// ---------------------------------------------------------------------------
// extern "C" surface
// ---------------------------------------------------------------------------

extern "C" const char* sundog_last_error(void) { return t_lastError.c_str(); }

extern "C" sundog_scene* sundog_scene_create(const char* base_dir) {
  return guarded<sundog_scene*>(nullptr, [&] {
    sundog_scene* h = new sundog_scene();
    h->scene.baseDir = (base_dir && *base_dir) ? base_dir : ".";
    return h;
  });
}

extern "C" void sundog_scene_destroy(sundog_scene* h) { delete h; }

extern "C" int sundog_set_render(sundog_scene* h, int width, int height, int spp, int max_depth, double clamp,
                                 int64_t seed, double exposure, int transparent_shadows) {
  return entry(SUNDOG_ERROR, h, kConfig, "render settings", [&](Scene& s) {
    RenderSettings& r = s.render;
    // non-positive counts, a negative seed, NaN and -1 keep the current value
    if (width > 0) r.width = width;
    if (height > 0) r.height = height;
    if (spp > 0) r.spp = spp;
    if (max_depth > 0) r.maxDepth = max_depth;
    r.clampVal = num(clamp, r.clampVal);
    if (seed >= 0) r.seed = (unsigned)seed;
    r.exposure = num(exposure, r.exposure);
    if (transparent_shadows >= 0) r.transparentShadows = transparent_shadows == 1;
    return SUNDOG_OK;
  });
}

extern "C" int sundog_set_camera(sundog_scene* h, const double lookfrom[3], const double lookat[3],
                                 const double up[3], double vfov, double aperture, double focus_dist) {
  return entry(SUNDOG_ERROR, h, kConfig, "camera", [&](Scene& s) {
    if (!lookfrom || !lookat) sceneFail("camera: lookfrom and lookat are required");
    CameraSettings& c = s.camera;
    c.lookfrom = vec(lookfrom, c.lookfrom);
    c.lookat = vec(lookat, c.lookat);
    c.up = vec(up, c.up);
    c.vfov = num(vfov, 40.0f);
    c.aperture = num(aperture, 0.0f);
    c.focusDist = num(focus_dist, 0.0f);
    h->hasCamera = true;
    return SUNDOG_OK;
  });
}

extern "C" int sundog_set_background_solid(sundog_scene* h, const double color[3]) {
  return entry(SUNDOG_ERROR, h, kConfig, "background", [&](Scene& s) {
    if (!color) sceneFail("background: color is required");
    s.bgKind = BG_SOLID;
    s.bgA = vec(color, v3(0.0f));
    return SUNDOG_OK;
  });
}

extern "C" int sundog_set_background_gradient(sundog_scene* h, const double horizon[3], const double zenith[3]) {
  return entry(SUNDOG_ERROR, h, kConfig, "background", [&](Scene& s) {
    if (!horizon || !zenith) sceneFail("background: horizon and zenith are required");
    s.bgKind = BG_GRADIENT;
    s.bgA = vec(horizon, v3(0.0f));
    s.bgB = vec(zenith, v3(0.0f));
    return SUNDOG_OK;
  });
}

extern "C" int sundog_set_background_envmap(sundog_scene* h, const char* file, double rotate_deg, double intensity,
                                            int importance) {
  return entry(SUNDOG_ERROR, h, kConfig, "background", [&](Scene& s) {
    if (!file || !*file) sceneFail("envmap: file is required");
    EnvSettings e;
    e.file = file;
    e.rotateDeg = num(rotate_deg, 0.0f);
    e.intensity = num(intensity, 1.0f);
    e.importance = importance != 0;
    s.bgKind = BG_ENVMAP;
    s.env = e;
    if (e.intensity < 0.0f) sceneFail("envmap: intensity must be >= 0");
    return SUNDOG_OK;
  });
}

extern "C" int sundog_set_physics(sundog_scene* h, const double gravity[3], double timestep, double max_time,
                                  double friction, double restitution, int pos_iters, int vel_iters,
                                  double sleep_threshold, double stop_time) {
  return entry(SUNDOG_ERROR, h, kConfig, "physics settings", [&](Scene& s) {
    PhysicsSettings& p = s.physics;
    p.enabled = true;
    p.gravity = vec(gravity, p.gravity);
    p.timestep = num(timestep, p.timestep);
    p.maxTime = num(max_time, p.maxTime);
    p.friction = num(friction, p.friction);
    p.restitution = num(restitution, p.restitution);
    if (pos_iters >= 0) p.posIters = pos_iters;
    if (vel_iters >= 0) p.velIters = vel_iters;
    p.sleepThreshold = num(sleep_threshold, p.sleepThreshold);
    p.stopTime = num(stop_time, p.stopTime);
    if (p.timestep <= 0.0f || p.maxTime <= 0.0f) sceneFail("physics: timestep and max_time must be positive");
    if (p.stopTime < 0.0f) sceneFail("physics: stop_time must be >= 0 (0 = settle to sleep)");
    return SUNDOG_OK;
  });
}

extern "C" int sundog_add_texture_solid(sundog_scene* h, const double color[3]) {
  return entry(-1, h, kConfig, "textures", [&](Scene& s) {
    if (!color) sceneFail("texture: color is required");
    Texture t;
    t.kind = TX_SOLID;
    t.a = vec(color, v3(0.0f));
    return addTexture(s, t);
  });
}

extern "C" int sundog_add_texture_checker(sundog_scene* h, const double a[3], const double b[3],
                                          const double scale[2]) {
  return entry(-1, h, kConfig, "textures", [&](Scene& s) {
    if (!a || !b) sceneFail("texture: a and b are required");
    Texture t;
    t.kind = TX_CHECKER;
    t.a = vec(a, v3(0.0f));
    t.b = vec(b, v3(0.0f));
    t.scale = pair(scale, t.scale);
    return addTexture(s, t);
  });
}

extern "C" int sundog_add_texture_grid(sundog_scene* h, const double a[3], const double b[3],
                                       const double scale[2], double width) {
  return entry(-1, h, kConfig, "textures", [&](Scene& s) {
    if (!a || !b) sceneFail("texture: a and b are required");
    Texture t;
    t.kind = TX_GRID;
    t.a = vec(a, v3(0.0f));
    t.b = vec(b, v3(0.0f));
    t.scale = pair(scale, t.scale);
    t.width = num(width, t.width);
    return addTexture(s, t);
  });
}

extern "C" int sundog_add_texture_image(sundog_scene* h, const char* file, int srgb) {
  return entry(-1, h, kConfig, "textures", [&](Scene& s) {
    if (!file || !*file) sceneFail("texture: file is required");
    Texture t;
    t.kind = TX_IMAGE;
    t.file = file;
    t.srgb = srgb != 0;
    return addTexture(s, t);
  });
}

extern "C" int sundog_add_material_lambert(sundog_scene* h, const double color[3], int tex_id) {
  MaterialArgs a;
  a.color = color;
  a.texId = tex_id;
  return entry(-1, h, kConfig, "materials", [&](Scene& s) { return addMaterial(s, MT_LAMBERT, a); });
}

extern "C" int sundog_add_material_metal(sundog_scene* h, const double color[3], int tex_id, double roughness) {
  MaterialArgs a;
  a.color = color;
  a.texId = tex_id;
  a.roughness = roughness;
  return entry(-1, h, kConfig, "materials", [&](Scene& s) { return addMaterial(s, MT_METAL, a); });
}

extern "C" int sundog_add_material_dielectric(sundog_scene* h, double ior, const double absorb[3],
                                              const double color[3], int tex_id, double roughness) {
  MaterialArgs a;
  a.color = color;
  a.texId = tex_id;
  a.ior = ior;
  a.absorb = absorb;
  a.roughness = roughness;
  return entry(-1, h, kConfig, "materials", [&](Scene& s) { return addMaterial(s, MT_DIELECTRIC, a); });
}

extern "C" int sundog_add_material_emissive(sundog_scene* h, const double color[3], int tex_id, double intensity,
                                            int two_sided) {
  MaterialArgs a;
  a.color = color;
  a.texId = tex_id;
  a.intensity = intensity;
  a.twoSided = two_sided;
  return entry(-1, h, kConfig, "materials", [&](Scene& s) { return addMaterial(s, MT_EMISSIVE, a); });
}

extern "C" int sundog_add_material_water(sundog_scene* h, double ior, const double absorb[3], double wave_amp,
                                         double wave_freq, const double color[3]) {
  MaterialArgs a;
  a.color = color;
  a.ior = ior;
  a.absorb = absorb;
  a.waveAmp = wave_amp;
  a.waveFreq = wave_freq;
  return entry(-1, h, kConfig, "materials", [&](Scene& s) { return addMaterial(s, MT_WATER, a); });
}

extern "C" int sundog_add_material_plastic(sundog_scene* h, const double color[3], int tex_id, double roughness) {
  MaterialArgs a;
  a.color = color;
  a.texId = tex_id;
  a.roughness = roughness;
  return entry(-1, h, kConfig, "materials", [&](Scene& s) { return addMaterial(s, MT_PLASTIC, a); });
}

extern "C" int sundog_add_mesh(sundog_scene* h, const char* obj_file, int smooth_normals, const char* usemtl) {
  return entry(-1, h, kConfig, "meshes", [&](Scene& s) {
    if (!obj_file || !*obj_file) sceneFail("mesh: obj file is required");
    MeshRef m;
    m.objFile = obj_file;
    m.usemtl = usemtl ? usemtl : "";
    m.smoothNormals = smooth_normals != 0;
    s.meshes.push_back(m);
    return registerId(s.meshes.size() - 1);
  });
}

// A face's material: an id, or the pass-through sentinel.
static uint16_t faceMaterial(const Scene& s, int id) {
  if (id == SUNDOG_MAT_NONE) return kMatNone;
  if (id < 0 || id >= (int)s.materials.size()) sceneFail("object: material id out of range");
  return (uint16_t)id;
}

static void readBody(const sundog_physics_body& b, PhysicsBody& p) {
  p.enabled = true;
  p.dynamic = b.dynamic == 1;
  p.density = num(b.density, p.density);
  if (b.has_velocity) p.velocity = vec(b.velocity, p.velocity);
  if (b.has_angular_velocity) p.angularVelocity = vec(b.angular_velocity, p.angularVelocity);
  p.friction = num(b.friction, p.friction);
  p.restitution = num(b.restitution, p.restitution);
  p.thickness = num(b.thickness, p.thickness);
}

extern "C" int sundog_add_object(sundog_scene* h, int geom_kind, int mesh_id, int mat_front, int mat_back,
                                 int cutout_tex_id, const sundog_xform_step* steps, int num_steps, int nee,
                                 const sundog_physics_body* physics) {
  return entry(-1, h, kObjects, "objects", [&](Scene& s) {
    if (geom_kind < 0 || geom_kind > GK_MESH) sceneFail("object: unknown shape");
    Object o;
    o.geomKind = geom_kind;
    if (geom_kind == GK_MESH) {
      if (mesh_id < 0 || mesh_id >= (int)s.meshes.size()) sceneFail("object: mesh id out of range");
      o.meshId = mesh_id;
    }
    o.matFront = faceMaterial(s, mat_front);
    o.matBack = mat_back == SUNDOG_MAT_DEFAULT ? o.matFront : faceMaterial(s, mat_back);
    if (cutout_tex_id != -1) {
      if (cutout_tex_id < 0 || cutout_tex_id >= (int)s.textures.size())
        sceneFail("object: cutout texture id out of range");
      o.cutoutTexId = cutout_tex_id;
    }
    if (num_steps > 0 && !steps) sceneFail("object: transform steps missing");
    std::vector<XformStep> xs;
    for (int i = 0; i < num_steps; i++)
      xs.push_back(XformStep{steps[i].kind, (float)steps[i].a, (float)steps[i].b, (float)steps[i].c});
    o.xform = composeTransform(xs.data(), (int)xs.size());
    o.nee = nee != 0;
    if (physics) readBody(*physics, o.physics);
    const int id = registerId(s.objects.size());
    addObjectDerived(s, o);
    return id;
  });
}

extern "C" int sundog_add_flame(sundog_scene* h, const double base[3], double height, double radius,
                                double intensity, double sigma, double noise_scale, int64_t seed,
                                double light_intensity) {
  return entry(-1, h, kFlames, "flames", [&](Scene& s) {
    if (!base) sceneFail("flame: base is required");
    Flame f;
    f.base = vec(base, v3(0.0f));
    f.height = (float)height;
    f.radius = (float)radius;
    f.intensity = num(intensity, f.intensity);
    f.sigma = num(sigma, f.sigma);
    f.noiseScale = num(noise_scale, f.noiseScale);
    f.seed = seed >= 0 ? (unsigned)seed : 0u;
    const int id = registerId(s.flames.size());
    addFlameDerived(s, f, num(light_intensity, 12.0f));
    return id;
  });
}

// An explicit light after validating its colour-like term.
static int addExplicitLight(Scene& s, Light l, const char* who, const char* term) {
  if (!nonNegative(l.L)) failFor(who, (std::string(term) + " components must be >= 0").c_str());
  s.lights.push_back(l);
  return registerId(s.lights.size() - 1);
}

extern "C" int sundog_add_point_light(sundog_scene* h, const double position[3], const double intensity[3],
                                      double radius) {
  return entry(-1, h, kLights, "lights", [&](Scene& s) {
    if (!position || !intensity) sceneFail("point light: position and intensity are required");
    Light l;
    l.kind = LT_POINT;
    l.p = vec(position, v3(0.0f));
    l.radius = num(radius, 0.0f);
    l.L = vec(intensity, v3(0.0f));
    return addExplicitLight(s, l, "point light", "intensity");
  });
}

extern "C" int sundog_add_distant_light(sundog_scene* h, const double direction[3], const double radiance[3]) {
  return entry(-1, h, kLights, "lights", [&](Scene& s) {
    if (!direction || !radiance) sceneFail("distant light: direction and radiance are required");
    Light l;
    l.kind = LT_DISTANT;
    l.dir = normalize(vec(direction, v3(0.0f)));
    l.L = vec(radiance, v3(0.0f));
    return addExplicitLight(s, l, "distant light", "radiance");
  });
}

extern "C" int sundog_scene_validate(sundog_scene* h) {
  return guarded(SUNDOG_ERROR, [&] {
    if (!h) sceneFail("null scene handle");
    if (!h->hasCamera) sceneFail("missing camera");
    return SUNDOG_OK;
  });
}

extern "C" int sundog_set_flame_march(sundog_scene* h, int flame_id, int octaves, double step, int bounds) {
  return guarded(SUNDOG_ERROR, [&] {
    if (!h) sceneFail("null scene handle");
    if (flame_id < 0 || flame_id >= (int)h->scene.flames.size()) sceneFail("flame march: flame id out of range");
    if (octaves < 1 || octaves > 8) sceneFail("flame march: octaves must be in [1, 8]");
    if (bounds != SUNDOG_FLAME_BOUNDS_CYLINDER && bounds != SUNDOG_FLAME_BOUNDS_BOX)
      sceneFail("flame march: unknown bounds kind");
    Flame& f = h->scene.flames[flame_id];
    f.octaves = octaves;
    f.step = std::isnan(step) ? 0.0f : (float)step;
    f.bounds = bounds;
    return SUNDOG_OK;
  });
}

extern "C" int sundog_set_water_octaves(sundog_scene* h, int material_id, int octaves) {
  return guarded(SUNDOG_ERROR, [&] {
    if (!h) sceneFail("null scene handle");
    if (material_id < 0 || material_id >= (int)h->scene.materials.size())
      sceneFail("water octaves: material id out of range");
    if (h->scene.materials[material_id].kind != MT_WATER) sceneFail("water octaves: material is not water");
    if (octaves < 1 || octaves > 8) sceneFail("water octaves: octaves must be in [1, 8]");
    h->scene.materials[material_id].waveOctaves = octaves;
    return SUNDOG_OK;
  });
}

// ---------------------------------------------------------------------------
// sundog_scene_describe: JSON dump of the derived scene state (GPU-free)
// ---------------------------------------------------------------------------

namespace {

struct Json {
  std::string s;
  void key(const char* k) { s += "\""; s += k; s += "\":"; }
  void num(double v) {
    char b[64];
    if (std::isnan(v)) snprintf(b, sizeof b, "null");
    else snprintf(b, sizeof b, "%.9g", v);
    s += b;
  }
  void v3(const vec3& v) {
    s += "[";
    num(v.x); s += ","; num(v.y); s += ","; num(v.z);
    s += "]";
  }
  void str(const std::string& v) {
    s += "\"";
    for (char c : v) {
      if (c == '"' || c == '\\') { s += '\\'; s += c; }
      else if ((unsigned char)c < 0x20) s += ' ';
      else s += c;
    }
    s += "\"";
  }
};

}  // namespace

extern "C" int sundog_scene_describe(sundog_scene* h, char* buf, size_t cap, size_t* needed) {
  return guarded(SUNDOG_ERROR, [&] {
    if (!h) sceneFail("null scene handle");
    const Scene& S = h->scene;
    Json j;
    j.s += "{";
    j.key("phase"); j.num(h->phase); j.s += ",";
    j.key("has_camera"); j.s += h->hasCamera ? "true" : "false"; j.s += ",";
    j.key("base_dir"); j.str(S.baseDir); j.s += ",";
    const RenderSettings& r = S.render;
    j.key("render"); j.s += "{";
    j.key("width"); j.num(r.width); j.s += ",";
    j.key("height"); j.num(r.height); j.s += ",";
    j.key("spp"); j.num(r.spp); j.s += ",";
    j.key("max_depth"); j.num(r.maxDepth); j.s += ",";
    j.key("clamp"); j.num(r.clampVal); j.s += ",";
    j.key("seed"); j.num(r.seed); j.s += ",";
    j.key("exposure"); j.num(r.exposure); j.s += ",";
    j.key("transparent_shadows"); j.s += r.transparentShadows ? "true" : "false"; j.s += ",";
    j.key("chunk"); j.num(r.chunk); j.s += ",";
    j.key("denoise"); j.s += r.denoise ? "true" : "false";
    j.s += "},";
    const CameraSettings& c = S.camera;
    j.key("camera"); j.s += "{";
    j.key("lookfrom"); j.v3(c.lookfrom); j.s += ",";
    j.key("lookat"); j.v3(c.lookat); j.s += ",";
    j.key("up"); j.v3(c.up); j.s += ",";
    j.key("vfov"); j.num(c.vfov); j.s += ",";
    j.key("aperture"); j.num(c.aperture); j.s += ",";
    j.key("focus_dist"); j.num(c.focusDist);
    if (h->hasCamera) {
      Camera d = makeCamera(c, r.width, r.height);
      j.s += ",";
      j.key("origin"); j.v3(d.origin); j.s += ",";
      j.key("lower_left"); j.v3(d.lowerLeft); j.s += ",";
      j.key("horizontal"); j.v3(d.horizontal); j.s += ",";
      j.key("vertical"); j.v3(d.vertical); j.s += ",";
      j.key("u"); j.v3(d.u); j.s += ",";
      j.key("v"); j.v3(d.v); j.s += ",";
      j.key("w"); j.v3(d.w); j.s += ",";
      j.key("lens_radius"); j.num(d.lensRadius);
    }
    j.s += "},";
    j.key("background"); j.s += "{";
    j.key("kind"); j.num(S.bgKind); j.s += ",";
    j.key("a"); j.v3(S.bgA); j.s += ",";
    j.key("b"); j.v3(S.bgB); j.s += ",";
    j.key("env_file"); j.str(S.env.file); j.s += ",";
    j.key("env_rotate_deg"); j.num(S.env.rotateDeg); j.s += ",";
    j.key("env_intensity"); j.num(S.env.intensity); j.s += ",";
    j.key("env_importance"); j.s += S.env.importance ? "true" : "false";
    j.s += "},";
    j.key("physics"); j.s += "{";
    j.key("enabled"); j.s += S.physics.enabled ? "true" : "false"; j.s += ",";
    j.key("gravity"); j.v3(S.physics.gravity); j.s += ",";
    j.key("timestep"); j.num(S.physics.timestep); j.s += ",";
    j.key("max_time"); j.num(S.physics.maxTime); j.s += ",";
    j.key("stop_time"); j.num(S.physics.stopTime); j.s += ",";
    j.key("friction"); j.num(S.physics.friction); j.s += ",";
    j.key("restitution"); j.num(S.physics.restitution); j.s += ",";
    j.key("pos_iters"); j.num(S.physics.posIters); j.s += ",";
    j.key("vel_iters"); j.num(S.physics.velIters);
    j.s += "},";
    j.key("textures"); j.s += "[";
    for (size_t i = 0; i < S.textures.size(); i++) {
      const Texture& t = S.textures[i];
      if (i) j.s += ",";
      j.s += "{";
      j.key("kind"); j.num(t.kind); j.s += ",";
      j.key("a"); j.v3(t.a); j.s += ",";
      j.key("b"); j.v3(t.b); j.s += ",";
      j.key("scale"); j.s += "["; j.num(t.scale.x); j.s += ","; j.num(t.scale.y); j.s += "],";
      j.key("width"); j.num(t.width); j.s += ",";
      j.key("file"); j.str(t.file); j.s += ",";
      j.key("srgb"); j.s += t.srgb ? "true" : "false";
      j.s += "}";
    }
    j.s += "],";
    j.key("materials"); j.s += "[";
    for (size_t i = 0; i < S.materials.size(); i++) {
      const Material& m = S.materials[i];
      if (i) j.s += ",";
      j.s += "{";
      j.key("kind"); j.num(m.kind); j.s += ",";
      j.key("tex_id"); j.num(m.texId); j.s += ",";
      j.key("color"); j.v3(m.color); j.s += ",";
      j.key("roughness"); j.num(m.roughness); j.s += ",";
      j.key("ior"); j.num(m.ior); j.s += ",";
      j.key("intensity"); j.num(m.intensity); j.s += ",";
      j.key("two_sided"); j.num(m.twoSided); j.s += ",";
      j.key("absorb"); j.v3(m.absorb); j.s += ",";
      j.key("wave_amp"); j.num(m.waveAmp); j.s += ",";
      j.key("wave_freq"); j.num(m.waveFreq); j.s += ",";
      j.key("wave_octaves"); j.num(m.waveOctaves);
      j.s += "}";
    }
    j.s += "],";
    j.key("meshes"); j.s += "[";
    for (size_t i = 0; i < S.meshes.size(); i++) {
      if (i) j.s += ",";
      j.s += "{";
      j.key("obj"); j.str(S.meshes[i].objFile); j.s += ",";
      j.key("usemtl"); j.str(S.meshes[i].usemtl); j.s += ",";
      j.key("smooth"); j.s += S.meshes[i].smoothNormals ? "true" : "false";
      j.s += "}";
    }
    j.s += "],";
    j.key("objects"); j.s += "[";
    for (size_t i = 0; i < S.objects.size(); i++) {
      const Object& o = S.objects[i];
      if (i) j.s += ",";
      j.s += "{";
      j.key("geom_kind"); j.num(o.geomKind); j.s += ",";
      j.key("mesh_id"); j.num(o.meshId); j.s += ",";
      j.key("mat_front"); j.num(o.matFront); j.s += ",";
      j.key("mat_back"); j.num(o.matBack); j.s += ",";
      j.key("cutout_tex"); j.num(o.cutoutTexId); j.s += ",";
      j.key("light_id"); j.num(o.lightId); j.s += ",";
      j.key("nee"); j.s += o.nee ? "true" : "false"; j.s += ",";
      j.key("physics"); j.s += o.physics.enabled ? (o.physics.dynamic ? "\"dynamic\"" : "\"static\"") : "null";
      j.s += ",";
      j.key("xform"); j.s += "[";
      for (int k = 0; k < 12; k++) {
        if (k) j.s += ",";
        j.num(o.xform.m[k]);
      }
      j.s += "]}";
    }
    j.s += "],";
    j.key("lights"); j.s += "[";
    for (size_t i = 0; i < S.lights.size(); i++) {
      const Light& l = S.lights[i];
      if (i) j.s += ",";
      j.s += "{";
      j.key("kind"); j.num(l.kind); j.s += ",";
      j.key("p"); j.v3(l.p); j.s += ",";
      j.key("u"); j.v3(l.u); j.s += ",";
      j.key("v"); j.v3(l.v); j.s += ",";
      j.key("n"); j.v3(l.n); j.s += ",";
      j.key("L"); j.v3(l.L); j.s += ",";
      j.key("radius"); j.num(l.radius); j.s += ",";
      j.key("area"); j.num(l.area); j.s += ",";
      j.key("dir"); j.v3(l.dir); j.s += ",";
      j.key("two_sided"); j.num(l.twoSided); j.s += ",";
      j.key("tex_id"); j.num(l.texId); j.s += ",";
      j.key("flame_id"); j.num(l.flameId); j.s += ",";
      j.key("object_id"); j.num(l.objectId);
      j.s += "}";
    }
    j.s += "],";
    j.key("flames"); j.s += "[";
    for (size_t i = 0; i < S.flames.size(); i++) {
      const Flame& f = S.flames[i];
      if (i) j.s += ",";
      j.s += "{";
      j.key("base"); j.v3(f.base); j.s += ",";
      j.key("height"); j.num(f.height); j.s += ",";
      j.key("radius"); j.num(f.radius); j.s += ",";
      j.key("intensity"); j.num(f.intensity); j.s += ",";
      j.key("sigma"); j.num(f.sigma); j.s += ",";
      j.key("noise_scale"); j.num(f.noiseScale); j.s += ",";
      j.key("seed"); j.num(f.seed); j.s += ",";
      j.key("octaves"); j.num(f.octaves); j.s += ",";
      j.key("step"); j.num(f.step); j.s += ",";
      j.key("bounds"); j.num(f.bounds);
      j.s += "}";
    }
    j.s += "]}";
    if (needed) *needed = j.s.size() + 1;
    if (buf && cap > 0) {
      size_t n = std::min(cap - 1, j.s.size());
      memcpy(buf, j.s.data(), n);
      buf[n] = '\0';
    }
    return SUNDOG_OK;
  });
}
