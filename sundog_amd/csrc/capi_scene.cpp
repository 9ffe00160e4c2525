// Scene-construction half of the C ABI. GPU-free: no HIP call is made in
// this translation unit, so it is exercised by the CPU test suite.
//
// Behaviour restated from the reference:
//   defaults and sentinel handling  src/capi_scene.cpp:33-458, src/scene.h:11-28
//   phase state machine             src/capi_internal.h:36-42
//   error model (thread-local msg)  src/capi_scene.cpp:16-21, src/capi_internal.h:46-56
//   derivation (transforms, NEE light registration, flame lights, camera)
//                                   src/scene_build.cpp:13-181
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

bool isset(double v) { return !std::isnan(v); }
float nf(double v, float dflt) { return std::isnan(v) ? dflt : (float)v; }
vec3 nf3(const double* p, vec3 dflt) {
  return p ? vec3{(float)p[0], (float)p[1], (float)p[2]} : dflt;
}

void phaseAdvance(sundog_scene* h, int needed, const char* what) {
  if (!h) throw std::runtime_error("scene: null scene handle");
  if (h->phase > needed)
    throw std::runtime_error(std::string("scene: ") + what +
                             " must come before objects/flames/lights "
                             "(call order: config, objects, flames, lights)");
  h->phase = needed;
}

template <class F>
int guardInt(int errval, F&& f) {
  try {
    return f();
  } catch (const std::exception& e) {
    t_lastError = e.what();
    return errval;
  } catch (...) {
    t_lastError = "unknown error";
    return errval;
  }
}

}  // namespace

namespace sdk {

void setLastError(const std::string& msg) { t_lastError = msg; }

void sceneFail(const std::string& msg) { throw std::runtime_error("scene: " + msg); }

Camera makeCamera(const CameraSettings& cs, int width, int height) {
  Camera c{};
  float aspect = (float)width / (float)height;
  float theta = cs.vfov * kPi / 180.0f;
  float halfH = tanf(theta / 2.0f);
  float halfW = aspect * halfH;
  float focus = cs.focusDist > 0.0f ? cs.focusDist : length(cs.lookfrom - cs.lookat);
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

static Affine stepMatrix(const XformStep& st) {
  const float d2r = kPi / 180.0f;
  switch (st.kind) {
    case SUNDOG_XF_SCALE:
      return Affine{{st.a, 0, 0, 0, 0, st.b, 0, 0, 0, 0, st.c, 0}};
    case SUNDOG_XF_TRANSLATE:
      return Affine{{1, 0, 0, st.a, 0, 1, 0, st.b, 0, 0, 1, st.c}};
    case SUNDOG_XF_ROTATE_X: {
      float c = cosf(st.a * d2r), s = sinf(st.a * d2r);
      return Affine{{1, 0, 0, 0, 0, c, -s, 0, 0, s, c, 0}};
    }
    case SUNDOG_XF_ROTATE_Y: {
      float c = cosf(st.a * d2r), s = sinf(st.a * d2r);
      return Affine{{c, 0, s, 0, 0, 1, 0, 0, -s, 0, c, 0}};
    }
    case SUNDOG_XF_ROTATE_Z: {
      float c = cosf(st.a * d2r), s = sinf(st.a * d2r);
      return Affine{{c, -s, 0, 0, s, c, 0, 0, 0, 0, 1, 0}};
    }
    default:
      sceneFail("unknown transform step kind");
  }
}

Affine composeTransform(const XformStep* steps, int n) {
  Affine m = Affine::identity();
  for (int i = 0; i < n; i++) m = compose(stepMatrix(steps[i]), m);  // later wraps earlier
  return m;
}

static const char* shapeName(int k) {
  switch (k) {
    case GK_SPHERE: return "sphere";
    case GK_RECT: return "rect";
    case GK_DISK: return "disk";
    case GK_CYLINDER: return "cylinder";
    case GK_PARABOLA: return "parabola";
    default: return "mesh";
  }
}

void addObjectDerived(Scene& s, Object& o) {
  if (o.matFront == kMatNone && o.matBack == kMatNone)
    sceneFail("object has no material on either face");
  if (o.physics.enabled) {
    if (!s.physics.enabled)
      sceneFail("physics: object opts in but the scene has no top-level physics block");
    if (o.geomKind == GK_DISK || o.geomKind == GK_CYLINDER || o.geomKind == GK_PARABOLA)
      sceneFail(std::string("physics: shape '") + shapeName(o.geomKind) +
                "' is not supported as a collider (use sphere, rect, or mesh)");
    if (o.geomKind == GK_RECT && o.physics.dynamic)
      sceneFail("physics: rect colliders are static-only (zero-thickness plates tunnel)");
    if (o.physics.dynamic && o.physics.density <= 0.0f)
      sceneFail("physics: dynamic body needs positive density");
    if (o.physics.thickness <= 0.0f) sceneFail("physics: thickness must be positive");
  }

  uint16_t em = kMatNone;
  if (o.matFront != kMatNone && s.materials[o.matFront].kind == MT_EMISSIVE) em = o.matFront;
  bool registers = em != kMatNone && o.nee &&
                   (o.geomKind == GK_RECT || o.geomKind == GK_DISK ||
                    o.geomKind == GK_SPHERE || o.geomKind == GK_MESH);
  if (o.physics.dynamic && registers)
    sceneFail("physics: a dynamic object cannot be an NEE area light (the light "
              "frame is baked at parse time); set nee:false or keep it static");
  if (o.cutoutTexId >= 0 && registers)
    sceneFail("an alpha-cutout object cannot be an NEE area light (NEE would "
              "sample emission inside the holes); set nee:false");
  if (registers) {
    const Material& m = s.materials[em];
    Light l;
    l.texId = -1;
    l.flameId = -1;
    l.p = o.xform.point(v3(0, 0, 0));
    vec3 ux = o.xform.vector(v3(1, 0, 0));
    vec3 uy = o.xform.vector(v3(0, 1, 0));
    vec3 uz = o.xform.vector(v3(0, 0, 1));
    float ds = detSign3(ux, uy, uz);
    l.L = m.color * m.intensity;
    l.twoSided = m.twoSided;
    if (m.texId >= 0) {
      if (o.geomKind == GK_SPHERE)
        sceneFail("textured emissive sphere cannot be an NEE light (uv frame is "
                  "object-space); use a solid color or set nee:false");
      l.texId = m.texId;
      l.L = v3(m.intensity);
    }
    if (o.geomKind == GK_RECT) {
      l.kind = LT_RECT;
      l.u = ux;
      l.v = uz;
      l.n = normalize(cross(uz, ux)) * ds;
      l.area = 4.0f * length(cross(ux, uz));
    } else if (o.geomKind == GK_DISK) {
      float ru = length(ux), rz = length(uz);
      if (fabsf(ru - rz) > 1e-3f * fmaxf(ru, rz))
        sceneFail("disk NEE light requires uniform XZ scale (pdf would be wrong); set nee:false");
      l.kind = LT_DISK;
      l.u = ux;
      l.v = uz;
      l.n = normalize(cross(uz, ux)) * ds;
      l.area = kPi * ru * rz;
    } else if (o.geomKind == GK_SPHERE) {
      float rx = length(ux), rz = length(uz), ry = length(uy);
      if (fabsf(rx - ry) > 1e-3f * rx || fabsf(rx - rz) > 1e-3f * rx)
        sceneFail("sphere NEE light requires uniform scale; set nee:false");
      l.kind = LT_SPHERE;
      l.radius = rx;
    } else {
      l.kind = LT_MESH;  // geometry patched when the mesh is loaded
      l.objectId = (int)s.objects.size();
    }
    o.lightId = (int)s.lights.size();
    s.lights.push_back(l);
  }
  s.objects.push_back(o);
}

void addFlameDerived(Scene& s, const Flame& f, float lightIntensity) {
  if (f.height <= 0.0f || f.radius <= 0.0f)
    sceneFail("flame: height and radius must be positive");
  if (f.intensity < 0.0f || f.sigma < 0.0f)
    sceneFail("flame: intensity and sigma must be >= 0");
  const float frac[2] = {0.65f, 0.35f};
  const float hfac[2] = {0.35f, 0.70f};
  const vec3 warm[2] = {v3(1.0f, 0.50f, 0.15f), v3(1.0f, 0.62f, 0.25f)};
  for (int k = 0; k < 2; k++) {
    Light l;
    l.texId = -1;
    l.flameId = (int)s.flames.size();
    l.kind = LT_POINT;
    l.p = f.base + v3(0.0f, hfac[k] * f.height, 0.0f);
    l.radius = 0.3f * f.radius;
    l.L = warm[k] * (lightIntensity * frac[k]);
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
  try {
    sundog_scene* h = new sundog_scene();
    h->scene.baseDir = (base_dir && *base_dir) ? base_dir : ".";
    return h;
  } catch (const std::exception& e) {
    t_lastError = e.what();
    return nullptr;
  }
}

extern "C" void sundog_scene_destroy(sundog_scene* h) { delete h; }

extern "C" int sundog_set_render(sundog_scene* h, int width, int height, int spp,
                                 int max_depth, double clamp, int64_t seed,
                                 double exposure, int transparent_shadows) {
  return guardInt(SUNDOG_ERROR, [&] {
    phaseAdvance(h, 0, "render settings");
    RenderSettings& r = h->scene.render;
    if (width > 0) r.width = width;
    if (height > 0) r.height = height;
    if (spp > 0) r.spp = spp;
    if (max_depth > 0) r.maxDepth = max_depth;
    if (isset(clamp)) r.clampVal = (float)clamp;
    if (seed >= 0) r.seed = (unsigned)seed;
    if (isset(exposure)) r.exposure = (float)exposure;
    if (transparent_shadows >= 0) r.transparentShadows = transparent_shadows == 1;
    return SUNDOG_OK;
  });
}

extern "C" int sundog_set_camera(sundog_scene* h, const double lookfrom[3],
                                 const double lookat[3], const double up[3],
                                 double vfov, double aperture, double focus_dist) {
  return guardInt(SUNDOG_ERROR, [&] {
    phaseAdvance(h, 0, "camera");
    if (!lookfrom || !lookat) sceneFail("camera: lookfrom and lookat are required");
    CameraSettings& c = h->scene.camera;
    c.lookfrom = nf3(lookfrom, c.lookfrom);
    c.lookat = nf3(lookat, c.lookat);
    c.up = nf3(up, c.up);
    c.vfov = nf(vfov, 40.0f);
    c.aperture = nf(aperture, 0.0f);
    c.focusDist = nf(focus_dist, 0.0f);
    h->hasCamera = true;
    return SUNDOG_OK;
  });
}

extern "C" int sundog_set_background_solid(sundog_scene* h, const double color[3]) {
  return guardInt(SUNDOG_ERROR, [&] {
    phaseAdvance(h, 0, "background");
    if (!color) sceneFail("background: color is required");
    h->scene.bgKind = BG_SOLID;
    h->scene.bgA = nf3(color, v3(0.0f));
    return SUNDOG_OK;
  });
}

extern "C" int sundog_set_background_gradient(sundog_scene* h, const double horizon[3],
                                              const double zenith[3]) {
  return guardInt(SUNDOG_ERROR, [&] {
    phaseAdvance(h, 0, "background");
    if (!horizon || !zenith) sceneFail("background: horizon and zenith are required");
    h->scene.bgKind = BG_GRADIENT;
    h->scene.bgA = nf3(horizon, v3(0.0f));
    h->scene.bgB = nf3(zenith, v3(0.0f));
    return SUNDOG_OK;
  });
}

extern "C" int sundog_set_background_envmap(sundog_scene* h, const char* file,
                                            double rotate_deg, double intensity,
                                            int importance) {
  return guardInt(SUNDOG_ERROR, [&] {
    phaseAdvance(h, 0, "background");
    if (!file || !*file) sceneFail("envmap: file is required");
    h->scene.bgKind = BG_ENVMAP;
    h->scene.env.file = file;
    h->scene.env.rotateDeg = nf(rotate_deg, 0.0f);
    h->scene.env.intensity = nf(intensity, 1.0f);
    h->scene.env.importance = importance != 0;
    if (h->scene.env.intensity < 0.0f) sceneFail("envmap: intensity must be >= 0");
    return SUNDOG_OK;
  });
}

extern "C" int sundog_set_physics(sundog_scene* h, const double gravity[3], double timestep,
                                  double max_time, double friction, double restitution,
                                  int pos_iters, int vel_iters, double sleep_threshold,
                                  double stop_time) {
  return guardInt(SUNDOG_ERROR, [&] {
    phaseAdvance(h, 0, "physics settings");
    PhysicsSettings& p = h->scene.physics;
    p.enabled = true;
    p.gravity = nf3(gravity, p.gravity);
    p.timestep = nf(timestep, p.timestep);
    p.maxTime = nf(max_time, p.maxTime);
    p.friction = nf(friction, p.friction);
    p.restitution = nf(restitution, p.restitution);
    if (pos_iters >= 0) p.posIters = pos_iters;
    if (vel_iters >= 0) p.velIters = vel_iters;
    p.sleepThreshold = nf(sleep_threshold, p.sleepThreshold);
    p.stopTime = nf(stop_time, p.stopTime);
    if (p.timestep <= 0.0f || p.maxTime <= 0.0f)
      sceneFail("physics: timestep and max_time must be positive");
    if (p.stopTime < 0.0f) sceneFail("physics: stop_time must be >= 0 (0 = settle to sleep)");
    return SUNDOG_OK;
  });
}

static int pushTexture(sundog_scene* h, const Texture& t) {
  phaseAdvance(h, 0, "textures");
  int id = (int)h->scene.textures.size();
  h->scene.textures.push_back(t);
  return id;
}

extern "C" int sundog_add_texture_solid(sundog_scene* h, const double color[3]) {
  return guardInt(-1, [&] {
    if (!color) sceneFail("texture: color is required");
    Texture t;
    t.kind = TX_SOLID;
    t.a = nf3(color, v3(0.0f));
    return pushTexture(h, t);
  });
}

extern "C" int sundog_add_texture_checker(sundog_scene* h, const double a[3],
                                          const double b[3], const double scale[2]) {
  return guardInt(-1, [&] {
    if (!a || !b) sceneFail("texture: a and b are required");
    Texture t;
    t.kind = TX_CHECKER;
    t.a = nf3(a, v3(0.0f));
    t.b = nf3(b, v3(0.0f));
    t.scale = scale ? vec2{(float)scale[0], (float)scale[1]} : vec2{8.0f, 8.0f};
    return pushTexture(h, t);
  });
}

extern "C" int sundog_add_texture_grid(sundog_scene* h, const double a[3], const double b[3],
                                       const double scale[2], double width) {
  return guardInt(-1, [&] {
    if (!a || !b) sceneFail("texture: a and b are required");
    Texture t;
    t.kind = TX_GRID;
    t.a = nf3(a, v3(0.0f));
    t.b = nf3(b, v3(0.0f));
    t.scale = scale ? vec2{(float)scale[0], (float)scale[1]} : vec2{8.0f, 8.0f};
    t.width = nf(width, 0.05f);
    return pushTexture(h, t);
  });
}

extern "C" int sundog_add_texture_image(sundog_scene* h, const char* file, int srgb) {
  return guardInt(-1, [&] {
    if (!file || !*file) sceneFail("texture: file is required");
    Texture t;
    t.kind = TX_IMAGE;
    t.file = file;
    t.srgb = srgb != 0;
    return pushTexture(h, t);
  });
}

static Material baseMaterial() { return Material{}; }

static int pushMaterial(sundog_scene* h, const Material& m, int texId) {
  phaseAdvance(h, 0, "materials");
  if (texId != -1 && (texId < 0 || texId >= (int)h->scene.textures.size()))
    sceneFail("material: texture id out of range");
  if (h->scene.materials.size() >= kMatNone) sceneFail("too many materials");
  int id = (int)h->scene.materials.size();
  h->scene.materials.push_back(m);
  return id;
}

extern "C" int sundog_add_material_lambert(sundog_scene* h, const double color[3], int tex_id) {
  return guardInt(-1, [&] {
    Material m = baseMaterial();
    m.kind = MT_LAMBERT;
    m.color = nf3(color, m.color);
    m.texId = tex_id;
    return pushMaterial(h, m, tex_id);
  });
}

extern "C" int sundog_add_material_metal(sundog_scene* h, const double color[3], int tex_id,
                                         double roughness) {
  return guardInt(-1, [&] {
    Material m = baseMaterial();
    m.kind = MT_METAL;
    m.color = nf3(color, m.color);
    m.texId = tex_id;
    m.roughness = nf(roughness, m.roughness);
    return pushMaterial(h, m, tex_id);
  });
}

extern "C" int sundog_add_material_dielectric(sundog_scene* h, double ior,
                                              const double absorb[3], const double color[3],
                                              int tex_id, double roughness) {
  return guardInt(-1, [&] {
    Material m = baseMaterial();
    m.kind = MT_DIELECTRIC;
    m.color = nf3(color, m.color);
    m.texId = tex_id;
    m.ior = nf(ior, m.ior);
    m.roughness = nf(roughness, m.roughness);
    if (absorb) {
      m.absorb = nf3(absorb, v3(0.0f));
      if (m.absorb.x < 0.0f || m.absorb.y < 0.0f || m.absorb.z < 0.0f)
        sceneFail("dielectric: absorb components must be >= 0");
    }
    return pushMaterial(h, m, tex_id);
  });
}

extern "C" int sundog_add_material_emissive(sundog_scene* h, const double color[3], int tex_id,
                                            double intensity, int two_sided) {
  return guardInt(-1, [&] {
    Material m = baseMaterial();
    m.kind = MT_EMISSIVE;
    m.color = nf3(color, m.color);
    m.texId = tex_id;
    m.intensity = nf(intensity, m.intensity);
    if (m.intensity < 0.0f) sceneFail("emissive: intensity must be >= 0");
    m.twoSided = two_sided == 1 ? 1 : 0;
    return pushMaterial(h, m, tex_id);
  });
}

extern "C" int sundog_add_material_water(sundog_scene* h, double ior, const double absorb[3],
                                         double wave_amp, double wave_freq,
                                         const double color[3]) {
  return guardInt(-1, [&] {
    Material m = baseMaterial();
    m.kind = MT_WATER;
    m.ior = nf(ior, 1.33f);
    m.absorb = nf3(absorb, v3(0.45f, 0.08f, 0.035f));
    m.waveAmp = nf(wave_amp, 0.05f);
    m.waveFreq = nf(wave_freq, 2.0f);
    m.color = nf3(color, v3(1.0f));
    m.texId = -1;
    if (m.waveAmp < 0.0f || m.waveFreq < 0.0f)
      sceneFail("water: wave_amp and wave_freq must be >= 0");
    if (m.absorb.x < 0.0f || m.absorb.y < 0.0f || m.absorb.z < 0.0f)
      sceneFail("water: absorb components must be >= 0");
    return pushMaterial(h, m, -1);
  });
}

extern "C" int sundog_add_material_plastic(sundog_scene* h, const double color[3], int tex_id,
                                           double roughness) {
  return guardInt(-1, [&] {
    Material m = baseMaterial();
    m.kind = MT_PLASTIC;
    m.color = nf3(color, m.color);
    m.texId = tex_id;
    m.roughness = nf(roughness, 0.15f);
    return pushMaterial(h, m, tex_id);
  });
}

extern "C" int sundog_add_mesh(sundog_scene* h, const char* obj_file, int smooth_normals,
                               const char* usemtl) {
  return guardInt(-1, [&] {
    phaseAdvance(h, 0, "meshes");
    if (!obj_file || !*obj_file) sceneFail("mesh: obj file is required");
    MeshRef m;
    m.objFile = obj_file;
    m.usemtl = usemtl ? usemtl : "";
    m.smoothNormals = smooth_normals != 0;
    int id = (int)h->scene.meshes.size();
    h->scene.meshes.push_back(m);
    return id;
  });
}

extern "C" int sundog_add_object(sundog_scene* h, int geom_kind, int mesh_id, int mat_front,
                                 int mat_back, int cutout_tex_id,
                                 const sundog_xform_step* steps, int num_steps, int nee,
                                 const sundog_physics_body* physics) {
  return guardInt(-1, [&] {
    phaseAdvance(h, 1, "objects");
    Object o;
    if (geom_kind < 0 || geom_kind > GK_MESH) sceneFail("object: unknown shape");
    o.geomKind = geom_kind;
    if (geom_kind == GK_MESH) {
      if (mesh_id < 0 || mesh_id >= (int)h->scene.meshes.size())
        sceneFail("object: mesh id out of range");
      o.meshId = mesh_id;
    }
    auto matId = [&](int id) -> uint16_t {
      if (id == SUNDOG_MAT_NONE) return kMatNone;
      if (id < 0 || id >= (int)h->scene.materials.size())
        sceneFail("object: material id out of range");
      return (uint16_t)id;
    };
    o.matFront = matId(mat_front);
    o.matBack = mat_back == SUNDOG_MAT_DEFAULT ? o.matFront : matId(mat_back);
    if (cutout_tex_id != -1) {
      if (cutout_tex_id < 0 || cutout_tex_id >= (int)h->scene.textures.size())
        sceneFail("object: cutout texture id out of range");
      o.cutoutTexId = cutout_tex_id;
    }
    if (num_steps > 0 && !steps) sceneFail("object: transform steps missing");
    std::vector<XformStep> xs((size_t)std::max(num_steps, 0));
    for (int i = 0; i < num_steps; i++) {
      xs[i].kind = steps[i].kind;
      xs[i].a = (float)steps[i].a;
      xs[i].b = (float)steps[i].b;
      xs[i].c = (float)steps[i].c;
    }
    o.xform = composeTransform(xs.data(), num_steps);
    o.nee = nee != 0;
    if (physics) {
      o.physics.enabled = true;
      o.physics.dynamic = physics->dynamic == 1;
      o.physics.density = nf(physics->density, o.physics.density);
      if (physics->has_velocity) o.physics.velocity = nf3(physics->velocity, o.physics.velocity);
      if (physics->has_angular_velocity)
        o.physics.angularVelocity = nf3(physics->angular_velocity, o.physics.angularVelocity);
      o.physics.friction = nf(physics->friction, o.physics.friction);
      o.physics.restitution = nf(physics->restitution, o.physics.restitution);
      o.physics.thickness = nf(physics->thickness, o.physics.thickness);
    }
    int id = (int)h->scene.objects.size();
    addObjectDerived(h->scene, o);
    return id;
  });
}

extern "C" int sundog_add_flame(sundog_scene* h, const double base[3], double height,
                                double radius, double intensity, double sigma,
                                double noise_scale, int64_t seed, double light_intensity) {
  return guardInt(-1, [&] {
    phaseAdvance(h, 2, "flames");
    if (!base) sceneFail("flame: base is required");
    Flame f;
    f.base = nf3(base, v3(0.0f));
    f.height = (float)height;
    f.radius = (float)radius;
    f.intensity = nf(intensity, 20.0f);
    f.sigma = nf(sigma, 4.0f);
    f.noiseScale = nf(noise_scale, 3.0f);
    f.seed = seed >= 0 ? (unsigned)seed : 0u;
    float li = nf(light_intensity, 12.0f);
    int id = (int)h->scene.flames.size();
    addFlameDerived(h->scene, f, li);
    return id;
  });
}

extern "C" int sundog_add_point_light(sundog_scene* h, const double position[3],
                                      const double intensity[3], double radius) {
  return guardInt(-1, [&] {
    phaseAdvance(h, 3, "lights");
    if (!position || !intensity) sceneFail("point light: position and intensity are required");
    Light l;
    l.kind = LT_POINT;
    l.p = nf3(position, v3(0.0f));
    l.radius = nf(radius, 0.0f);
    l.L = nf3(intensity, v3(0.0f));
    if (l.L.x < 0.0f || l.L.y < 0.0f || l.L.z < 0.0f)
      sceneFail("point light: intensity components must be >= 0");
    int id = (int)h->scene.lights.size();
    h->scene.lights.push_back(l);
    return id;
  });
}

extern "C" int sundog_add_distant_light(sundog_scene* h, const double direction[3],
                                        const double radiance[3]) {
  return guardInt(-1, [&] {
    phaseAdvance(h, 3, "lights");
    if (!direction || !radiance) sceneFail("distant light: direction and radiance are required");
    Light l;
    l.kind = LT_DISTANT;
    l.dir = normalize(nf3(direction, v3(0.0f)));
    l.L = nf3(radiance, v3(0.0f));
    if (l.L.x < 0.0f || l.L.y < 0.0f || l.L.z < 0.0f)
      sceneFail("distant light: radiance components must be >= 0");
    int id = (int)h->scene.lights.size();
    h->scene.lights.push_back(l);
    return id;
  });
}

extern "C" int sundog_scene_validate(sundog_scene* h) {
  return guardInt(SUNDOG_ERROR, [&] {
    if (!h) sceneFail("null scene handle");
    if (!h->hasCamera) sceneFail("missing camera");
    return SUNDOG_OK;
  });
}

extern "C" int sundog_set_flame_march(sundog_scene* h, int flame_id, int octaves, double step,
                                      int bounds) {
  return guardInt(SUNDOG_ERROR, [&] {
    if (!h) sceneFail("null scene handle");
    if (flame_id < 0 || flame_id >= (int)h->scene.flames.size())
      sceneFail("flame march: flame id out of range");
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
  return guardInt(SUNDOG_ERROR, [&] {
    if (!h) sceneFail("null scene handle");
    if (material_id < 0 || material_id >= (int)h->scene.materials.size())
      sceneFail("water octaves: material id out of range");
    if (h->scene.materials[material_id].kind != MT_WATER)
      sceneFail("water octaves: material is not water");
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
  return guardInt(SUNDOG_ERROR, [&] {
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
