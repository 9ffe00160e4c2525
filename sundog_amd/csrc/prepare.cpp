// Render preparation and the flat scene export (GPU-free).
#include "prepare.h"

#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>

#include "../../include/sundog_flat.h"

namespace sdk {

void applyOverrides(Scene& s, const sundog_render_options* opt) {
  if (!opt) return;
  RenderSettings& r = s.render;
  if (opt->spp > 0) r.spp = opt->spp;
  if (opt->width > 0) r.width = opt->width;
  if (opt->height > 0) r.height = opt->height;
  if (opt->seed >= 0) r.seed = (unsigned)opt->seed;
  if (opt->denoise >= 0) r.denoise = opt->denoise == 1;
  if (opt->transparent_shadows >= 0) r.transparentShadows = opt->transparent_shadows == 1;
  if (!std::isnan(opt->clamp) && opt->clamp >= 0.0) r.clampVal = (float)opt->clamp;
}

Affine invertAffine(const Affine& m) {
  double a[9] = {m.m[0], m.m[1], m.m[2], m.m[4], m.m[5], m.m[6], m.m[8], m.m[9], m.m[10]};
  double c00 = a[4] * a[8] - a[5] * a[7];
  double c01 = a[5] * a[6] - a[3] * a[8];
  double c02 = a[3] * a[7] - a[4] * a[6];
  double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
  if (!(std::fabs(det) > 0.0)) throw std::runtime_error("object transform is singular");
  double id = 1.0 / det;
  double inv[9];
  inv[0] = c00 * id;
  inv[1] = (a[2] * a[7] - a[1] * a[8]) * id;
  inv[2] = (a[1] * a[5] - a[2] * a[4]) * id;
  inv[3] = c01 * id;
  inv[4] = (a[0] * a[8] - a[2] * a[6]) * id;
  inv[5] = (a[2] * a[3] - a[0] * a[5]) * id;
  inv[6] = c02 * id;
  inv[7] = (a[1] * a[6] - a[0] * a[7]) * id;
  inv[8] = (a[0] * a[4] - a[1] * a[3]) * id;
  double t[3] = {m.m[3], m.m[7], m.m[11]};
  Affine r{};
  for (int i = 0; i < 3; i++) {
    r.m[i * 4 + 0] = (float)inv[i * 3 + 0];
    r.m[i * 4 + 1] = (float)inv[i * 3 + 1];
    r.m[i * 4 + 2] = (float)inv[i * 3 + 2];
    r.m[i * 4 + 3] = (float)(-(inv[i * 3 + 0] * t[0] + inv[i * 3 + 1] * t[1] + inv[i * 3 + 2] * t[2]));
  }
  return r;
}

void prepareScene(const Scene& s, Prepared& out) {
  auto t0 = std::chrono::steady_clock::now();
  out.scene = s;
  out.meshes.clear();
  for (const MeshRef& m : s.meshes)
    out.meshes.push_back(loadObj(resolvePath(s.baseDir, m.objFile), m.smoothNormals, m.usemtl));
  out.texImages.assign(s.textures.size(), HostImage{});
  for (size_t i = 0; i < s.textures.size(); i++)
    if (s.textures[i].kind == TX_IMAGE)
      out.texImages[i] = loadTextureImage(resolvePath(s.baseDir, s.textures[i].file), s.textures[i].srgb);
  if (s.bgKind == BG_ENVMAP) out.env = loadRadianceHdr(resolvePath(s.baseDir, s.env.file));
  out.inverses.clear();
  for (const Object& o : s.objects) out.inverses.push_back(invertAffine(o.xform));

  // Mesh area lights: world-space triangle CDF (reference src/capi_render.cpp:119-172).
  out.meshLights.clear();
  for (size_t oi = 0; oi < s.objects.size(); oi++) {
    const Object& o = s.objects[oi];
    if (o.geomKind != GK_MESH || o.lightId < 0) continue;
    const HostMesh& m = out.meshes[o.meshId];
    MeshLightPatch p;
    p.lightId = o.lightId;
    p.objectId = (int)oi;
    p.worldPos.resize(m.positions.size());
    for (size_t i = 0; i < m.positions.size(); i++) p.worldPos[i] = o.xform.point(m.positions[i]);
    size_t nt = m.numTris();
    p.cdf.assign(nt + 1, 0.0f);
    double total = 0.0;
    for (size_t t = 0; t < nt; t++) {
      const vec3& a = p.worldPos[m.indices[3 * t]];
      const vec3& b = p.worldPos[m.indices[3 * t + 1]];
      const vec3& c = p.worldPos[m.indices[3 * t + 2]];
      total += 0.5 * (double)length(cross(b - a, c - a));
      p.cdf[t + 1] = (float)total;
    }
    if (!(total > 0.0))
      throw std::runtime_error("mesh NEE light has zero surface area: " + s.meshes[o.meshId].objFile);
    for (size_t t = 1; t < nt; t++) p.cdf[t] = (float)(p.cdf[t] / total);
    p.cdf[nt] = 1.0f;
    p.area = (float)total;
    p.ngSign = detSign3(o.xform.vector(v3(1, 0, 0)), o.xform.vector(v3(0, 1, 0)),
                        o.xform.vector(v3(0, 0, 1)));
    out.meshLights.push_back(std::move(p));
  }
  out.loadMs = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace sdk

// ---------------------------------------------------------------------------
// flat scene export
// ---------------------------------------------------------------------------

namespace {

struct FlatOwner {
  sundog_flat_scene flat{};
  sdk::Prepared prep;
  std::vector<sundog_flat_texture> textures;
  std::vector<sundog_flat_material> materials;
  std::vector<sundog_flat_light> lights;
  std::vector<sundog_flat_flame> flames;
  std::vector<sundog_flat_object> objects;
  std::vector<sundog_flat_mesh> meshes;
};

void put3(float* d, sdk::vec3 v) {
  d[0] = v.x;
  d[1] = v.y;
  d[2] = v.z;
}

}  // namespace

extern "C" sundog_flat_scene* sundog_scene_flatten(sundog_scene* h, const sundog_render_options* opt) {
  using namespace sdk;
  try {
    if (!h) throw std::runtime_error("flatten: null scene handle");
    if (sundog_scene_validate(h) != SUNDOG_OK) throw std::runtime_error(sundog_last_error());
    std::unique_ptr<FlatOwner> own(new FlatOwner());
    Scene s = h->scene;
    applyOverrides(s, opt);
    prepareScene(s, own->prep);
    const Scene& S = own->prep.scene;
    sundog_flat_scene& f = own->flat;
    f.width = S.render.width;
    f.height = S.render.height;
    f.spp = S.render.spp;
    f.max_depth = S.render.maxDepth;
    f.clamp = S.render.clampVal;
    f.seed = S.render.seed;
    f.transparent_shadows = S.render.transparentShadows ? 1 : 0;
    f.exposure = S.render.exposure;
    Camera c = makeCamera(S.camera, S.render.width, S.render.height);
    put3(f.cam + 0, c.origin);
    put3(f.cam + 3, c.lowerLeft);
    put3(f.cam + 6, c.horizontal);
    put3(f.cam + 9, c.vertical);
    put3(f.cam + 12, c.u);
    put3(f.cam + 15, c.v);
    put3(f.cam + 18, c.w);
    f.cam[21] = c.lensRadius;
    f.bg_kind = S.bgKind;
    put3(f.bg_a, S.bgA);
    put3(f.bg_b, S.bgB);
    if (S.bgKind == BG_ENVMAP) {
      f.env_w = own->prep.env.width;
      f.env_h = own->prep.env.height;
      f.env_rgba = own->prep.env.rgba.data();
      f.env_rotate = S.env.rotateDeg * kPi / 180.0f;
      f.env_intensity = S.env.intensity;
      f.env_importance = S.env.importance ? 1 : 0;
    }
    for (size_t i = 0; i < S.textures.size(); i++) {
      const Texture& t = S.textures[i];
      sundog_flat_texture ft{};
      ft.kind = t.kind;
      put3(ft.a, t.a);
      put3(ft.b, t.b);
      ft.scale[0] = t.scale.x;
      ft.scale[1] = t.scale.y;
      ft.width = t.width;
      ft.img_w = own->prep.texImages[i].width;
      ft.img_h = own->prep.texImages[i].height;
      ft.img = own->prep.texImages[i].rgba.empty() ? nullptr : own->prep.texImages[i].rgba.data();
      own->textures.push_back(ft);
    }
    for (const Material& m : S.materials) {
      sundog_flat_material fm{};
      fm.kind = m.kind;
      fm.tex_id = m.texId;
      put3(fm.color, m.color);
      fm.roughness = m.roughness;
      fm.ior = m.ior;
      fm.intensity = m.intensity;
      fm.two_sided = m.twoSided;
      put3(fm.absorb, m.absorb);
      fm.wave_amp = m.waveAmp;
      fm.wave_freq = m.waveFreq;
      fm.wave_octaves = m.waveOctaves;
      own->materials.push_back(fm);
    }
    for (const Light& l : S.lights) {
      sundog_flat_light fl{};
      fl.kind = l.kind;
      put3(fl.p, l.p);
      put3(fl.u, l.u);
      put3(fl.v, l.v);
      put3(fl.n, l.n);
      put3(fl.L, l.L);
      fl.radius = l.radius;
      fl.area = l.area;
      put3(fl.dir, l.dir);
      fl.two_sided = l.twoSided;
      fl.tex_id = l.texId;
      fl.flame_id = l.flameId;
      fl.object_id = l.objectId;
      own->lights.push_back(fl);
    }
    for (const Flame& fl : S.flames) {
      sundog_flat_flame ff{};
      put3(ff.base, fl.base);
      ff.height = fl.height;
      ff.radius = fl.radius;
      ff.intensity = fl.intensity;
      ff.sigma = fl.sigma;
      ff.noise_scale = fl.noiseScale;
      ff.seed = fl.seed;
      ff.octaves = fl.octaves;
      ff.step = fl.step;
      ff.bounds = fl.bounds;
      own->flames.push_back(ff);
    }
    for (size_t i = 0; i < S.objects.size(); i++) {
      const Object& o = S.objects[i];
      sundog_flat_object fo{};
      fo.geom_kind = o.geomKind;
      fo.mesh_id = o.meshId;
      fo.mat_front = o.matFront;
      fo.mat_back = o.matBack;
      fo.cutout_tex = o.cutoutTexId;
      fo.light_id = o.lightId;
      memcpy(fo.xform, o.xform.m, sizeof fo.xform);
      memcpy(fo.inv, own->prep.inverses[i].m, sizeof fo.inv);
      own->objects.push_back(fo);
    }
    for (const HostMesh& m : own->prep.meshes) {
      sundog_flat_mesh fm{};
      fm.num_verts = (int32_t)m.positions.size();
      fm.num_tris = (int32_t)m.numTris();
      fm.positions = &m.positions[0].x;
      fm.indices = m.indices.data();
      fm.normals = m.normals.empty() ? nullptr : &m.normals[0].x;
      fm.uvs = m.uvs.empty() ? nullptr : &m.uvs[0].x;
      own->meshes.push_back(fm);
    }
    f.num_textures = (int32_t)own->textures.size();
    f.num_materials = (int32_t)own->materials.size();
    f.num_lights = (int32_t)own->lights.size();
    f.num_flames = (int32_t)own->flames.size();
    f.num_objects = (int32_t)own->objects.size();
    f.num_meshes = (int32_t)own->meshes.size();
    f.textures = own->textures.data();
    f.materials = own->materials.data();
    f.lights = own->lights.data();
    f.flames = own->flames.data();
    f.objects = own->objects.data();
    f.meshes = own->meshes.data();
    FlatOwner* raw = own.release();
    return &raw->flat;
  } catch (const std::exception& e) {
    sdk::setLastError(e.what());
    return nullptr;
  }
}

extern "C" void sundog_flat_free(sundog_flat_scene* f) {
  if (!f) return;
  // flat is the first member of FlatOwner
  delete reinterpret_cast<FlatOwner*>(f);
}
