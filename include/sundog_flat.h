/* sundog_flat.h — plain-array view of a derived scene.
 *
 * Produced by sundog_scene_flatten() after the scene-construction calls have
 * composed transforms, registered NEE lights and applied render overrides
 * (reference src/capi_scene.cpp, src/scene_build.cpp, src/capi_render.cpp:73-83).
 * Meshes are loaded (OBJ), mesh-light CDF inputs are implicit in the mesh
 * arrays, and the environment image is decoded to linear float RGBA.
 *
 * Every record is a fixed-size block of 32-bit fields so the layout is the
 * same for C, C++ and ctypes. All enums use the reference values
 * (device/params.h:12-16).
 */
#ifndef SUNDOG_FLAT_H
#define SUNDOG_FLAT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int32_t kind;        /* 0 solid, 1 image, 2 checker, 3 grid */
  float a[3], b[3];
  float scale[2];
  float width;
  int32_t img_w, img_h; /* image textures: decoded RGBA float, row 0 = top */
  const float* img;
} sundog_flat_texture;

typedef struct {
  int32_t kind;        /* 0 lambert, 1 metal, 2 dielectric, 3 emissive, 4 water, 5 plastic */
  int32_t tex_id;
  float color[3];
  float roughness;
  float ior;
  float intensity;
  int32_t two_sided;
  float absorb[3];
  float wave_amp;
  float wave_freq;
  int32_t wave_octaves; /* 3 = reference */
} sundog_flat_material;

typedef struct {
  int32_t kind;        /* 0 rect, 1 disk, 2 sphere, 3 point, 4 distant, 5 mesh */
  float p[3], u[3], v[3], n[3];
  float L[3];
  float radius;
  float area;          /* mesh lights: filled at render time from the mesh */
  float dir[3];
  int32_t two_sided;
  int32_t tex_id;
  int32_t flame_id;
  int32_t object_id;   /* mesh lights: the emitting object, else -1 */
} sundog_flat_light;

typedef struct {
  float base[3];
  float height, radius, intensity, sigma, noise_scale;
  uint32_t seed;
  int32_t octaves;     /* 3 = reference */
  float step;          /* <= 0: 32 fixed steps (reference) */
  int32_t bounds;      /* 0 cylinder (reference), 1 box */
} sundog_flat_flame;

typedef struct {
  int32_t geom_kind;   /* 0..5 */
  int32_t mesh_id;
  int32_t mat_front, mat_back; /* 0xFFFF = pass-through face */
  int32_t cutout_tex;
  int32_t light_id;
  float xform[12];     /* row-major 3x4 object -> world */
  float inv[12];       /* world -> object: double-precision inverse of xform,
                          rounded once; both renderers transform rays with it */
} sundog_flat_object;

typedef struct {
  int32_t num_verts, num_tris;
  const float* positions;   /* num_verts x 3 */
  const uint32_t* indices;  /* num_tris x 3 */
  const float* normals;     /* num_verts x 3 or NULL */
  const float* uvs;         /* num_verts x 2 or NULL */
} sundog_flat_mesh;

typedef struct sundog_flat_scene {
  int32_t width, height, spp, max_depth;
  float clamp;
  uint32_t seed;
  int32_t transparent_shadows;
  float exposure;
  /* camera: origin, lower_left, horizontal, vertical, u, v, w, lens_radius */
  float cam[22];
  int32_t bg_kind;     /* 0 solid, 1 gradient, 2 envmap */
  float bg_a[3], bg_b[3];
  int32_t env_w, env_h;
  const float* env_rgba; /* env_w*env_h*4 linear float, row 0 = zenith */
  float env_rotate;      /* radians */
  float env_intensity;
  int32_t env_importance;
  int32_t num_textures, num_materials, num_lights, num_flames, num_objects, num_meshes;
  const sundog_flat_texture* textures;
  const sundog_flat_material* materials;
  const sundog_flat_light* lights;
  const sundog_flat_flame* flames;
  const sundog_flat_object* objects;
  const sundog_flat_mesh* meshes;
} sundog_flat_scene;

#ifdef __cplusplus
}
#endif

#endif /* SUNDOG_FLAT_H */
