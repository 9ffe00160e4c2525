/* sundog_api.h — public C ABI of libsundog.so (MI355X / gfx950 build).
 *
 * Part 1 declares the reference ABI with identical names, argument types and
 * struct layouts, so a ctypes call program written against the reference
 * (scenes/scenelib.py:724-880) drives this library unchanged. Each group
 * cites the reference declaration it replaces (src/sundog_api.h:<line>).
 *
 * Conventions (reference src/sundog_api.h:1-15, src/capi_internal.h:26-56):
 *   - doubles are narrowed to float exactly once at the boundary;
 *   - "unset" is NaN for doubles, NULL for vectors, -1 for tri-state ints;
 *   - set_* / validate / probe / render return 0 (ok) or 1 (error);
 *     add_* return an id >= 0, or -1; create returns NULL on error;
 *   - the message of the last failure is in sundog_last_error()
 *     (thread-local);
 *   - call order is config/registries -> objects -> flames -> explicit
 *     lights -> render, which freezes the light-list layout
 *     [auto NEE area lights in object order][2 per flame][explicit].
 *
 * Part 2 holds additive entry points with no counterpart in the reference:
 * resident render sessions (scene data and film kept in HBM across passes),
 * ray-query and sampling probes used by parity tests, and the flat scene
 * view that the CPU oracle consumes. They change no Part-1 symbol.
 */
#ifndef SUNDOG_API_H
#define SUNDOG_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
#define SUNDOG_API extern "C" __attribute__((visibility("default")))
#else
#define SUNDOG_API __attribute__((visibility("default")))
#endif

/* ======================================================================== */
/* Part 1 — reference ABI                                                   */
/* ======================================================================== */

/* error model: reference src/sundog_api.h:27-30 */
#define SUNDOG_OK 0
#define SUNDOG_ERROR 1
SUNDOG_API const char* sundog_last_error(void);

/* lifetime: reference src/sundog_api.h:32-35 */
typedef struct sundog_scene sundog_scene;
SUNDOG_API sundog_scene* sundog_scene_create(const char* base_dir);
SUNDOG_API void sundog_scene_destroy(sundog_scene*);

/* enums: reference src/sundog_api.h:37-54 */
enum {
  SUNDOG_GK_SPHERE = 0,
  SUNDOG_GK_RECT = 1,
  SUNDOG_GK_DISK = 2,
  SUNDOG_GK_CYLINDER = 3,
  SUNDOG_GK_PARABOLA = 4,
  SUNDOG_GK_MESH = 5,
};
enum {
  SUNDOG_XF_SCALE = 0,
  SUNDOG_XF_TRANSLATE = 1,
  SUNDOG_XF_ROTATE_X = 2,
  SUNDOG_XF_ROTATE_Y = 3,
  SUNDOG_XF_ROTATE_Z = 4,
};
#define SUNDOG_MAT_NONE (-1)
#define SUNDOG_MAT_DEFAULT (-2)

/* config blocks: reference src/sundog_api.h:56-76 */
SUNDOG_API int sundog_set_render(sundog_scene*, int width, int height, int spp,
                                 int max_depth, double clamp, int64_t seed,
                                 double exposure, int transparent_shadows);
SUNDOG_API int sundog_set_camera(sundog_scene*, const double lookfrom[3],
                                 const double lookat[3], const double up[3],
                                 double vfov, double aperture, double focus_dist);
SUNDOG_API int sundog_set_background_solid(sundog_scene*, const double color[3]);
SUNDOG_API int sundog_set_background_gradient(sundog_scene*, const double horizon[3],
                                              const double zenith[3]);
SUNDOG_API int sundog_set_background_envmap(sundog_scene*, const char* file,
                                            double rotate_deg, double intensity,
                                            int importance);
SUNDOG_API int sundog_set_physics(sundog_scene*, const double gravity[3],
                                  double timestep, double max_time,
                                  double friction, double restitution,
                                  int pos_iters, int vel_iters,
                                  double sleep_threshold, double stop_time);

/* registries: reference src/sundog_api.h:78-111 */
SUNDOG_API int sundog_add_texture_solid(sundog_scene*, const double color[3]);
SUNDOG_API int sundog_add_texture_checker(sundog_scene*, const double a[3],
                                          const double b[3], const double scale[2]);
SUNDOG_API int sundog_add_texture_grid(sundog_scene*, const double a[3],
                                       const double b[3], const double scale[2],
                                       double width);
SUNDOG_API int sundog_add_texture_image(sundog_scene*, const char* file, int srgb);
SUNDOG_API int sundog_add_material_lambert(sundog_scene*, const double color[3],
                                           int tex_id);
SUNDOG_API int sundog_add_material_metal(sundog_scene*, const double color[3],
                                         int tex_id, double roughness);
SUNDOG_API int sundog_add_material_dielectric(sundog_scene*, double ior,
                                              const double absorb[3],
                                              const double color[3], int tex_id,
                                              double roughness);
SUNDOG_API int sundog_add_material_emissive(sundog_scene*, const double color[3],
                                            int tex_id, double intensity,
                                            int two_sided);
SUNDOG_API int sundog_add_material_water(sundog_scene*, double ior,
                                         const double absorb[3], double wave_amp,
                                         double wave_freq, const double color[3]);
SUNDOG_API int sundog_add_material_plastic(sundog_scene*, const double color[3],
                                           int tex_id, double roughness);
SUNDOG_API int sundog_add_mesh(sundog_scene*, const char* obj_file,
                               int smooth_normals, const char* usemtl);

/* objects: reference src/sundog_api.h:113-138 */
typedef struct {
  int32_t kind;
  double a, b, c;
} sundog_xform_step;

typedef struct {
  int32_t dynamic;
  double density;
  int32_t has_velocity, has_angular_velocity;
  double velocity[3], angular_velocity[3];
  double friction, restitution;
  double thickness;
} sundog_physics_body;

SUNDOG_API int sundog_add_object(sundog_scene*, int geom_kind, int mesh_id,
                                 int mat_front, int mat_back, int cutout_tex_id,
                                 const sundog_xform_step* steps, int num_steps,
                                 int nee, const sundog_physics_body* physics);

/* flames and explicit lights: reference src/sundog_api.h:140-152 */
SUNDOG_API int sundog_add_flame(sundog_scene*, const double base[3], double height,
                                double radius, double intensity, double sigma,
                                double noise_scale, int64_t seed,
                                double light_intensity);
SUNDOG_API int sundog_add_point_light(sundog_scene*, const double position[3],
                                      const double intensity[3], double radius);
SUNDOG_API int sundog_add_distant_light(sundog_scene*, const double direction[3],
                                        const double radiance[3]);

/* validation / probe / render: reference src/sundog_api.h:154-179 */
SUNDOG_API int sundog_scene_validate(sundog_scene*);

typedef struct {
  char name[256];
  int32_t cc_major, cc_minor;           /* gfx950 reports 9, 5 */
  int32_t driver_version, runtime_version;
  uint64_t total_mem_mb;
  uint32_t optix_header_version;        /* always 0: no OptiX */
  uint32_t rtcore_version;              /* always 0: no RT cores */
} sundog_device_info;
SUNDOG_API int sundog_probe(sundog_device_info* out);

typedef struct {
  const char* out_path;
  const char* stats_path;
  const char* aov_albedo_path, *aov_normal_path;
  const char* scene_name;
  int32_t spp, width, height;
  int64_t seed;
  int32_t denoise, transparent_shadows;
  double clamp;
  double physics_time;
  int32_t quiet;
} sundog_render_options;
SUNDOG_API int sundog_render(sundog_scene*, const sundog_render_options*);

/* ======================================================================== */
/* Part 2 — additive entry points (no counterpart in the reference)         */
/* ======================================================================== */

/* Scene-level march/ripple quality knobs. The reference hard-codes a 32-step
 * march over a bounding cylinder with 3-octave fbm for flames
 * (device/volume.cuh:90-115) and 3-octave fbm for water ripples
 * (device/noise.cuh:67-79); those stay the defaults. These calls select the
 * step-length / box-bounds / octave variants. */
#define SUNDOG_FLAME_BOUNDS_CYLINDER 0
#define SUNDOG_FLAME_BOUNDS_BOX 1
SUNDOG_API int sundog_set_flame_march(sundog_scene*, int flame_id, int octaves,
                                      double step /* <=0: 32 fixed steps */,
                                      int bounds /* SUNDOG_FLAME_BOUNDS_* */);
SUNDOG_API int sundog_set_water_octaves(sundog_scene*, int material_id, int octaves);

/* JSON description of the derived scene state (render settings, camera and
 * its derived frame, background, registries, objects with composed
 * transforms and light back-links, the light list, flames). GPU-free.
 * Writes at most cap bytes (NUL-terminated) and reports the size needed. */
SUNDOG_API int sundog_scene_describe(sundog_scene*, char* buf, size_t cap, size_t* needed);

/* ---- resident render sessions ----------------------------------------------
 * A session owns every device resource of one render (BVHs, instance table,
 * materials, lights, env map + CDFs, film) and keeps them in HBM until it is
 * destroyed. sundog_render is create + progressive passes + download + write.
 * Render settings come from the scene after applying `opt` overrides exactly
 * as sundog_render does (opt may be NULL). The scene handle is not retained. */
typedef struct sundog_session sundog_session;

#define SUNDOG_KERNEL_MEGA 0      /* one work-item per pixel, regenerated paths */
#define SUNDOG_KERNEL_WAVEFRONT 1 /* staged queues, material-sorted shading */

typedef struct {
  uint64_t rays_traced;     /* radiance + traced shadow queries */
  uint64_t paths;           /* pixels x samples rendered since reset */
  double render_ms;         /* device time of render passes since reset */
  double build_ms;          /* BVH build + uploads at create */
  uint64_t device_bytes;    /* bytes of resident device allocations */
  int32_t stack_depth;      /* traversal stack entries needed (checked) */
  int32_t tlas_nodes, blas_nodes;
  int32_t pad;
  /* traversal work since reset, counted only while count_work is on:
   * 4-wide node visits (4 box tests each), triangle tests, quadric tests,
   * world->object instance transforms */
  uint64_t node_visits, tri_tests, quad_tests, inst_transforms;
  int32_t bvh_builder;      /* 0: host binned SAH (default), 1: GPU LBVH, 2: GPU LBVH meshes + host SAH
                               top, 3: GPU PLOC, 4: GPU PLOC meshes + host SAH top */
  int32_t pad2;
  double bvh_gpu_ms;        /* wall time of the last GPU LBVH rebuild */
  int32_t waves;            /* megakernel occupancy variant in use (0: auto mode still timing) */
  int32_t stack_lds;        /* LDS traversal-stack entries per lane */
  int32_t sample_groups;    /* megakernel sample groups per tile in use (0: auto mode still timing) */
  int32_t pad3;
} sundog_session_stats;

SUNDOG_API sundog_session* sundog_session_create(sundog_scene*,
                                                 const sundog_render_options* opt);
SUNDOG_API void sundog_session_destroy(sundog_session*);
/* Clears the film and counters. */
SUNDOG_API int sundog_session_reset(sundog_session*);
/* Renders samples [sample_offset, sample_offset+count) of every pixel into
 * the film (one progressive pass). Asynchronous on the session stream. */
SUNDOG_API int sundog_session_render(sundog_session*, int sample_offset, int count);
SUNDOG_API int sundog_session_set_kernel(sundog_session*, int kernel);
SUNDOG_API int sundog_session_set_count_rays(sundog_session*, int on);
/* Launch-shape knobs picked by the autotuner. Keys: "waves" (2..5: the
 * minimum waves per SIMD the megakernel's register budget allows; 0, the
 * default, picks one per scene by timing the first passes), "wf_waves" (4|5: the same for the wavefront query kernels), "stack"
 * (4..24: LDS traversal-stack entries per lane of the megakernel; default
 * 12), "groups" (1..16: blocks per tile sharing a pass's samples, which a
 * separate kernel then folds in sample order; 0, the default, picks from
 * 1, 2, 3, 4, 6, 8, 16 by timing passes after the occupancy variant is chosen;
 * films are bit-identical for every choice), and
 * "bvh" (1: rebuild both BVH levels on the GPU with the LBVH builder; 2:
 * the mesh levels with the GPU LBVH and the top level with the host SAH
 * builder over instance boxes computed on the GPU; 3: both levels on the GPU
 * with PLOC clustering; 4: the mesh levels with GPU PLOC and the top level
 * with the host SAH builder), and "ploc_radius" (1..32, default 16:
 * neighbours searched on either side by the PLOC build; set before "bvh"). */
SUNDOG_API int sundog_session_set_tuning(sundog_session*, const char* key, int value);
/* Times the megakernel's occupancy variants on passes of `count` samples
 * from sample 0 (one untimed warm-up pass, then 5, 4, 3 and 2 waves per
 * SIMD twice each, the faster pass counting), then, when `count` <= 16 and
 * "groups" is 0, the sample groups 1, 2, 3, 4, 6, 8, 16 at that variant the same
 * way and then the occupancy variants again at the chosen groups; keeps the
 * fastest for later passes and clears the film. A value
 * fixed by "waves" or "groups" is kept and not timed. *waves_out receives
 * the occupancy variant in use (the groups are in the stats). */
SUNDOG_API int sundog_session_autotune(sundog_session*, int count, int* waves_out);
/* Runs the instrumented kernel variant that counts traversal work. */
SUNDOG_API int sundog_session_set_count_work(sundog_session*, int on);
/* Waits for outstanding work. */
SUNDOG_API int sundog_session_sync(sundog_session*);
/* Copies width*height float4 (RGBA) into each non-NULL host buffer. */
SUNDOG_API int sundog_session_download(sundog_session*, float* beauty,
                                       float* albedo, float* normal);
SUNDOG_API int sundog_session_get_stats(sundog_session*, sundog_session_stats*);
/* The resident two-level BVH as the traversal kernels read it (for tree
 * quality and build-determinism checks). */
typedef struct {
  int32_t num_nodes;      /* 128-byte 4-wide nodes: the top level at [0, tlas_nodes), the meshes' after it */
  int32_t tlas_nodes;
  int32_t num_tris;       /* leaf-ordered triangles: 3 float4 each, v0.w = the triangle's index bits */
  int32_t num_instances;
  uint32_t tlas_root;     /* packed reference of the top-level root */
  int32_t pad;
} sundog_bvh_layout;
/* Fills `out`; copies the node array (num_nodes x 128 bytes: six float4 rows
 * of child bounds lo.x, hi.x, lo.y, hi.y, lo.z, hi.z, then the int4 of packed
 * child references and the int4 of child primitive counts) into `nodes` and
 * the triangle pool (num_tris x 12 floats) into `tris` when non-NULL and
 * large enough (an error otherwise), and each instance's mesh root node (-1
 * for quadrics) into the first blas_roots_cap entries of blas_roots when
 * non-NULL. Packed references: bit 31 leaf, bits 27-30 count - 1, bits 0-26
 * the first primitive (a triangle position, or an instance in the top
 * level); inner references are node indices; 0xFFFFFFFF is an empty slot. */
SUNDOG_API int sundog_session_bvh(sundog_session*, sundog_bvh_layout* out, void* nodes, size_t nodes_bytes,
                                  float* tris, size_t tris_bytes, int32_t* blas_roots, int blas_roots_cap);
SUNDOG_API int sundog_session_size(sundog_session*, int* width, int* height, int* spp);
/* The HIP stream (hipStream_t) the session launches on, for event timing. */
SUNDOG_API void* sundog_session_stream(sundog_session*);

/* Closest-hit query of n world rays (rays: n x 8 floats ox,oy,oz,tmin,
 * dx,dy,dz,tmax) through the session's two-level BVH. Outputs per ray: t
 * (INFINITY on miss), instance id (-1 on miss), primitive id (triangle index
 * in its mesh; 0 or 1 = root index for quadrics). brute != 0 tests every
 * primitive of every instance instead of traversing. Radiance-ray masking
 * applies (pass-through faces are skipped). Synchronous. */
SUNDOG_API int sundog_session_trace(sundog_session*, const float* rays, int n,
                                    int brute, float* t_out, int32_t* inst_out,
                                    int32_t* prim_out);
/* The closest-hit query above for rays and outputs already in device
 * memory, allocating nothing: no counterpart in the reference; it is for
 * callers that cannot allocate (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). d_rays: n x 8 floats; d_t, d_inst,
 * d_prim: n entries each, filled as sundog_session_trace fills its outputs.
 * Asynchronous on the session's stream (sundog_session_stream): synchronize
 * before reading the outputs. The traversal stack's overflow lives in a
 * static scratch array of SUNDOG_TRACE_SCRATCH_BYTES per GPU, shared by every
 * call on that GPU: calls that overlap on one GPU race. A call whose rays
 * would need more overflow than the array holds (deep trees times n) is
 * refused with an error that names the largest n the scene allows. */
#define SUNDOG_TRACE_SCRATCH_BYTES (32u << 20)
SUNDOG_API int sundog_session_trace_noalloc(sundog_session*, const float* d_rays, int n,
                                            float* d_t, int32_t* d_inst, int32_t* d_prim);
/* Shadow transmittance of n segments (same ray layout; tmax = segment end)
 * including the transparent-shadow accumulation; out: n x 3 floats. */
SUNDOG_API int sundog_session_shadow(sundog_session*, const float* rays, int n,
                                     float* tr_out);
/* Environment importance sampling on device: for n pairs (u0,u1) returns the
 * marginal row and conditional column picked by the binary searches. */
SUNDOG_API int sundog_session_env_bins(sundog_session*, const float* u, int n,
                                       int32_t* row_out, int32_t* col_out);
/* The same bins by a chosen search, with the search kernel's time (ms, HIP
 * events): mode 0 = the renderer's per-lane 16-ary table, mode 1 = a
 * wave-cooperative search (the wave serves its lanes' lookups in turn, all
 * 64 lanes probing keys of one CDF per step). Both return the bisection
 * bins. No counterpart in the reference (its search is per thread). */
SUNDOG_API int sundog_session_env_bins_mode(sundog_session*, const float* u, int n, int mode,
                                            int32_t* row_out, int32_t* col_out, float* kernel_ms);
/* Light sampling on device: n samples of light `light` (an index into the
 * scene's light list) from the points xs (n x 3); sample i draws from the
 * PCG32 stream init(seed0 + i, seq). out: 12 floats per sample — wi (3),
 * distance, Li (3), solid-angle pdf (0 when invalid), flags (bit 0 valid,
 * bit 1 delta), draws taken, the MIS-side pdf lightPdfSolidAngle returns for
 * the sampled point (NaN for mesh and delta lights), 0. The sampler is the
 * one the render kernels call (device/light_sample.cuh:24-226). */
SUNDOG_API int sundog_session_light_samples(sundog_session*, int light, const float* xs, int n,
                                            uint64_t seed0, uint64_t seq, float* out);
/* Flame density field on device: extinction sigma (n floats) and emission
 * (n x 3 floats) of flame `flame` at n points (x, y, z), the values the
 * flame march integrates (device/volume.cuh:27-153). */
SUNDOG_API int sundog_session_flame_field(sundog_session*, int flame, const float* pts,
                                          int n, float* sigma_out,
                                          float* emission_out);

/* ---- serving: resident scenes + continuous batching ------------------------
 * A server keeps every added scene resident (BVHs, tables, textures) and
 * runs a worker thread that owns one GPU stream. Requests (scene, resolution,
 * spp, seed) are admitted FIFO while their film fits the HBM budget
 * (48 bytes per pixel: beauty + albedo + normal float4). Each batch renders
 * the next pass_spp samples of every admitted request in one launch of the
 * persistent tile-queue kernel; requests join and leave between batches.
 * A request's film is bit-identical to rendering it alone. Times are
 * milliseconds since the server was created. Requests larger than the whole
 * budget complete immediately with status REJECTED. */
typedef struct sundog_server sundog_server;

#define SUNDOG_REQ_QUEUED 0
#define SUNDOG_REQ_RUNNING 1
#define SUNDOG_REQ_DONE 2
#define SUNDOG_REQ_REJECTED 3
#define SUNDOG_REQ_FAILED 4

typedef struct {
  int64_t id;
  int32_t status;
  int32_t scene;
  int32_t width, height, spp;
  int32_t pad;
  double t_submit_ms, t_admit_ms, t_done_ms;
  uint64_t paths;
} sundog_request_info;

typedef struct {
  uint64_t paths, rays_traced, batches, work_items;
  double kernel_ms, busy_ms;
  uint64_t used_bytes, budget_bytes;
  int32_t active, waiting, tile_size, waves, pass_spp, blocks_per_cu;
  int32_t stack, graph_passes;
  uint64_t graph_launches;
} sundog_server_stats;

/* Autotuner result: the launch shape it kept and the rates it measured on
 * the calibration request (paths/s of the kernel time). */
typedef struct {
  int32_t tile, waves, stack, blocks_per_cu, graph_passes, trials;
  double start_paths_per_s, best_paths_per_s;
} sundog_tune_result;

SUNDOG_API sundog_server* sundog_server_create(uint64_t hbm_budget_bytes, int pass_spp,
                                               int tile_size /* 8 | 16 | 32 */);
SUNDOG_API void sundog_server_destroy(sundog_server*);
/* Before the first submit. Returns the scene id, or -1. */
SUNDOG_API int sundog_server_add_scene(sundog_server*, sundog_scene*, const sundog_render_options* opt);
/* Keys: "tile" (8|16|32: pixels per side of a work item), "waves" (2..5:
 * occupancy variant of the persistent kernel), "stack" (4..24: LDS
 * traversal-stack entries per lane), "blocks" (queue depth: resident
 * persistent blocks per CU, 0 = as many as fit), "graph_passes" (1..64:
 * progressive passes captured into one graph launch), "pass_spp" (> 0),
 * "serial_waves" (0: picked per request by timing its first passes, the
 * default; or 2..5: occupancy variant of the per-request baseline), and
 * "groups" (sample groups per pass: 0, the default, splits each tile's pass
 * over several work items when a batch's tiles cannot fill the resident
 * waves; 1 never splits; 2..16 always splits that many ways, at most
 * pass_spp; passes of more than 16 samples are never split; films are
 * bit-identical either way). */
SUNDOG_API int sundog_server_set_tuning(sundog_server*, const char* key, int value);
/* Before the first submit: renders a calibration request of scene_id
 * (width x height x spp) under candidate launch shapes by coordinate descent
 * over waves, stack, tile, blocks and graph_passes, and keeps the fastest. */
SUNDOG_API int sundog_server_autotune(sundog_server*, int scene_id, int width, int height, int spp,
                                      sundog_tune_result* out);
/* Returns the request id (>= 0) or -1; seed < 0 uses the scene's seed. */
SUNDOG_API int64_t sundog_server_submit(sundog_server*, int scene_id, int width, int height, int spp,
                                        int64_t seed);
/* Drains up to max completion records (done, rejected or failed); returns the count. */
SUNDOG_API int sundog_server_poll(sundog_server*, sundog_request_info* out, int max);
/* Copies a finished request's films (width*height float4 each; NULL skips). */
SUNDOG_API int sundog_server_fetch(sundog_server*, int64_t id, float* beauty, float* albedo, float* normal);
/* Frees a finished request's host-side record. */
SUNDOG_API int sundog_server_release(sundog_server*, int64_t id);
/* Blocks until no request is queued or running (timeout_s <= 0: no limit). */
SUNDOG_API int sundog_server_wait_idle(sundog_server*, double timeout_s);
SUNDOG_API int sundog_server_get_stats(sundog_server*, sundog_server_stats*);
/* Baseline: one request rendered alone on a resident scene with the
 * megakernel in 16-spp chunks, synchronising after each chunk. Only while
 * the batching loop is idle. */
SUNDOG_API int sundog_server_render_serial(sundog_server*, int scene_id, int width, int height, int spp,
                                           int64_t seed, float* beauty, double* ms_out);

/* PCG32 on device: n outputs of the stream init(seed, seq), raw u32 draws. */
SUNDOG_API int sundog_device_pcg32(uint64_t seed, uint64_t seq, int n, uint32_t* out);

/* ---- flat scene view -------------------------------------------------------
 * A read-only view of the derived scene (after render overrides), with
 * meshes loaded and the environment image decoded. Pointers stay valid until
 * sundog_flat_free. Layouts are plain arrays of floats/ints documented in
 * sundog_flat.h. */
struct sundog_flat_scene;
SUNDOG_API struct sundog_flat_scene* sundog_scene_flatten(sundog_scene*,
                                                          const sundog_render_options* opt);
SUNDOG_API void sundog_flat_free(struct sundog_flat_scene*);

/* ---- HDR output transfer ------------------------------------------------------
 * The chain of the reference's output writer (src/film.cpp:96-115,
 * src/transfer.h): exposure, BT.709 -> BT.2020 gamut, 1.0 -> 203 nits,
 * SMPTE ST 2084 (PQ), 12-bit code values. sundog_render writes these codes
 * as a 16-bit PPM (maxval 4095) beside the linear PFM. */
SUNDOG_API float sundog_pq_oetf(float linear);  /* 1.0 = 10,000 nits; clamps, NaN -> 0 */
SUNDOG_API float sundog_pq_eotf(float code);
SUNDOG_API void sundog_bt709_to_2020(const float in[3], float out[3]);
/* width*height RGBA float film -> width*height*3 12-bit codes. */
SUNDOG_API int sundog_encode_pq12(const float* rgba, int width, int height, float exposure, uint16_t* out);

/* ---- learned denoiser -------------------------------------------------------
 * Replaces the vendor AI denoiser behind src/denoise.cpp:18-53 (HDR beauty +
 * albedo + normal guides in, denoised HDR out) with a BF16 U-Net whose
 * forward and backward passes are this library's own MFMA kernels.
 *
 * Image batches are NHWC float32 on the host: input [n][h][w][9] = beauty
 * RGB, albedo RGB, normal XYZ; target and output [n][h][w][3] (linear HDR).
 * h and w must be multiples of 16. Parameters are one flat float32 vector
 * (per convolution: weights [cout][3*3][cin], then bias[cout]; cin padded to
 * 16 for the first layer, cout padded to 32 for the output layer). */
typedef struct sundog_denoiser sundog_denoiser;
/* He-normal initialisation from a PCG32 stream; NULL on error. */
SUNDOG_API sundog_denoiser* sundog_denoiser_create(uint64_t seed);
SUNDOG_API void sundog_denoiser_destroy(sundog_denoiser*);
SUNDOG_API int64_t sundog_denoiser_param_count(sundog_denoiser*);
SUNDOG_API int sundog_denoiser_get_params(sundog_denoiser*, float* out);
SUNDOG_API int sundog_denoiser_set_params(sundog_denoiser*, const float* in);
/* One forward + backward of the L1 loss in log(1+x) space: loss, the full
 * gradient vector and the network output (log space, [n][h][w][3]); each
 * output pointer may be NULL. Parameters are not changed. */
SUNDOG_API int sundog_denoiser_forward_backward(sundog_denoiser*, const float* input, const float* target, int n,
                                                int h, int w, float* loss, float* grads, float* pred);
/* forward_backward followed by one Adam step (beta 0.9 / 0.999). */
SUNDOG_API int sundog_denoiser_train_step(sundog_denoiser*, const float* input, const float* target, int n, int h,
                                          int w, float lr, float* loss);
/* Denoised linear HDR [n][h][w][3]. */
SUNDOG_API int sundog_denoiser_infer(sundog_denoiser*, const float* input, int n, int h, int w, float* out);
/* Mean ms of `iters` forward passes at (n, h, w) on device-resident input,
 * and the algorithmic FLOPs of one forward. */
SUNDOG_API int sundog_denoiser_time_forward(sundog_denoiser*, int n, int h, int w, int iters, double* ms,
                                            double* flops);
SUNDOG_API int sundog_denoiser_save(sundog_denoiser*, const char* path);
SUNDOG_API int sundog_denoiser_load(sundog_denoiser*, const char* path);

#endif /* SUNDOG_API_H */
