/* oracle.h — CPU oracle for the sundog render path.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product (sundog_amd/, libsundog.so)
 * includes, links or calls this code; only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg use it, as the oracle.
 *
 * The oracle is a scalar C++17 + OpenMP restatement of the reference render
 * path (device/programs.cu:311-572 and the device helper headers), consuming
 * the flat scene view (include/sundog_flat.h). It has its own BVH (binary,
 * median split) and shares no code with the GPU implementation.
 */
#ifndef SUNDOG_ORACLE_H
#define SUNDOG_ORACLE_H

#include <stdint.h>

#include "../include/sundog_flat.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int32_t sample_offset, sample_count; /* samples [offset, offset+count) */
  int32_t tile_stride;  /* render every tile_stride-th 16x16 tile (1 = all) */
  int32_t threads;      /* 0 = OpenMP default */
  int32_t brute;        /* 1 = brute-force intersection instead of the BVH */
} oracle_params;

typedef struct {
  uint64_t radiance_rays, shadow_rays, paths;
  double seconds;
  int32_t threads;
} oracle_stats;

/* Renders into full-frame float4 films (width*height*4 floats each); only the
 * selected tiles are written. Films accumulate as running means, so passes
 * can be chained like the GPU's progressive passes. */
int oracle_render(const sundog_flat_scene* s, const oracle_params* p, float* accum,
                  float* albedo, float* normal, oracle_stats* st);

/* Closest hit of n rays (8 floats each: ox,oy,oz,tmin,dx,dy,dz,tmax). */
int oracle_trace(const sundog_flat_scene* s, const float* rays, int n, int brute,
                 float* t_out, int32_t* inst_out, int32_t* prim_out);
/* Shadow transmittance (n x 3 floats). */
int oracle_shadow(const sundog_flat_scene* s, const float* rays, int n, float* tr_out);
/* Environment CDF bins for n (u0, u1) pairs. */
int oracle_env_bins(const sundog_flat_scene* s, const float* u, int n, int32_t* row, int32_t* col);
/* Environment CDFs as built by the oracle (sizes h+1 and h*(w+1)). */
int oracle_env_cdf(const sundog_flat_scene* s, float* marg, float* cond);
/* Raw PCG32 draws of stream init(seed, seq). */
void oracle_pcg32(uint64_t seed, uint64_t seq, int n, uint32_t* out);
/* BSDF probe: sample / eval / pdf for a unit-frame configuration. */
int oracle_bsdf(const sundog_flat_material* m, const float albedo[3], const float wo[3],
                const float wi[3], float eta, float f_out[3], float* pdf_out);
int oracle_bsdf_sample(const sundog_flat_material* m, const float albedo[3], const float raydir[3],
                       int frontface, float eta_ext, uint64_t seed, uint64_t seq,
                       float wi_out[3], float weight_out[3], float* pdf_out, int* delta_out,
                       int* valid_out, int* draws_out);
/* Flame transmittance probe on the flat scene's flames. */
float oracle_flame_transmittance(const sundog_flat_scene* s, const float o[3], const float d[3],
                                 float tend, uint64_t seed, uint64_t seq, int skip, int* draws_out);
/* Environment sample for fixed uniforms (u0, u1): direction, its pdf, and
 * the pdf the MIS side recomputes for that direction. */
int oracle_env_sample(const sundog_flat_scene* s, float u0, float u1, float wi_out[3],
                      float* pdf_out, float* pdf_back_out);
/* Emission + transmittance march over the flat scene's flames. */
float oracle_march_flames(const sundog_flat_scene* s, const float o[3], const float d[3], float tend,
                          uint64_t seed, uint64_t seq, float trans_out[3], float L_out[3],
                          int* draws_out);
/* Texture `tex` of the flat scene at n (u, v) pairs: n x 4 floats. */
int oracle_texture(const sundog_flat_scene* s, int tex, const float* uv, int n, float* out);
/* The bilinear image filter alone (wrap in u; wrap or clamp in v). */
void oracle_bilinear(const float* img, int w, int h, const float* uv, int n, int wrap_v, float* out);
/* Value noise (octaves = 0) or fbm of `octaves` octaves at n points. */
void oracle_noise(const float* p, int n, uint32_t seed, int octaves, float* out);
/* The water material's perturbed shading normal. */
void oracle_water_normal(const float x[3], const float ns[3], const float wo[3], float amp, float freq,
                         int octaves, float out[3]);
/* Flame density and emission at n points (sigma n, emission n x 3). */
int oracle_flame_field(const sundog_flat_scene* s, int flame, const float* p, int n, float* sigma_out,
                       float* emission_out);
/* Mesh-light samples for n fixed uniform triples (pick, r0, r1) seen from x:
 * direction, distance, solid-angle pdf (0 when invalid), picked triangle, and
 * the light's total world-space area. */
int oracle_mesh_light_sample(const sundog_flat_scene* s, int light, const float x[3], const float* u, int n,
                             float* wi_out, float* dist_out, float* pdf_out, int32_t* tri_out, float* area_out);
/* Light samples of light `light` from n points xs (n x 3); sample i draws
 * from init(seed0 + i, seq). Outputs: wi (n x 3), dist, Li (n x 3), pdf (0 when
 * invalid), flags (bit 0 valid, bit 1 delta), draws taken, and the MIS-side
 * solid-angle pdf at the sampled point (NaN for mesh and delta lights). */
int oracle_light_sample(const sundog_flat_scene* s, int light, const float* xs, int n, uint64_t seed0,
                        uint64_t seq, float* wi_out, float* dist_out, float* li_out, float* pdf_out,
                        int32_t* flags_out, int32_t* draws_out, float* pdf_back_out);
/* MIS-side light pdf for n (from, hit point, hit normal) triples (n x 9). */
int oracle_light_pdf(const sundog_flat_scene* s, int light, const float* q, int n, float* pdf_out);
/* Camera jitter (jx, jy) in [0,1)^2 of sample s of pixel (px, py). */
int oracle_camera_jitter(const sundog_flat_scene* s, int px, int py, int sample, float* jx, float* jy);
/* Last error message. */
const char* oracle_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
