/*
 * oracle.h - C ABI of the CPU oracle (test infrastructure only).
 *
 * The oracle is a scalar C++ restatement of the reference's trace kernel
 * (src/render/trace.slang:195-1309), templated on the arithmetic type:
 * float gives the FP32 parity oracle, double the FP64 convergence oracle.
 * It is the checker of the HIP kernels; no product code calls it.
 */
#ifndef PT_ORACLE_H
#define PT_ORACLE_H

#include <stdint.h>

#include "../include/pt_api.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Per-render counters (summed over all paths). */
typedef struct oracle_stats {
    uint64_t closest_rays;  /* closest-hit traversals */
    uint64_t shadow_rays;   /* any-hit traversals */
    uint64_t pair_tests;    /* child pairs box-tested (64 B each) */
    uint64_t tri_tests;     /* triangle tests (36 B each) */
} oracle_stats;

/* Trace `samples` samples of a tile starting at 1-based sample index
 * params->sample_count, accumulating into accum (tile_w*tile_h float4) with
 * the reference's rule acc = (sample_count > 1 ? acc : 0) + L. If storage is
 * non-NULL it receives EXPOSURE * acc / sample_count after the last sample.
 * threads <= 0 uses all OpenMP threads. stats may be NULL. */
int oracle_render_f32(const pt_scene_desc* scene, const PtTraceParameters* params,
                      uint32_t tile_w, uint32_t tile_h, uint32_t samples,
                      float* accum, float* storage, int threads, oracle_stats* stats);
/* Same in FP64 (accum is double). */
int oracle_render_f64(const pt_scene_desc* scene, const PtTraceParameters* params,
                      uint32_t tile_w, uint32_t tile_h, uint32_t samples,
                      double* accum, int threads, oracle_stats* stats);
/* Trace only the rows [row_begin, row_end) of the tile (bounded CPU samples). */
int oracle_render_rows_f32(const pt_scene_desc* scene, const PtTraceParameters* params,
                           uint32_t tile_w, uint32_t tile_h, uint32_t row_begin, uint32_t row_end,
                           uint32_t samples, float* accum, int threads, oracle_stats* stats);

/* Primary-ray closest hit per tile pixel for sample index params->sample_count. */
int oracle_primary_hits(const pt_scene_desc* scene, const PtTraceParameters* params,
                        uint32_t tile_w, uint32_t tile_h, int32_t* tri, float* t);

/* Closest hit by BVH traversal for rays (origin xyz, dir xyz per ray). */
int oracle_intersect(const pt_scene_desc* scene, const float* rays, uint32_t count,
                     int32_t* hit, float* t, float* v, float* w,
                     uint32_t* pair_tests, uint32_t* tri_tests);
/* Closest hit by testing every triangle in index order (no BVH). */
int oracle_intersect_brute(const pt_scene_desc* scene, const float* rays, uint32_t count,
                           int32_t* hit, float* t);
/* Any-hit by BVH traversal, skipping triangle skip[i] (-1: none). */
int oracle_occluded(const pt_scene_desc* scene, const float* rays, const float* tmax,
                    const int32_t* skip, uint32_t count, uint8_t* occluded);

/* Individual kernel functions, exposed for closed-form tests (FP32). */
float oracle_fresnel_exact(float ior, float cos_i);
float oracle_ggx_lambda(const float v[3], float ax, float ay);
void oracle_sample_vndf(const float wo[3], float ax, float ay, float r, float two_pi_r, float h_out[3]);
float oracle_e_fon_exact(float mu, float r);
/* The GGX distribution D(h) of evaluate_lobes (trace.slang:249-250). */
float oracle_ggx_d(const float h[3], float ax, float ay);
/* estimate_e_dielectric (trace.slang:396-411) with 1 + Lambda(wo) computed
 * from wo; draws 32 random numbers from *rng, which it advances. */
void oracle_estimate_e_dielectric(const float wo[3], float ax, float ay, float specular_ior_modulated,
                                  const float specular_color[3], uint32_t* rng, float out[3]);
/* alpha_anisotropic (trace.slang:369-375). */
void oracle_alpha_anisotropic(float roughness, float anisotropy, float* ax, float* ay);
/* evaluate_skybox: bilinear sky fetch times average luminance (trace.slang:473-475). */
void oracle_evaluate_skybox(const pt_scene_desc* scene, const float dir[3], float out[3]);
/* The sRGB decode table (Rgba8UnormSrgb) the oracle's textures use. */
float oracle_srgb_decode(uint32_t byte);
/* orthonormal_tangent (trace.slang:207-218). */
void oracle_orthonormal_tangent(const float tangent[3], const float normal[3], float out[3]);
/* The atlas fetch: nearest texel, clamp-to-edge, LOD 0 (world_resources.rs:447-450). */
void oracle_sample_atlas(const uint8_t* atlases, uint32_t size, int layer, float u, float v, int srgb, float out[4]);
/* The primary ray direction of pixel (px, py) (trace.slang:726-748); draws
 * the two jitter numbers from *rng. */
void oracle_primary_dir(const pt_scene_desc* scene, uint32_t px, uint32_t py, uint32_t iw, uint32_t ih,
                        uint32_t* rng, float out[3]);
float oracle_specular_ior_modulated(float ior, float weight);
float oracle_power_heuristic(float pdf, float other);
/* evaluate_lobes for a material (base_color rgb, base_weight, metalness,
 * diffuse roughness, specular rgb, specular weight, specular_ior,
 * spec roughness, spec roughness anisotropy) at wo, wi; E_dielectric given.
 * Writes f[3], pdf. */
void oracle_evaluate_lobes(const float params[13], const float wo[3], const float wi[3],
                           const float e_dielectric[3], float f_out[3], float* pdf_out);
/* Sample a BSDF direction like the bounce does (rands r0 = two_pi_rand/2pi,
 * r1, r2); writes wi and returns 1, or 0 if the sample ends the path. */
int oracle_sample_bsdf(const float params[13], const float wo[3], const float e_dielectric[3],
                       float r0, float r1, float r2, float wi_out[3]);
/* Sample a light from `origin` with RNG state *rng; writes direction,
 * distance, emission, pdf; returns the triangle index (-1: sky). */
int oracle_sample_light(const pt_scene_desc* scene, const float origin[3], uint32_t* rng,
                        float dir_out[3], float* distance, float emission_out[3], float* pdf);
float oracle_skybox_light_pdf(const pt_scene_desc* scene, const float dir[3], float category_probability);
float oracle_triangle_light_pdf(const pt_scene_desc* scene, uint32_t tri, const float dir[3],
                                float distance_squared, float probability);
float oracle_rand(uint32_t* state);
uint32_t oracle_seed(uint32_t px, uint32_t py, uint32_t sample_count);
/* The deterministic FP32 libm shared with the kernels (pt_math.h), over
 * arrays: fn 0 sin, 1 cos, 2 acos, 3 atan2(x, y), 4 pow(x, y), 5 log2, 6 exp2. */
int oracle_math(int fn, const float* x, const float* y, float* out, uint32_t n);
/* ACES + sRGB encode of one linear value into [0, 1] (display.wgsl:19-46). */
float oracle_display_channel(float linear);

#ifdef __cplusplus
}
#endif

#endif /* PT_ORACLE_H */
