/* Oracle: the CPU reference tracer's C entry points (TEST INFRASTRUCTURE). */
#ifndef LUCIDA_ORACLE_TRACER_H
#define LUCIDA_ORACLE_TRACER_H

#include <stddef.h>
#include <stdint.h>

#include "../include/lucida.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct RenderCounts {
    uint64_t extension_rays, shadow_rays, node_visits, triangle_tests;
} RenderCounts;

uint32_t oracle_sobol(uint32_t index, uint32_t dim);
uint32_t oracle_owen_scramble(uint32_t x, uint32_t seed);
uint32_t oracle_hash_u32(uint32_t x);
float oracle_sobol_sample(uint32_t index, uint32_t dim, uint32_t seed);
/* Draws from a Sampler (sobol.rs:287-361): pattern[i] = 1 for next_f32, 2 for
 * next_vec2 (two outputs). */
void oracle_sampler_draws(uint32_t kind, uint32_t frame, uint32_t pixel, uint32_t sample, const uint32_t* pattern,
                          size_t n, float* out);
/* energy::compute_table (bsdf.rs:1200-1227), 32 x 32, row = roughness. */
void oracle_ggx_table(float* out);
float oracle_directional_albedo_grid(float roughness, float cos_o, uint32_t grid);
void oracle_offset_ray_origin(const float p[3], const float n[3], float out[3]);

/* integrator::render_with_guides restricted to rows [row_begin, row_end):
 * rgb receives the per-pixel mean (3 floats per pixel), guides (optional) 8
 * floats per pixel: albedo[3], normal[3], depth, steps. The scene's BVH must
 * be compacted (leaves index triangles directly) or absent (brute force). */
void oracle_render(const lucida_scene_desc* d, const GpuUniforms* u, const float background[3], uint32_t threads,
                   uint32_t row_begin, uint32_t row_end, float* rgb, float* guides, RenderCounts* counts);

/* The same render over an explicit list of image rows (each row one work
 * item of the thread pool); rgb receives row k of the output for image row
 * rows[k] (3 floats per pixel). */
void oracle_render_rows(const lucida_scene_desc* d, const GpuUniforms* u, const float background[3], uint32_t threads,
                        const uint32_t* rows, uint32_t n_rows, float* rgb, RenderCounts* counts);

/* One ray per 8 floats (origin.xyz, unused, direction.xyz, t_max) against the
 * scene without shading: hits receive (t, u, v, id bits) per ray, id =
 * triangle, 0x80000000 | analytic primitive, or 0xFFFFFFFF for a miss;
 * instances (optional) the instance primitive index of a triangle hit in an
 * instanced scene. With any_hit, id is 0 for an occluded ray. */
void oracle_scene_cast(const lucida_scene_desc* d, const float* rays, size_t n, int any_hit, float* hits,
                       uint32_t* instances);

/* denoise::denoise (denoise.rs:134) on a film and its guides (3 floats per
 * pixel for film/albedo/normal/out, 1 for depth). */
void oracle_denoise(const float* film, const float* albedo, const float* normal, const float* depth, uint32_t width,
                    uint32_t height, const lucida_denoise_params* p, float* out);

/* bsdf.rs:314 / :358 / :555 closed forms, and the reflected / transmitted
 * halves of a surface's directional albedo (mean luminance weight of
 * surface::sample split by the side wi lands on). */
float oracle_fresnel_dielectric(float cos_i, float eta);
void oracle_fresnel_conductor(float cos_i, const float eta[3], const float k[3], float out[3]);
int oracle_refract(const float wo[3], const float n[3], float eta, float out[3]);
void oracle_bsdf_split_albedo(const GpuMaterial* m, uint32_t front, const float wo[3], uint32_t n, uint32_t seed,
                              double out[2]);

/* envmap.rs: the oracle's own CDF build (EnvMap::new, returns total_weight;
 * conditional holds (w+1)*h floats, marginal h+1), pdf/radiance at a
 * direction, one sample (out = direction[3], radiance[3], returns pdf), and
 * the uv <-> direction maps. */
float oracle_env_build(uint32_t width, uint32_t height, const float* pixels, float* conditional, float* marginal);
float oracle_env_pdf(uint32_t w, uint32_t h, const float* pixels, const float* cond, const float* marg, float total,
                     const float dir[3], float radiance_out[3]);
float oracle_env_sample(uint32_t w, uint32_t h, const float* pixels, const float* cond, const float* marg, float total,
                        float u, float v, float out[6]);
void oracle_env_direction_from_uv(float u, float v, float out[3]);
void oracle_env_uv_from_direction(const float d[3], float out[2]);

/* The convergence meter (shaders/stats/convergence.wgsl) over a film of
 * GpuAccum records (16 floats per pixel). */
double oracle_convergence(const float* accum, size_t pixels);

/* image::compare (image.rs:178): out = {rmse, max_abs, mean_rel, mean_a,
 * mean_b, mean_abs}. */
void oracle_compare(const float* a, const float* b, size_t pixels, double* out6);

#ifdef __cplusplus
}
#endif
#endif
