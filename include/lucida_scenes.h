/*
 * lucida_scenes.h — host-side scene generators and scene-description helpers.
 *
 * These build the inputs the hot path consumes: a SceneBlob laid out per
 * crates/core/src/gpu_layout.rs, the light list (light.rs:62 build_lights) and
 * the camera uniforms (camera.rs:55 Camera::write_uniforms, integrator.rs:153
 * build_uniforms). They replace the scene side of the reference's
 * crates/core/src/scenes.rs and mesh.rs for the workloads this build measures.
 * No GPU is touched here.
 */
#ifndef LUCIDA_SCENES_H
#define LUCIDA_SCENES_H

#include "lucida.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lucida_blob lucida_blob;

/* Generator knobs. Zero fields take each scene's documented default. */
typedef struct lucida_blob_params {
    uint32_t seed;       /* material / displacement seed                        */
    uint32_t frequency;  /* geodesic frequency of the displaced sphere (20 f^2 tris) */
    float displacement;  /* radial noise amplitude of the displaced sphere       */
} lucida_blob_params;

/* Scene names:
 *   "cornell-box"   the reference's cornell_box() (scenes.rs:56), analytic
 *   "cornell-mesh"  the reference's cornell_mesh() (scenes.rs:174), triangles
 *                   in generation order (no BVH, no compaction)
 *   "box-c1"        the 1x1x1 triangle-quad Cornell box with two spheres
 *   "sphere-c2"     the noise-displaced geodesic sphere in a 1,024-triangle
 *                   room lit by 16 emissive quads
 *   "icosphere-L"   a bare icosphere of subdivision level L (mesh.rs:169)
 *   "instances-c3"  4,096 placements of a 9,680-triangle displaced sphere
 *                   (frequency 22) in a 100^3 box, 30% rough dielectric, lit
 *                   by a 2048 x 1024 log-normal environment with three hot spots
 *   "instances-small" the same recipe at 64 placements of a 320-triangle
 *                   sphere (frequency 4) in a 10^3 box, environment 256 x 128
 * Instanced scenes carry meshes and placements (lucida_blob_instancing) and
 * are uploaded with lucida_scene_create_instanced. */
lucida_status lucida_blob_create(const char* name, const lucida_blob_params* params, lucida_blob** out);
void lucida_blob_destroy(lucida_blob* blob);

/* Pointers into the blob's arrays (valid until it is destroyed or modified). */
lucida_status lucida_blob_desc(const lucida_blob* blob, lucida_scene_desc* out);

/* build_uniforms (integrator.rs:153) for a render of this scene: camera,
 * counts, background and environment fields. The caller sets sample_offset,
 * samples_per_launch, max_depth, frame_seed, sampling_mode and sampler_kind. */
lucida_status lucida_blob_uniforms(const lucida_blob* blob, uint32_t width, uint32_t height, GpuUniforms* out);

/* Installs a BVH over the blob's triangles: nodes (compacted, leaves index
 * triangles directly) and the triangles in leaf order. */
lucida_status lucida_blob_set_bvh(lucida_blob* blob, const GpuBvhNode* nodes, size_t num_nodes,
                                  const GpuTriangle* triangles_in_leaf_order);

/* The meshes and placements of an instanced scene (empty otherwise). */
lucida_status lucida_blob_instancing(const lucida_blob* blob, const lucida_mesh_range** meshes, size_t* num_meshes,
                                     const lucida_instance_placement** placements, size_t* num_placements);

/* Replaces the blob's environment map (RGBA32F, w*h) and builds its 2D CDF
 * (envmap.rs:112 EnvMap::new). */
lucida_status lucida_blob_set_env(lucida_blob* blob, uint32_t width, uint32_t height, const float* rgba);

#ifdef __cplusplus
}
#endif

#endif /* LUCIDA_SCENES_H */
