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
 *   "bvh-stress"    bvh_stress() (scenes.rs:218): 72 level-4 icospheres
 *                   (368,640 triangles) in the analytic Cornell box
 *   "furnace-test"  furnace_test() (scenes.rs:275): white metal spheres of
 *                   roughness 0..1 under a uniform environment of 1
 *   "metal-sweep"   metal_sweep() (scenes.rs:327): copper and gold conductors
 *   "mis-scene"     mis_scene() (scenes.rs:410): Veach's plates and emitters
 *   "glass-box"     glass_box() (scenes.rs:536): clear, diamond, frosted and
 *                   tinted dielectrics in a Cornell box
 *   "sunset"        sunset() (scenes.rs:651): lit only by procedural_sky
 *                   (envmap.rs:343), 256 x 128
 *   "instance-forest" instance_forest() (scenes.rs:733): 49 placements of one
 *                   icosphere under a procedural sky (instanced)
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

/* Camera (camera.rs:14-26): eye, look-at point, up hint, vertical field of
 * view in degrees, lens aperture (diameter), focus distance. */
typedef struct lucida_camera {
    float eye[3];
    float look_at[3];
    float up[3];
    float vfov_deg;
    float aperture;
    float focus_distance;
} lucida_camera;

/* Camera::write_uniforms (camera.rs:55): writes cam_origin, cam_upper_left,
 * cam_horizontal, cam_vertical, lens_radius and focus_distance of *u for an
 * image of the given aspect (width / height); other fields are untouched. */
lucida_status lucida_camera_write_uniforms(const lucida_camera* cam, float aspect, GpuUniforms* u);

/* The camera a built-in scene carries (SceneDef::camera, scenes.rs:11-18). */
lucida_status lucida_blob_camera(const lucida_blob* blob, lucida_camera* out);

/* Replaces the blob's environment map (RGBA32F, w*h) and builds its 2D CDF
 * (envmap.rs:112 EnvMap::new). */
lucida_status lucida_blob_set_env(lucida_blob* blob, uint32_t width, uint32_t height, const float* rgba);

#ifdef __cplusplus
}
#endif

#endif /* LUCIDA_SCENES_H */
