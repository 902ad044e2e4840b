// Host-side scene blob: the reference's SceneBlob (gpu_layout.rs:706) plus the
// camera and environment map a SceneDef carries (scenes.rs:11-18).
#pragma once
#include <vector>

#include "../../include/lucida.h"
#include "hvec.hpp"

namespace lucida {

struct Blob {
    std::vector<GpuMaterial> materials;
    std::vector<GpuPrimitive> primitives;
    std::vector<GpuInstance> instances;
    std::vector<float> positions;  // vec4 per vertex
    std::vector<GpuVertexAttr> vertex_attrs;
    std::vector<GpuTriangle> triangles;
    std::vector<GpuBvhNode> bvh_nodes;
    std::vector<GpuLight> lights;
    uint32_t tlas_root = 0;
    // Instanced scenes: meshes and placements, built into BLAS/TLAS at upload.
    std::vector<lucida_mesh_range> meshes;
    std::vector<lucida_instance_placement> placements;

    uint32_t env_width = 0, env_height = 0;
    std::vector<float> env_pixels, env_conditional, env_marginal;
    float env_total_weight = 0.0f;

    f3 cam_eye, cam_look_at, cam_up{0, 1, 0};
    float cam_vfov_deg = 40.0f, cam_aperture = 0.0f, cam_focus_distance = 1.0f;
    f3 background;
};

}  // namespace lucida
