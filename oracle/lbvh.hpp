// Oracle: linear BVH builder (crates/core/src/lbvh.rs). TEST INFRASTRUCTURE.
#pragma once
#include <vector>

#include "bvh.hpp"

namespace oracle {

uint32_t expand_bits(uint32_t v);
uint32_t morton3d(Vec3 p);
Vec3 normalise(Vec3 p, const Aabb& b);
int longest_common_prefix(const uint32_t* codes, size_t n, int i, int j);
void build_hierarchy(const uint32_t* codes, size_t n, std::vector<uint32_t>& karras, std::vector<uint32_t>& parent);

struct LbvhResult {
    Aabb scene_bounds;
    std::vector<uint32_t> codes, order, karras, parent, subtree;
    std::vector<Aabb> node_bounds;
    std::vector<GpuBvhNode> nodes;
    std::vector<uint32_t> prim_indices;
    BuildStats stats;
};

void lbvh_build(const GpuTriangle* tris, size_t n, const float* positions, size_t max_leaf, LbvhResult& out);

}  // namespace oracle
