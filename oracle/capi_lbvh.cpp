// C entry points of the CPU oracle's LBVH and BVH utilities.
// TEST INFRASTRUCTURE: CPU oracle only.
#include <algorithm>
#include <cstring>
#include <numeric>

#include "lbvh.hpp"
#include "oracle.h"

using namespace oracle;

static void store_aabb(const Aabb& a, LucidaAabb* o) {
    a.min.store(o->lo);
    a.max.store(o->hi);
    o->_p0 = 0.0f;
    o->_p1 = 0.0f;
}

extern "C" {

uint32_t oracle_expand_bits(uint32_t v) { return expand_bits(v); }
uint32_t oracle_morton3d(float x, float y, float z) { return morton3d(Vec3(x, y, z)); }
void oracle_normalise(const float p[3], const float lo[3], const float hi[3], float out[3]) {
    Aabb b;
    b.min = Vec3::from(lo);
    b.max = Vec3::from(hi);
    normalise(Vec3::from(p), b).store(out);
}
int oracle_lcp(const uint32_t* codes, size_t n, int i, int j) { return longest_common_prefix(codes, n, i, j); }

void oracle_triangle_bounds(const GpuTriangle* tris, size_t n, const float* positions, LucidaAabb* out) {
    for (size_t i = 0; i < n; ++i) store_aabb(triangle_bounds(tris[i], positions), &out[i]);
}

size_t oracle_lbvh_build(const GpuTriangle* tris, size_t n, const float* positions, size_t max_leaf,
                         uint32_t* codes, uint32_t* order, uint32_t* karras, uint32_t* parent,
                         LucidaAabb* node_bounds, uint32_t* subtree, LucidaAabb* scene_bounds, GpuBvhNode* nodes,
                         uint32_t* prim_indices, float* sah) {
    LbvhResult r;
    lbvh_build(tris, n, positions, max_leaf, r);
    if (n == 0) return 0;
    if (codes) std::memcpy(codes, r.codes.data(), n * 4);
    if (order) std::memcpy(order, r.order.data(), n * 4);
    if (karras && !r.karras.empty()) std::memcpy(karras, r.karras.data(), r.karras.size() * 4);
    if (parent) std::memcpy(parent, r.parent.data(), r.parent.size() * 4);
    if (node_bounds)
        for (size_t i = 0; i < r.node_bounds.size(); ++i) store_aabb(r.node_bounds[i], &node_bounds[i]);
    if (subtree && !r.subtree.empty()) std::memcpy(subtree, r.subtree.data(), r.subtree.size() * 4);
    if (scene_bounds) store_aabb(r.scene_bounds, scene_bounds);
    if (nodes) std::memcpy(nodes, r.nodes.data(), r.nodes.size() * sizeof(GpuBvhNode));
    if (prim_indices) std::memcpy(prim_indices, r.prim_indices.data(), r.prim_indices.size() * 4);
    if (sah) *sah = r.stats.sah_cost;
    return r.nodes.size();
}

float oracle_sah_cost(const GpuBvhNode* nodes, size_t n) { return sah_cost(std::vector<GpuBvhNode>(nodes, nodes + n)); }

size_t oracle_validate(const GpuBvhNode* nodes, size_t n, const uint32_t* prim_indices, size_t num_prims,
                       size_t triangle_count) {
    return validate(std::vector<GpuBvhNode>(nodes, nodes + n),
                    std::vector<uint32_t>(prim_indices, prim_indices + num_prims), triangle_count)
        .size();
}

void oracle_sort_pairs(const uint32_t* keys, const uint32_t* vals, size_t n, uint32_t* keys_out, uint32_t* vals_out) {
    std::vector<std::pair<uint32_t, uint32_t>> v(n);
    for (size_t i = 0; i < n; ++i) v[i] = {keys[i], vals[i]};
    std::stable_sort(v.begin(), v.end(), [](auto& a, auto& b) { return a.first < b.first; });
    for (size_t i = 0; i < n; ++i) {
        keys_out[i] = v[i].first;
        vals_out[i] = v[i].second;
    }
}

}  // extern "C"
