// Oracle: axis-aligned boxes, triangle bounds, SAH cost, BVH validation and
// the ray/box and ray/triangle tests, restated from the reference's
// crates/core/src/bvh.rs (Aabb :44-160, safe_inv_dir :165, triangle_bounds
// :202, sah_cost :469, compact_primitives :517, validate :544,
// intersect_triangle :930, brute_force_intersect :974, intersect_with_stats
// :1045, occluded :1137).
//
// TEST INFRASTRUCTURE: CPU oracle only; never linked into the product.
#pragma once
#include <cstdint>
#include <limits>
#include <string>
#include <vector>

#include "../include/lucida.h"
#include "vec.hpp"

namespace oracle {

struct Aabb {
    Vec3 min = Vec3::splat(std::numeric_limits<float>::infinity());
    Vec3 max = Vec3::splat(-std::numeric_limits<float>::infinity());

    void grow_point(Vec3 p) { min = vmin(min, p); max = vmax(max, p); }
    void grow(const Aabb& o) { min = vmin(min, o.min); max = vmax(max, o.max); }
    bool is_empty() const { return min.x > max.x; }
    float surface_area() const {
        if (is_empty()) return 0.0f;
        Vec3 d = max - min;
        return 2.0f * (d.x * d.y + d.y * d.z + d.z * d.x);
    }
    Vec3 centroid() const { return (min + max) * 0.5f; }
    // bvh.rs:100 — grow each axis by 1e-6 * (1 + max(|lo|, |hi|)).
    void pad() {
        if (is_empty()) return;
        const float RELATIVE = 1.0e-6f;
        for (int a = 0; a < 3; ++a) {
            float scale = 1.0f + std::fmax(std::fabs(min[a]), std::fabs(max[a]));
            float p = RELATIVE * scale;
            min[a] -= p;
            max[a] += p;
        }
    }
    bool contains(const Aabb& o, float eps) const {
        if (o.is_empty()) return true;
        return min.x <= o.min.x + eps && min.y <= o.min.y + eps && min.z <= o.min.z + eps &&
               max.x >= o.max.x - eps && max.y >= o.max.y - eps && max.z >= o.max.z - eps;
    }
    // bvh.rs:127 — slab test; returns entry distance or a negative miss flag
    // through `hit`.
    bool hit(Vec3 origin, Vec3 inv_dir, float t_min, float t_max, float* entry_out) const {
        Vec3 t0 = (min - origin) * inv_dir;
        Vec3 t1 = (max - origin) * inv_dir;
        Vec3 nr = vmin(t0, t1);
        Vec3 fr = vmax(t0, t1);
        float entry = std::fmax(max_element(nr), t_min);
        float exit = std::fmin(min_element(fr), t_max);
        if (entry <= exit) {
            *entry_out = entry;
            return true;
        }
        return false;
    }
};

inline Vec3 safe_inv_dir(Vec3 d) {
    const float BIG = 1.0e16f, SMALL = 1.0f / 1.0e16f;
    auto g = [&](float v) { return std::fabs(v) > SMALL ? 1.0f / v : std::copysign(BIG, v); };
    return Vec3(g(d.x), g(d.y), g(d.z));
}

inline Vec3 pos3(const float* positions, uint32_t i) {
    return Vec3(positions[4 * (size_t)i], positions[4 * (size_t)i + 1], positions[4 * (size_t)i + 2]);
}

inline Aabb triangle_bounds(const GpuTriangle& t, const float* positions) {
    Aabb b;
    b.grow_point(pos3(positions, t.i0));
    b.grow_point(pos3(positions, t.i1));
    b.grow_point(pos3(positions, t.i2));
    b.pad();
    return b;
}

inline Aabb node_aabb(const GpuBvhNode& n) {
    Aabb b;
    b.min = Vec3::from(n.bounds_min);
    b.max = Vec3::from(n.bounds_max);
    return b;
}

struct BuildStats {
    size_t triangles = 0, nodes = 0, leaves = 0, max_depth = 0, max_leaf_size = 0;
    float mean_leaf_size = 0.0f;
    float sah_cost = 0.0f;
};

// bvh.rs:469 — traversal cost 1 per internal node, 1 per primitive in a leaf,
// weighted by area relative to the root.
float sah_cost(const std::vector<GpuBvhNode>& nodes);
// bvh.rs:544 — structural problems, empty when the tree is sound.
std::vector<std::string> validate(const std::vector<GpuBvhNode>& nodes, const std::vector<uint32_t>& prim_indices,
                                  size_t triangle_count);

struct Ray {
    Vec3 origin, dir;
};

struct TriHit {
    float t, u, v;
    uint32_t triangle;
};

struct TraversalStats {
    uint32_t node_visits = 0, triangle_tests = 0;
};

// bvh.rs:930 — double-sided Moller-Trumbore.
bool intersect_triangle(Vec3 p0, Vec3 p1, Vec3 p2, const Ray& ray, float t_min, float t_max, float* t, float* u,
                        float* v);

// bvh.rs:1045 / :1137 over a compacted tree (prim_indices given explicitly).
struct BvhView {
    const GpuBvhNode* nodes = nullptr;
    size_t num_nodes = 0;
    const uint32_t* prim_indices = nullptr;  // nullptr: identity (compacted)
    const GpuTriangle* triangles = nullptr;
    size_t num_triangles = 0;
    const float* positions = nullptr;
};
bool bvh_intersect(const BvhView& b, uint32_t root, const Ray& ray, float t_min, float t_max, TriHit* out,
                   TraversalStats* stats);
bool bvh_occluded(const BvhView& b, uint32_t root, const Ray& ray, float t_min, float t_max);
bool brute_force_intersect(const BvhView& b, const Ray& ray, float t_min, float t_max, TriHit* out);

}  // namespace oracle
