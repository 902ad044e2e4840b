// Oracle: BVH utilities and traversal, restated from the reference's
// crates/core/src/bvh.rs. TEST INFRASTRUCTURE: CPU oracle only.
#include "bvh.hpp"

#include <cstdio>

namespace oracle {

// bvh.rs:469-493
float sah_cost(const std::vector<GpuBvhNode>& nodes) {
    if (nodes.empty()) return 0.0f;
    float root_area = node_aabb(nodes[0]).surface_area();
    if (root_area <= 0.0f) return 0.0f;
    float total = 0.0f;
    for (const auto& n : nodes) {
        float area = node_aabb(n).surface_area();
        float own = n.count == 0 ? 1.0f : (float)n.count * 1.0f;
        total += area / root_area * own;
    }
    return total;
}

// bvh.rs:544-606
std::vector<std::string> validate(const std::vector<GpuBvhNode>& nodes, const std::vector<uint32_t>& prim_indices,
                                  size_t triangle_count) {
    std::vector<std::string> problems;
    char buf[256];
    if (nodes.empty()) {
        if (triangle_count > 0) problems.push_back("triangles but no nodes");
        return problems;
    }
    std::vector<uint32_t> seen(triangle_count, 0);
    for (size_t i = 0; i < nodes.size(); ++i) {
        const auto& node = nodes[i];
        if (node.count == 0) {
            size_t l = node.left_first;
            if (l + 1 >= nodes.size()) {
                snprintf(buf, sizeof buf, "node %zu has out-of-range children at %zu", i, l);
                problems.push_back(buf);
                continue;
            }
            Aabb parent = node_aabb(node);
            float eps = 1e-3f * std::fmax(std::sqrt(parent.surface_area()), 1.0f);
            for (size_t c : {l, l + 1}) {
                if (!parent.contains(node_aabb(nodes[c]), eps)) {
                    snprintf(buf, sizeof buf, "node %zu does not contain child %zu", i, c);
                    problems.push_back(buf);
                }
            }
        } else {
            size_t start = node.left_first, end = start + node.count;
            if (end > prim_indices.size()) {
                snprintf(buf, sizeof buf, "leaf %zu range %zu..%zu is out of bounds", i, start, end);
                problems.push_back(buf);
                continue;
            }
            for (size_t k = start; k < end; ++k) {
                uint32_t pi = prim_indices[k];
                if (pi >= triangle_count) {
                    snprintf(buf, sizeof buf, "leaf %zu references triangle %u, out of range", i, pi);
                    problems.push_back(buf);
                } else {
                    seen[pi] += 1;
                }
            }
        }
    }
    size_t missing = 0, dup = 0;
    for (auto c : seen) {
        if (c == 0) ++missing;
        if (c > 1) ++dup;
    }
    if (missing) problems.push_back(std::to_string(missing) + " triangles are in no leaf");
    if (dup) problems.push_back(std::to_string(dup) + " triangles appear in more than one leaf");
    return problems;
}

// bvh.rs:930-966
bool intersect_triangle(Vec3 p0, Vec3 p1, Vec3 p2, const Ray& ray, float t_min, float t_max, float* t_out,
                        float* u_out, float* v_out) {
    Vec3 e1 = p1 - p0;
    Vec3 e2 = p2 - p0;
    Vec3 pv = cross(ray.dir, e2);
    float det = dot(e1, pv);
    if (std::fabs(det) < 1.0e-12f) return false;
    float inv_det = 1.0f / det;
    Vec3 tv = ray.origin - p0;
    float u = dot(tv, pv) * inv_det;
    if (!(u >= 0.0f && u <= 1.0f)) return false;
    Vec3 qv = cross(tv, e1);
    float v = dot(ray.dir, qv) * inv_det;
    if (v < 0.0f || u + v > 1.0f) return false;
    float t = dot(e2, qv) * inv_det;
    if (t < t_min || t > t_max) return false;
    *t_out = t;
    *u_out = u;
    *v_out = v;
    return true;
}

static inline void tri_pos(const BvhView& b, uint32_t tri, Vec3* p0, Vec3* p1, Vec3* p2) {
    const GpuTriangle& t = b.triangles[tri];
    *p0 = pos3(b.positions, t.i0);
    *p1 = pos3(b.positions, t.i1);
    *p2 = pos3(b.positions, t.i2);
}

// bvh.rs:974-995
bool brute_force_intersect(const BvhView& b, const Ray& ray, float t_min, float t_max, TriHit* out) {
    bool any = false;
    float closest = t_max;
    for (size_t i = 0; i < b.num_triangles; ++i) {
        Vec3 p0, p1, p2;
        tri_pos(b, (uint32_t)i, &p0, &p1, &p2);
        float t, u, v;
        if (intersect_triangle(p0, p1, p2, ray, t_min, closest, &t, &u, &v)) {
            closest = t;
            *out = TriHit{t, u, v, (uint32_t)i};
            any = true;
        }
    }
    return any;
}

static const size_t MAX_STACK = 64;  // bvh.rs MAX_STACK

// bvh.rs:1045-1130, generalised to start at `root` (the GPU's
// intersect_bvh_from, shaders/common/bvh.wgsl:137).
bool bvh_intersect(const BvhView& b, uint32_t root, const Ray& ray, float t_min, float t_max, TriHit* out,
                   TraversalStats* stats) {
    if (b.num_nodes == 0) return false;
    Vec3 inv_dir = safe_inv_dir(ray.dir);
    bool any = false;
    float closest = t_max;
    uint32_t stack[MAX_STACK];
    size_t sp = 0;
    uint32_t node = root;
    for (;;) {
        const GpuBvhNode& n = b.nodes[node];
        if (stats) stats->node_visits += 1;
        if (n.count > 0) {
            size_t start = n.left_first;
            for (size_t k = start; k < start + n.count; ++k) {
                uint32_t pi = b.prim_indices ? b.prim_indices[k] : (uint32_t)k;
                Vec3 p0, p1, p2;
                tri_pos(b, pi, &p0, &p1, &p2);
                if (stats) stats->triangle_tests += 1;
                float t, u, v;
                if (intersect_triangle(p0, p1, p2, ray, t_min, closest, &t, &u, &v)) {
                    closest = t;
                    *out = TriHit{t, u, v, pi};
                    any = true;
                }
            }
        } else {
            uint32_t c0 = n.left_first, c1 = c0 + 1;
            float a = 0, bb = 0;
            bool h0 = node_aabb(b.nodes[c0]).hit(ray.origin, inv_dir, t_min, closest, &a);
            bool h1 = node_aabb(b.nodes[c1]).hit(ray.origin, inv_dir, t_min, closest, &bb);
            if (h0 && h1) {
                uint32_t nearc = a <= bb ? c0 : c1;
                uint32_t farc = a <= bb ? c1 : c0;
                if (sp < MAX_STACK) stack[sp++] = farc;
                node = nearc;
                continue;
            } else if (h0) {
                node = c0;
                continue;
            } else if (h1) {
                node = c1;
                continue;
            }
        }
        if (sp == 0) break;
        node = stack[--sp];
    }
    return any;
}

// bvh.rs:1137-1190
bool bvh_occluded(const BvhView& b, uint32_t root, const Ray& ray, float t_min, float t_max) {
    if (b.num_nodes == 0) return false;
    Vec3 inv_dir = safe_inv_dir(ray.dir);
    uint32_t stack[MAX_STACK];
    size_t sp = 0;
    uint32_t node = root;
    for (;;) {
        const GpuBvhNode& n = b.nodes[node];
        if (n.count > 0) {
            size_t start = n.left_first;
            for (size_t k = start; k < start + n.count; ++k) {
                uint32_t pi = b.prim_indices ? b.prim_indices[k] : (uint32_t)k;
                Vec3 p0, p1, p2;
                tri_pos(b, pi, &p0, &p1, &p2);
                float t, u, v;
                if (intersect_triangle(p0, p1, p2, ray, t_min, t_max, &t, &u, &v)) return true;
            }
        } else {
            uint32_t c0 = n.left_first, c1 = c0 + 1;
            float e;
            bool h0 = node_aabb(b.nodes[c0]).hit(ray.origin, inv_dir, t_min, t_max, &e);
            bool h1 = node_aabb(b.nodes[c1]).hit(ray.origin, inv_dir, t_min, t_max, &e);
            if (h0) {
                if (h1 && sp < MAX_STACK) stack[sp++] = c1;
                node = c0;
                continue;
            }
            if (h1) {
                node = c1;
                continue;
            }
        }
        if (sp == 0) return false;
        node = stack[--sp];
    }
}

}  // namespace oracle
