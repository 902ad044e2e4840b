// Oracle: the linear BVH builder, restated from the reference's
// crates/core/src/lbvh.rs (expand_bits :72, morton3d :92, normalise :109,
// morton_codes :123, longest_common_prefix :149, LBVH_MAX_LEAF :185,
// build_hierarchy :211-292, relayout :315-406, Lbvh::build_with_leaf_size
// :422-547, Lbvh::from_parts :549-608).
//
// TEST INFRASTRUCTURE: CPU oracle only; never linked into the product.
#include "lbvh.hpp"

#include <algorithm>
#include <numeric>

namespace oracle {

uint32_t expand_bits(uint32_t v) {
    uint32_t x = v & 0x000003FFu;
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

uint32_t morton3d(Vec3 p) {
    const float SCALE = 1023.0f;
    auto q = [&](float v) { return f2u_sat(clampf(v * SCALE, 0.0f, SCALE)); };
    return (expand_bits(q(p.x)) << 2) | (expand_bits(q(p.y)) << 1) | expand_bits(q(p.z));
}

Vec3 normalise(Vec3 p, const Aabb& b) {
    Vec3 e = b.max - b.min;
    auto f = [](float v, float lo, float ext) { return ext > 0.0f ? (v - lo) / ext : 0.5f; };
    return Vec3(f(p.x, b.min.x, e.x), f(p.y, b.min.y, e.y), f(p.z, b.min.z, e.z));
}

int longest_common_prefix(const uint32_t* codes, size_t n, int i, int j) {
    if (j < 0 || j >= (int)n) return -1;
    uint32_t a = codes[i], b = codes[j];
    if (a == b) {
        uint32_t x = (uint32_t)i ^ (uint32_t)j;
        return 32 + (x == 0 ? 32 : __builtin_clz(x));
    }
    return __builtin_clz(a ^ b);
}

static int signum(int v) { return (v > 0) - (v < 0); }

// lbvh.rs:211-292
void build_hierarchy(const uint32_t* codes, size_t n, std::vector<uint32_t>& karras, std::vector<uint32_t>& parent) {
    karras.assign(2 * (n - 1), 0);
    parent.assign(2 * n - 1, 0xFFFFFFFFu);
    const uint32_t leaf_base = (uint32_t)(n - 1);
    auto delta = [&](int i, int j) { return longest_common_prefix(codes, n, i, j); };
    for (size_t ii = 0; ii + 1 < n; ++ii) {
        int i = (int)ii;
        int d = signum(delta(i, i + 1) - delta(i, i - 1));
        int delta_min = delta(i, i - d);
        int l_max = 2;
        while (delta(i, i + l_max * d) > delta_min) l_max *= 2;
        int l = 0;
        for (int t = l_max / 2; t >= 1; t /= 2)
            if (delta(i, i + (l + t) * d) > delta_min) l += t;
        int j = i + l * d;
        int first = std::min(i, j), last = std::max(i, j);
        int delta_node = delta(i, j);
        int s = 0;
        int t = l;
        while (t > 1) {
            t = (t + 1) / 2;
            if (delta(i, i + (s + t) * d) > delta_node) s += t;
        }
        int split = i + s * d + std::min(d, 0);
        uint32_t left = split == first ? leaf_base + (uint32_t)split : (uint32_t)split;
        uint32_t right = split + 1 == last ? leaf_base + (uint32_t)(split + 1) : (uint32_t)(split + 1);
        karras[2 * ii] = left;
        karras[2 * ii + 1] = right;
        parent[left] = (uint32_t)ii;
        parent[right] = (uint32_t)ii;
    }
}

// lbvh.rs:315-406
static void relayout(const std::vector<uint32_t>& karras, const std::vector<uint32_t>& subtree,
                     const std::vector<uint32_t>& sorted, const std::vector<Aabb>& bounds, size_t n, size_t max_leaf,
                     LbvhResult& out) {
    const uint32_t leaf_base = (uint32_t)(n - 1);
    out.nodes.clear();
    out.prim_indices.clear();
    out.stats = BuildStats{};
    out.stats.triangles = n;
    auto gather = [&](uint32_t root) {
        std::vector<uint32_t> stack{root};
        while (!stack.empty()) {
            uint32_t id = stack.back();
            stack.pop_back();
            if (id >= leaf_base) {
                out.prim_indices.push_back(sorted[id - leaf_base]);
            } else {
                stack.push_back(karras[2 * id + 1]);
                stack.push_back(karras[2 * id]);
            }
        }
    };
    auto size_of = [&](uint32_t id) -> uint32_t { return id >= leaf_base ? 1u : subtree[id]; };
    GpuBvhNode def{};
    def.bounds_min[0] = def.bounds_min[1] = def.bounds_min[2] = std::numeric_limits<float>::infinity();
    def.bounds_max[0] = def.bounds_max[1] = def.bounds_max[2] = -std::numeric_limits<float>::infinity();
    out.nodes.push_back(def);
    struct Item {
        uint32_t id;
        size_t slot, depth;
    };
    std::vector<Item> stack{{0u, 0, 1}};
    while (!stack.empty()) {
        Item it = stack.back();
        stack.pop_back();
        out.stats.max_depth = std::max(out.stats.max_depth, it.depth);
        const Aabb& b = bounds[it.id];
        size_t count = size_of(it.id);
        if (it.id >= leaf_base || count <= max_leaf) {
            uint32_t first = (uint32_t)out.prim_indices.size();
            gather(it.id);
            uint32_t got = (uint32_t)out.prim_indices.size() - first;
            GpuBvhNode nd{};
            b.min.store(nd.bounds_min);
            b.max.store(nd.bounds_max);
            nd.left_first = first;
            nd.count = got;
            out.nodes[it.slot] = nd;
            out.stats.leaves += 1;
            out.stats.max_leaf_size = std::max(out.stats.max_leaf_size, (size_t)got);
            continue;
        }
        size_t left_slot = out.nodes.size();
        out.nodes.push_back(def);
        out.nodes.push_back(def);
        GpuBvhNode nd{};
        b.min.store(nd.bounds_min);
        b.max.store(nd.bounds_max);
        nd.left_first = (uint32_t)left_slot;
        nd.count = 0;
        out.nodes[it.slot] = nd;
        stack.push_back({karras[2 * it.id + 1], left_slot + 1, it.depth + 1});
        stack.push_back({karras[2 * it.id], left_slot, it.depth + 1});
    }
    out.stats.nodes = out.nodes.size();
    out.stats.mean_leaf_size = out.stats.leaves > 0 ? (float)n / (float)out.stats.leaves : 0.0f;
}

// lbvh.rs:422-547 — stages kept separately so every one can be compared.
void lbvh_build(const GpuTriangle* tris, size_t n, const float* positions, size_t max_leaf, LbvhResult& out) {
    max_leaf = std::max<size_t>(max_leaf, 1);
    out = LbvhResult{};
    if (n == 0) return;
    std::vector<Aabb> tri_bounds(n);
    for (size_t i = 0; i < n; ++i) tri_bounds[i] = triangle_bounds(tris[i], positions);
    Aabb cb;
    for (const auto& b : tri_bounds) cb.grow_point(b.centroid());
    out.scene_bounds = cb;
    out.codes.resize(n);
    for (size_t i = 0; i < n; ++i) out.codes[i] = morton3d(normalise(tri_bounds[i].centroid(), cb));
    out.order.resize(n);
    std::iota(out.order.begin(), out.order.end(), 0u);
    // order.sort_by_key(|&i| (codes[i], i)): a total order, so any sort agrees.
    std::sort(out.order.begin(), out.order.end(), [&](uint32_t a, uint32_t b) {
        return out.codes[a] != out.codes[b] ? out.codes[a] < out.codes[b] : a < b;
    });
    std::vector<uint32_t> sorted_codes(n);
    for (size_t i = 0; i < n; ++i) sorted_codes[i] = out.codes[out.order[i]];

    if (n == 1) {
        GpuBvhNode nd{};
        tri_bounds[0].min.store(nd.bounds_min);
        tri_bounds[0].max.store(nd.bounds_max);
        nd.left_first = 0;
        nd.count = 1;
        out.nodes = {nd};
        out.prim_indices = {0};
        out.node_bounds = {tri_bounds[0]};
        out.parent = {0xFFFFFFFFu};
        out.stats = BuildStats{1, 1, 1, 1, 1, 1.0f, 1.0f};
        return;
    }

    build_hierarchy(sorted_codes.data(), n, out.karras, out.parent);
    const uint32_t leaf_base = (uint32_t)(n - 1);
    out.node_bounds.assign(2 * n - 1, Aabb{});
    out.subtree.assign(n - 1, 0);
    std::vector<uint32_t> visits(n - 1, 0);
    for (size_t leaf = 0; leaf < n; ++leaf) out.node_bounds[leaf_base + leaf] = tri_bounds[out.order[leaf]];
    for (size_t leaf = 0; leaf < n; ++leaf) {
        uint32_t node = out.parent[leaf_base + leaf];
        while (node != 0xFFFFFFFFu) {
            visits[node] += 1;
            if (visits[node] == 1) break;
            uint32_t l = out.karras[2 * node], r = out.karras[2 * node + 1];
            Aabb b = out.node_bounds[l];
            b.grow(out.node_bounds[r]);
            out.node_bounds[node] = b;
            auto sz = [&](uint32_t id) -> uint32_t { return id >= leaf_base ? 1u : out.subtree[id]; };
            out.subtree[node] = sz(l) + sz(r);
            node = out.parent[node];
        }
    }
    relayout(out.karras, out.subtree, out.order, out.node_bounds, n, max_leaf, out);
    out.stats.sah_cost = sah_cost(out.nodes);
}

}  // namespace oracle
