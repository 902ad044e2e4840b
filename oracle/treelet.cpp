// Oracle: treelet restructuring of the LBVH. TEST INFRASTRUCTURE: CPU oracle
// only; never linked into the product.
//
// No counterpart in the reference: its GPU path traces the LBVH as built
// (crates/core/src/lbvh.rs:422, relayout :315-406) and keeps its binned-SAH
// builder on the CPU (bvh.rs:267). This is the CPU statement of the device
// pass in lucida_amd/csrc/treelet.hip, which rebuilds the LBVH's topology for
// a lower surface-area cost after the bit-exact build: Karras and Aila,
// "Fast Parallel Construction of High-Quality Bounding Volume Hierarchies",
// HPG 2013 (treelets of up to 7 leaves grown by largest surface area, an
// optimal topology per treelet by dynamic programming over leaf subsets, and
// subtrees collapsed into leaves where the cost model prefers it).
//
// Parity with the device pass is exact: both follow one schedule (children
// before parents, so each node sees its subtree's final state), one set of
// tie-breaks (largest area then most primitives, then the first, when
// growing a treelet; least cost, then the most even split, then the smallest
// mask, when splitting; a leaf on equal cost), and the same float operations
// with contraction off. The optimality
// of the subset search is checked against brute force over every binary tree
// of up to 6 leaves, and the trees it emits against validate() and against
// closest hits over the unrestructured LBVH.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#include "lbvh.hpp"
#include "oracle.h"

namespace oracle {
namespace {

struct Box {
    float lo[3], hi[3];
};

inline float area(const Box& b) {
    const float dx = b.hi[0] - b.lo[0], dy = b.hi[1] - b.lo[1], dz = b.hi[2] - b.lo[2];
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}

inline void grow(Box& a, const Box& b) {
    for (int k = 0; k < 3; ++k) {
        a.lo[k] = std::fmin(a.lo[k], b.lo[k]);
        a.hi[k] = std::fmax(a.hi[k], b.hi[k]);
    }
}

inline int popcount(uint32_t v) { return __builtin_popcount(v); }

// A cost with its 4 lowest mantissa bits cleared (costs are not negative, so
// the keys order as the costs do).
inline float cost_key(float c) {
    uint32_t u;
    std::memcpy(&u, &c, 4);
    u &= 0xFFFFFFF0u;
    float r;
    std::memcpy(&r, &u, 4);
    return r;
}
inline int lowest(uint32_t v) { return __builtin_ctz(v); }

struct Tree {
    uint32_t n = 0, leaf_base = 0;
    std::vector<uint32_t> left, right;   // internal nodes [0, n-1)
    std::vector<uint32_t> parent, size;  // all nodes [0, 2n-1)
    std::vector<Box> box;
    std::vector<float> cost;
    std::vector<uint8_t> leafy;  // the cost model keeps this subtree as one leaf
};

struct Params {
    float ci, ct;
    uint32_t max_leaf, min_size;
};

// One node's cost from its children's, the leaf option when it is allowed.
void combine(Tree& t, const Params& p, uint32_t x) {
    const float a = area(t.box[x]);
    const float split = t.cost[t.left[x]] + t.cost[t.right[x]];
    const float inner = p.ci * a + split;
    const float leaf = p.ct * a * (float)t.size[x];
    const bool as_leaf = t.size[x] <= p.max_leaf && leaf <= inner;
    t.cost[x] = as_leaf ? leaf : inner;
    t.leafy[x] = as_leaf;
}

}  // namespace

// Optimal topology over k <= 7 treelet leaves: for every subset s of the
// leaves, c[s] is the least cost of a subtree over exactly those leaves and
// part[s] the left side of its best split. A partition's left side holds the
// subset's lowest leaf; the split kept is the least (cost, imbalance, mask),
// imbalance being the difference of the two sides' primitive counts and
// costs compared to 20 mantissa bits (cost_key), so costs equal up to their
// rounding (coinciding boxes, summed in different orders) take the most even
// split instead of building chains. A NaN cost is never kept; with every
// cost NaN the split is the lowest leaf alone.
void treelet_dp(int k, const Box* leaf_box, const float* leaf_cost, const uint32_t* leaf_size, float ci, float ct,
                uint32_t max_leaf, Box* box, float* c, uint8_t* part, uint8_t* leafy) {
    const uint32_t full = (1u << k) - 1;
    std::vector<uint32_t> count(full + 1, 0);
    std::vector<float> a(full + 1, 0.0f);
    for (uint32_t s = 1; s <= full; ++s) {
        Box b = leaf_box[lowest(s)];
        uint32_t cnt = 0;
        for (int i = 0; i < k; ++i)
            if (s >> i & 1u) {
                grow(b, leaf_box[i]);
                cnt += leaf_size[i];
            }
        box[s] = b;
        a[s] = area(b);
        count[s] = cnt;
    }
    for (int i = 0; i < k; ++i) {
        c[1u << i] = leaf_cost[i];
        part[1u << i] = 0;
        leafy[1u << i] = 0;
    }
    for (uint32_t s = 1; s <= full; ++s) {
        if (popcount(s) < 2) continue;
        const uint32_t low = s & (0u - s), rest = s ^ low;
        float best_key = std::numeric_limits<float>::infinity();
        uint32_t best_p = 0, best_im = 0xFFFFFFFFu;
        for (uint32_t q = 0; q != rest; q = (q - rest) & rest) {
            const uint32_t pm = low | q;
            const float cand = c[pm] + c[s ^ pm];
            if (!(cand == cand)) continue;  // NaN
            const float qc = cost_key(cand);
            const uint32_t np = count[pm], nq = count[s ^ pm];
            const uint32_t im = np > nq ? np - nq : nq - np;
            if (qc < best_key || (qc == best_key && im < best_im)) {
                best_key = qc;
                best_im = im;
                best_p = pm;
            }
        }
        float best = std::numeric_limits<float>::infinity();
        if (best_p == 0)
            best_p = low;  // every cost NaN
        else
            best = c[best_p] + c[s ^ best_p];
        const float inner = ci * a[s] + best;
        const float leaf = ct * a[s] * (float)count[s];
        const bool as_leaf = count[s] <= max_leaf && leaf <= inner;
        c[s] = as_leaf ? leaf : inner;
        part[s] = (uint8_t)best_p;
        leafy[s] = as_leaf;
    }
}

namespace {

void restructure(Tree& t, const Params& p, uint32_t x) {
    uint32_t T[7], I[6];
    int k = 2, ni = 1;
    T[0] = t.left[x];
    T[1] = t.right[x];
    I[0] = x;
    while (k < 7) {
        // the leaf of largest area, then most primitives (the first such)
        int best = -1;
        float best_a = -1.0f;
        uint32_t best_n = 0;
        for (int i = 0; i < k; ++i) {
            if (T[i] >= t.leaf_base) continue;
            const float a = area(t.box[T[i]]);
            const uint32_t sz = t.size[T[i]];
            if (a >= 0.0f && (a > best_a || (a == best_a && sz > best_n))) {
                best_a = a;
                best_n = sz;
                best = i;
            }
        }
        if (best < 0) break;
        const uint32_t e = T[best];
        I[ni++] = e;
        T[best] = t.left[e];
        T[k++] = t.right[e];
    }
    Box lb[7];
    float lc[7];
    uint32_t ls[7];
    for (int i = 0; i < k; ++i) {
        lb[i] = t.box[T[i]];
        lc[i] = t.cost[T[i]];
        ls[i] = t.size[T[i]];
    }
    Box box[128];
    float c[128];
    uint8_t part[128], leafy[128];
    treelet_dp(k, lb, lc, ls, p.ci, p.ct, p.max_leaf, box, c, part, leafy);
    std::vector<uint32_t> count(128, 0);
    for (uint32_t s = 1; s < (1u << k); ++s)
        for (int i = 0; i < k; ++i)
            if (s >> i & 1u) count[s] += ls[i];
    // Rebuild in pre-order, left side first, reusing the treelet's internal
    // nodes in the order they were taken.
    struct Item {
        uint32_t s, slot;
    };
    Item stack[8];
    int sp = 0, next = 1;
    stack[sp++] = {(1u << k) - 1, x};
    while (sp) {
        const Item it = stack[--sp];
        const uint32_t pm = part[it.s], qm = it.s ^ pm;
        const uint32_t lc_id = popcount(pm) == 1 ? T[lowest(pm)] : I[next++];
        const uint32_t rc_id = popcount(qm) == 1 ? T[lowest(qm)] : I[next++];
        t.left[it.slot] = lc_id;
        t.right[it.slot] = rc_id;
        t.parent[lc_id] = it.slot;
        t.parent[rc_id] = it.slot;
        t.box[it.slot] = box[it.s];
        t.size[it.slot] = count[it.s];
        t.cost[it.slot] = c[it.s];
        t.leafy[it.slot] = leafy[it.s];
        if (popcount(qm) > 1) stack[sp++] = {qm, rc_id};
        if (popcount(pm) > 1) stack[sp++] = {pm, lc_id};
    }
}

// Children before parents; a node's children are read when it is first
// reached, and only its own treelet changes them afterwards.
void round(Tree& t, const Params& p) {
    std::vector<std::pair<uint32_t, bool>> stack;
    stack.push_back({0u, false});
    while (!stack.empty()) {
        auto [x, expanded] = stack.back();
        stack.pop_back();
        if (x >= t.leaf_base) continue;
        if (!expanded) {
            stack.push_back({x, true});
            stack.push_back({t.right[x], false});
            stack.push_back({t.left[x], false});
            continue;
        }
        if (t.size[x] >= p.min_size)
            restructure(t, p, x);
        else
            combine(t, p, x);
    }
}

}  // namespace

size_t treelet_build(const GpuTriangle* tris, size_t n, const float* positions, uint32_t rounds, float ci, float ct,
                     uint32_t max_leaf, uint32_t min_size, GpuBvhNode* nodes_out, uint32_t* prims_out,
                     float* sah_out) {
    if (n == 0) return 0;
    LbvhResult r;
    lbvh_build(tris, n, positions, max_leaf, r);
    if (n == 1) {
        if (nodes_out) nodes_out[0] = r.nodes[0];
        if (prims_out) prims_out[0] = 0;
        if (sah_out) *sah_out = r.stats.sah_cost;
        return 1;
    }
    Tree t;
    t.n = (uint32_t)n;
    t.leaf_base = (uint32_t)(n - 1);
    t.left.resize(n - 1);
    t.right.resize(n - 1);
    for (size_t i = 0; i + 1 < n; ++i) {
        t.left[i] = r.karras[2 * i];
        t.right[i] = r.karras[2 * i + 1];
    }
    t.parent = r.parent;
    t.size.assign(2 * n - 1, 1u);
    for (size_t i = 0; i + 1 < n; ++i) t.size[i] = r.subtree[i];
    t.box.resize(2 * n - 1);
    for (size_t i = 0; i < 2 * n - 1; ++i) {
        r.node_bounds[i].min.store(t.box[i].lo);
        r.node_bounds[i].max.store(t.box[i].hi);
    }
    t.cost.assign(2 * n - 1, 0.0f);
    t.leafy.assign(2 * n - 1, 0);
    const Params p{ci, ct, max_leaf, std::max<uint32_t>(min_size, 3)};
    for (size_t i = t.leaf_base; i < 2 * n - 1; ++i) t.cost[i] = p.ct * area(t.box[i]);
    for (uint32_t k = 0; k < std::max<uint32_t>(rounds, 1); ++k) round(t, p);

    // Emit depth first, left first: an emitted internal node with pre-order
    // rank R among emitted internal nodes has its children at 1 + 2R; a leaf
    // lists its subtree's primitives in the same left-first order.
    std::vector<GpuBvhNode> out(1);
    std::vector<uint32_t> prims;
    prims.reserve(n);
    uint32_t rank = 0;
    struct Item {
        uint32_t id, slot;
    };
    std::vector<Item> stack{{0u, 0u}};
    std::vector<uint32_t> sub;
    while (!stack.empty()) {
        const Item it = stack.back();
        stack.pop_back();
        GpuBvhNode nd{};
        std::memcpy(nd.bounds_min, t.box[it.id].lo, 12);
        std::memcpy(nd.bounds_max, t.box[it.id].hi, 12);
        if (it.id >= t.leaf_base || t.leafy[it.id]) {
            nd.left_first = (uint32_t)prims.size();
            nd.count = t.size[it.id];
            sub.assign(1, it.id);
            while (!sub.empty()) {
                const uint32_t y = sub.back();
                sub.pop_back();
                if (y >= t.leaf_base) {
                    prims.push_back(r.order[y - t.leaf_base]);
                } else {
                    sub.push_back(t.right[y]);
                    sub.push_back(t.left[y]);
                }
            }
        } else {
            const uint32_t first = 1 + 2 * rank++;
            nd.left_first = first;
            nd.count = 0;
            if (out.size() < first + 2) out.resize(first + 2);
            stack.push_back({t.right[it.id], first + 1});
            stack.push_back({t.left[it.id], first});
        }
        out[it.slot] = nd;
    }
    out.resize(1 + 2 * (size_t)rank);
    if (nodes_out) std::memcpy(nodes_out, out.data(), out.size() * sizeof(GpuBvhNode));
    if (prims_out) std::memcpy(prims_out, prims.data(), n * 4);
    if (sah_out) *sah_out = sah_cost(out);
    return out.size();
}

}  // namespace oracle

using namespace oracle;

extern "C" {

size_t oracle_treelet_build(const GpuTriangle* tris, size_t n, const float* positions, uint32_t rounds,
                            GpuBvhNode* nodes, uint32_t* prim_indices, float* sah) {
    return treelet_build(tris, n, positions, rounds, LUCIDA_TREELET_CI, LUCIDA_TREELET_CT, LUCIDA_TREELET_MAX_LEAF,
                         LUCIDA_TREELET_MIN_SIZE, nodes, prim_indices, sah);
}

size_t oracle_treelet_build_with(const GpuTriangle* tris, size_t n, const float* positions, uint32_t rounds,
                                 float ci, float ct, uint32_t max_leaf, uint32_t min_size, GpuBvhNode* nodes,
                                 uint32_t* prim_indices, float* sah) {
    return treelet_build(tris, n, positions, rounds, ci, ct, max_leaf, min_size, nodes, prim_indices, sah);
}

void oracle_treelet_dp(int k, const LucidaAabb* leaf_box, const float* leaf_cost, const uint32_t* leaf_size,
                       float* cost_out, uint8_t* part_out, uint8_t* leafy_out) {
    Box lb[7];
    for (int i = 0; i < k; ++i)
        for (int a = 0; a < 3; ++a) {
            lb[i].lo[a] = leaf_box[i].lo[a];
            lb[i].hi[a] = leaf_box[i].hi[a];
        }
    Box box[128];
    float c[128];
    uint8_t part[128], leafy[128];
    treelet_dp(k, lb, leaf_cost, leaf_size, LUCIDA_TREELET_CI, LUCIDA_TREELET_CT, LUCIDA_TREELET_MAX_LEAF, box, c,
               part, leafy);
    const uint32_t m = 1u << k;
    std::memcpy(cost_out, c, m * 4);
    std::memcpy(part_out, part, m);
    std::memcpy(leafy_out, leafy, m);
    cost_out[0] = 0.0f;
    part_out[0] = 0;
    leafy_out[0] = 0;
}

}  // extern "C"
