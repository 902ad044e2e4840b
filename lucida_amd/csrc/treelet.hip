// Treelet restructuring of the LBVH on the device. No counterpart in the
// reference, whose GPU path traces the LBVH as built (crates/gpu/src/lbvh.rs,
// relayout crates/core/src/lbvh.rs:315-406); the method is Karras and Aila,
// "Fast Parallel Construction of High-Quality Bounding Volume Hierarchies",
// HPG 2013.
//
// After the bit-exact LBVH (Morton sort, Karras hierarchy, fitted boxes), each
// round climbs the tree bottom-up, one lane per primitive: at every internal
// node the first arrival stops and the second continues, so a node is handled
// once both of its subtrees are final. A node of at least
// LUCIDA_TREELET_MIN_SIZE primitives is handled by its whole wave: it grows a
// treelet of up to 7 leaves under the node (always opening the leaf of
// largest surface area), finds the treelet topology of least cost by dynamic
// programming over the 127 leaf subsets (subset boxes and counts in LDS, one
// subset size after another), and rewrites the treelet's internal nodes in
// place. Smaller nodes only take their cost from their children. A node
// costs CI x area plus its children's costs, or CT x area x primitives as a
// leaf when it holds at most LUCIDA_TREELET_MAX_LEAF; the cheaper one wins,
// the leaf on a tie. The emit pass then lays the tree out as the LBVH's
// relayout does (children adjacent, depth first, left first) with the chosen
// subtrees collapsed into leaves.
//
// The result is deterministic (a node's outcome depends only on its own
// subtree) and equals oracle/treelet.cpp's bit for bit: the same tie-breaks
// and the same float operations, compiled with contraction off.
#include <hip/hip_runtime.h>

#include "context.hpp"

namespace lucida {
namespace {

constexpr uint32_t NONE = 0xFFFFFFFFu;
constexpr int TL_BLOCK = 256;
constexpr int TL_WAVES = TL_BLOCK / 64;
constexpr int LEAVES = LUCIDA_TREELET_LEAVES;
constexpr uint32_t SUBSETS = 1u << LEAVES;
constexpr uint32_t MIN_SIZE = LUCIDA_TREELET_MIN_SIZE;
constexpr uint32_t MAX_LEAF = LUCIDA_TREELET_MAX_LEAF;
constexpr float CI = LUCIDA_TREELET_CI;
constexpr float CT = LUCIDA_TREELET_CT;
constexpr uint32_t LEAF_BIT = 0x100u;  // in a subset's part word: the subset is kept as one leaf
static_assert(LEAVES == 7 && MIN_SIZE >= 3, "one wave covers the 127 subsets of 7 leaves, two per lane");

// One wave's working set for one treelet.
struct TreeletLds {
    float box[6][SUBSETS];  // per subset: lo xyz, hi xyz
    float cost[SUBSETS];
    uint32_t count[SUBSETS];
    uint32_t part[SUBSETS];  // left side of the best split | LEAF_BIT
    uint32_t leaf[8];        // the treelet's leaves
    uint32_t inner[8];       // its internal nodes, the root first
    uint32_t stack_set[8], stack_slot[8];
    uint32_t rec_slot[8], rec_left[8], rec_right[8], rec_leaf[8];
};

__device__ inline float area(float lx, float ly, float lz, float hx, float hy, float hz) {
    const float dx = hx - lx, dy = hy - ly, dz = hz - lz;
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}

__device__ inline float node_area(const LucidaAabb* box, uint32_t id) {
    const float4* p = reinterpret_cast<const float4*>(box + id);
    const float4 a = p[0], b = p[1];
    return area(a.x, a.y, a.z, b.x, b.y, b.z);
}

// Orders this wave's LDS traffic (the lanes of one wave only).
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct TreeArrays {
    uint32_t n, leaf_base;
    uint2* child;       // internal nodes: (left, right)
    uint32_t* parent;   // all nodes
    LucidaAabb* box;    // all nodes
    uint32_t* size;     // internal nodes: primitives below
    float* cost;        // all nodes
    uint32_t* leafy;    // internal nodes: kept as one leaf
    uint32_t* inner;    // internal nodes: internal nodes emitted below (itself included)
};

__device__ inline uint32_t size_of(const TreeArrays& t, uint32_t id) { return id >= t.leaf_base ? 1u : t.size[id]; }
__device__ inline uint32_t inner_of(const TreeArrays& t, uint32_t id) { return id >= t.leaf_base ? 0u : t.inner[id]; }

// The least-cost subtree over subset s of the treelet's leaves, from the
// costs of its proper subsets (already in LDS). Partitions keep the lowest
// leaf on the left and are tried in ascending order of that side's mask.
__device__ inline void solve_subset(TreeletLds& L, uint32_t s) {
    const uint32_t low = s & (0u - s), rest = s ^ low;
    float best = __builtin_huge_valf();
    uint32_t best_p = low;  // a valid split even if no cost compares below infinity
    for (uint32_t q = 0; q != rest; q = (q - rest) & rest) {
        const uint32_t pm = low | q;
        const float cand = L.cost[pm] + L.cost[s ^ pm];
        if (cand < best) {
            best = cand;
            best_p = pm;
        }
    }
    const float a = area(L.box[0][s], L.box[1][s], L.box[2][s], L.box[3][s], L.box[4][s], L.box[5][s]);
    const float inner = CI * a + best;
    const float leaf = CT * a * (float)L.count[s];
    const bool as_leaf = L.count[s] <= MAX_LEAF && leaf <= inner;
    L.cost[s] = as_leaf ? leaf : inner;
    L.part[s] = best_p | (as_leaf ? LEAF_BIT : 0u);
}

// The whole wave rebuilds the treelet under node x (uniform).
__device__ void restructure(const TreeArrays& t, TreeletLds& L, uint32_t lane, uint32_t x) {
    wave_sync();  // the previous treelet's LDS reads are done
    // Grow: lane i < k holds leaf i; the leaf of largest area (the first
    // such) is replaced by its left child and its right child appended.
    const uint2 c0 = t.child[x];
    uint32_t id = lane == 0 ? c0.x : (lane == 1 ? c0.y : NONE);
    float id_area = -1.0f;  // internal leaves only; a primitive is never opened
    if (lane < 2 && id < t.leaf_base) id_area = node_area(t.box, id);
    uint32_t inner_id = lane == 0 ? x : NONE;
    int k = 2;
    while (k < LEAVES) {
        int best = -1;
        float best_a = -1.0f;
        for (int i = 0; i < k; ++i) {
            const float a = __shfl(id_area, i, 64);
            if (a > best_a) {
                best_a = a;
                best = i;
            }
        }
        if (best < 0) break;
        const uint32_t e = __shfl(id, best, 64);
        const uint2 ce = t.child[e];
        if (lane == (uint32_t)(k - 1)) inner_id = e;
        if (lane == (uint32_t)best) {
            id = ce.x;
            id_area = ce.x < t.leaf_base ? node_area(t.box, ce.x) : -1.0f;
        }
        if (lane == (uint32_t)k) {
            id = ce.y;
            id_area = ce.y < t.leaf_base ? node_area(t.box, ce.y) : -1.0f;
        }
        ++k;
    }
    // The leaves' boxes, costs and sizes.
    float lo0 = 0, lo1 = 0, lo2 = 0, hi0 = 0, hi1 = 0, hi2 = 0, my_cost = 0;
    uint32_t my_size = 0;
    if (lane < (uint32_t)k) {
        const float4* p = reinterpret_cast<const float4*>(t.box + id);
        const float4 a = p[0], b = p[1];
        lo0 = a.x; lo1 = a.y; lo2 = a.z;
        hi0 = b.x; hi1 = b.y; hi2 = b.z;
        my_cost = t.cost[id];
        my_size = size_of(t, id);
        L.leaf[lane] = id;
    }
    if (lane < (uint32_t)(k - 1)) L.inner[lane] = inner_id;
    // Subset boxes and counts: lane l covers subsets l + 1 and l + 65; each
    // box is the union of its leaves in ascending leaf order.
    const uint32_t full = (1u << k) - 1;
    const uint32_t s0 = lane + 1, s1 = lane + 65;
    const float inf = __builtin_huge_valf();
    float a0[6] = {inf, inf, inf, -inf, -inf, -inf}, a1[6] = {inf, inf, inf, -inf, -inf, -inf};
    uint32_t n0 = 0, n1 = 0;
    for (int i = 0; i < k; ++i) {
        const float q[6] = {__shfl(lo0, i, 64), __shfl(lo1, i, 64), __shfl(lo2, i, 64),
                            __shfl(hi0, i, 64), __shfl(hi1, i, 64), __shfl(hi2, i, 64)};
        const uint32_t qs = __shfl(my_size, i, 64);
        if (s0 >> i & 1u) {
            for (int j = 0; j < 3; ++j) {
                a0[j] = fminf(a0[j], q[j]);
                a0[j + 3] = fmaxf(a0[j + 3], q[j + 3]);
            }
            n0 += qs;
        }
        if (s1 >> i & 1u) {
            for (int j = 0; j < 3; ++j) {
                a1[j] = fminf(a1[j], q[j]);
                a1[j + 3] = fmaxf(a1[j + 3], q[j + 3]);
            }
            n1 += qs;
        }
    }
    if (s0 <= full) {
        for (int j = 0; j < 6; ++j) L.box[j][s0] = a0[j];
        L.count[s0] = n0;
    }
    if (s1 <= full) {
        for (int j = 0; j < 6; ++j) L.box[j][s1] = a1[j];
        L.count[s1] = n1;
    }
    if (lane < (uint32_t)k) {
        L.cost[1u << lane] = my_cost;
        L.part[1u << lane] = 0;
    }
    wave_sync();
    // Least costs, subsets of 2 leaves first, then 3, ...
    const int m0 = __popc(s0), m1 = __popc(s1);
    for (int m = 2; m <= k; ++m) {
        if (s0 <= full && m0 == m) solve_subset(L, s0);
        if (s1 <= full && m1 == m) solve_subset(L, s1);
        wave_sync();
    }
    // Rebuild in pre-order, left side first, reusing the treelet's internal
    // nodes in the order they were taken; then the emitted-internal counts in
    // reverse pre-order (children before parents).
    if (lane == 0) {
        uint32_t sp = 1, next = 1, nrec = 0;
        L.stack_set[0] = full;
        L.stack_slot[0] = x;
        while (sp) {
            --sp;
            const uint32_t s = L.stack_set[sp], slot = L.stack_slot[sp];
            const uint32_t pw = L.part[s];
            const uint32_t pm = pw & (SUBSETS - 1), qm = s ^ pm;
            const uint32_t lc = __popc(pm) == 1 ? L.leaf[__ffs(pm) - 1] : L.inner[next++];
            const uint32_t rc = __popc(qm) == 1 ? L.leaf[__ffs(qm) - 1] : L.inner[next++];
            t.child[slot] = make_uint2(lc, rc);
            t.parent[lc] = slot;
            t.parent[rc] = slot;
            float4* bp = reinterpret_cast<float4*>(t.box + slot);
            bp[0] = make_float4(L.box[0][s], L.box[1][s], L.box[2][s], 0.0f);
            bp[1] = make_float4(L.box[3][s], L.box[4][s], L.box[5][s], 0.0f);
            t.size[slot] = L.count[s];
            t.cost[slot] = L.cost[s];
            const uint32_t as_leaf = (pw & LEAF_BIT) ? 1u : 0u;
            t.leafy[slot] = as_leaf;
            L.rec_slot[nrec] = slot;
            L.rec_left[nrec] = lc;
            L.rec_right[nrec] = rc;
            L.rec_leaf[nrec] = as_leaf;
            ++nrec;
            if (__popc(qm) > 1) {
                L.stack_set[sp] = qm;
                L.stack_slot[sp] = rc;
                ++sp;
            }
            if (__popc(pm) > 1) {
                L.stack_set[sp] = pm;
                L.stack_slot[sp] = lc;
                ++sp;
            }
        }
        for (int r = (int)nrec - 1; r >= 0; --r)
            t.inner[L.rec_slot[r]] =
                L.rec_leaf[r] ? 0u : 1u + inner_of(t, L.rec_left[r]) + inner_of(t, L.rec_right[r]);
    }
}

// Longest climb a valid tree allows (its depth is far below this); a longer
// one means broken parent links, and the pass stops and reports it.
constexpr uint32_t MAX_CLIMB = 1u << 16;

// One restructuring round. Each lane starts at one primitive; the second
// arrival at a node continues (agent-scope release before the arrival count,
// acquire after it, as the subtrees were written by other waves, possibly on
// other XCDs). Every lane's climb ends at its first arrival or past the root,
// so the loop runs at most the tree's depth times.
__global__ __launch_bounds__(TL_BLOCK) void k_treelet_round(TreeArrays t, uint32_t* __restrict__ arrivals,
                                                            uint32_t* __restrict__ fault) {
    __shared__ TreeletLds lds[TL_WAVES];
    TreeletLds& L = lds[threadIdx.x / 64];
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t k = blockIdx.x * TL_BLOCK + threadIdx.x;
    bool active = k < t.n;
    uint32_t cur = NONE;
    if (active) {
        const uint32_t id = t.leaf_base + k;
        t.cost[id] = CT * node_area(t.box, id);
        cur = t.parent[id];
        active = cur != NONE;
    }
    for (uint32_t step = 0;; ++step) {
        if (step == MAX_CLIMB) {  // uniform: every lane of the wave counts the same steps
            if (lane == 0) atomicOr(fault, 1u);
            break;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        if (active) {
            const uint32_t old = __hip_atomic_fetch_add(&arrivals[cur], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            active = old != 0;  // the first arrival stops; the sibling's lane finishes this node
        }
        if (__ballot(active) == 0) break;
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        const bool big = active && t.size[cur] >= MIN_SIZE;
        uint64_t todo = __ballot(big);
        while (todo) {
            const int l = __ffsll((unsigned long long)todo) - 1;
            todo &= todo - 1;
            restructure(t, L, lane, __shfl(cur, l, 64));
        }
        if (active && !big) {
            const uint2 c = t.child[cur];
            const float a = node_area(t.box, cur);
            const float split = t.cost[c.x] + t.cost[c.y];
            const float inner = CI * a + split;
            const uint32_t sz = t.size[cur];
            const float leaf = CT * a * (float)sz;
            const bool as_leaf = sz <= MAX_LEAF && leaf <= inner;
            t.cost[cur] = as_leaf ? leaf : inner;
            t.leafy[cur] = as_leaf ? 1u : 0u;
            t.inner[cur] = as_leaf ? 0u : 1u + inner_of(t, c.x) + inner_of(t, c.y);
        }
        if (active) {
            cur = t.parent[cur];
            active = cur != NONE;
        }
    }
}

__global__ __launch_bounds__(256) void k_clear_words(uint32_t* __restrict__ p, uint32_t count) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < count) p[i] = 0u;
}

// Layout: an emitted internal node of pre-order rank R among emitted internal
// nodes has its children at 1 + 2R and 2 + 2R. Each node finds its rank, its
// parent's rank and its first primitive by walking to the root: R counts each
// ancestor plus the emitted internal nodes of every left sibling subtree on
// the way; the first primitive counts the primitives of those left siblings.
// A node under a collapsed ancestor is not emitted, but a primitive still
// takes its place in the leaf order.
__global__ __launch_bounds__(256) void k_treelet_emit(TreeArrays t, const uint32_t* __restrict__ sorted,
                                                      GpuBvhNode* __restrict__ out, uint32_t* __restrict__ order,
                                                      uint32_t* __restrict__ fault) {
    const uint32_t y = blockIdx.x * 256 + threadIdx.x;
    if (y >= 2 * t.n - 1) return;
    bool emitted = true, right = false;
    uint32_t rank = 0, first = 0, sibling_inner = 0, cur = y, steps = 0;
    while (cur != 0) {
        if (++steps == MAX_CLIMB) {
            atomicOr(fault, 2u);
            return;
        }
        const uint32_t p = t.parent[cur];
        const uint2 c = t.child[p];
        if (t.leafy[p]) emitted = false;
        const bool r = c.y == cur;
        const uint32_t li = r ? inner_of(t, c.x) : 0u;
        if (cur == y) {
            right = r;
            sibling_inner = li;
        }
        rank += 1 + li;
        if (r) first += size_of(t, c.x);
        cur = p;
    }
    if (y >= t.leaf_base) {
        if (first >= t.n) {
            atomicOr(fault, 4u);
            return;
        }
        order[first] = sorted[y - t.leaf_base];
    }
    if (!emitted) return;
    const uint32_t slot = y == 0 ? 0u : 1u + 2u * (rank - 1u - sibling_inner) + (right ? 1u : 0u);
    if (slot >= 2 * t.n - 1) {
        atomicOr(fault, 4u);
        return;
    }
    const float4* bp = reinterpret_cast<const float4*>(t.box + y);
    const float4 a = bp[0], b = bp[1];
    GpuBvhNode nd;
    nd.bounds_min[0] = a.x; nd.bounds_min[1] = a.y; nd.bounds_min[2] = a.z;
    nd.bounds_max[0] = b.x; nd.bounds_max[1] = b.y; nd.bounds_max[2] = b.z;
    if (y >= t.leaf_base || t.leafy[y]) {
        nd.left_first = first;
        nd.count = size_of(t, y);
    } else {
        nd.left_first = 1u + 2u * rank;
        nd.count = 0;
    }
    out[slot] = nd;
}

__global__ __launch_bounds__(256) void k_treelet_gather(const GpuTriangle* __restrict__ tris,
                                                        const uint32_t* __restrict__ order, uint32_t n,
                                                        GpuTriangle* __restrict__ out) {
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k < n) out[k] = tris[order[k]];
}

inline uint32_t blocks_for(size_t n, int block) { return (uint32_t)((n + block - 1) / block); }

}  // namespace

lucida_status bvh_build_treelet_device(lucida_ctx* ctx, const GpuTriangle* tris, size_t n, const float* positions,
                                       uint32_t rounds, GpuBvhNode* nodes, size_t* num_nodes,
                                       uint32_t* prim_indices, GpuTriangle* tris_sorted, float* stage_us) {
    if (n < 2)  // nothing to restructure: the LBVH's single leaf
        return lbvh_build_device(ctx, tris, n, positions, nodes, num_nodes, prim_indices, tris_sorted, nullptr);
    if (n > (1u << 30)) return set_error(LUCIDA_INVALID_ARG, "treelet BVH: more than 2^30 triangles");
    hipStream_t s = ctx->stream;
    hipEvent_t ev[4] = {};
    if (stage_us) {
        for (auto& e : ev) LUCIDA_HIP(hipEventCreate(&e), LUCIDA_VALIDATION);
        LUCIDA_HIP(hipEventRecord(ev[0], s), LUCIDA_VALIDATION);
    }
    lucida_status st = lbvh_fit_device(ctx, tris, n, positions);
    if (st != LUCIDA_OK) return st;
    if (stage_us) LUCIDA_HIP(hipEventRecord(ev[1], s), LUCIDA_VALIDATION);
    const size_t internal = n - 1;
    LUCIDA_HIP(ctx->tl_cost.ensure((2 * n - 1) * 4), LUCIDA_OOM);
    LUCIDA_HIP(ctx->tl_leafy.ensure((internal + 1) * 4), LUCIDA_OOM);  // + the fault word
    LUCIDA_HIP(ctx->tl_order.ensure(n * 4), LUCIDA_OOM);
    TreeArrays t;
    t.n = (uint32_t)n;
    t.leaf_base = (uint32_t)internal;
    t.child = ctx->lb_karras.as<uint2>();
    t.parent = ctx->lb_parent.as<uint32_t>();
    t.box = ctx->lb_nodebox.as<LucidaAabb>();
    t.size = ctx->lb_subtree.as<uint32_t>();
    t.inner = t.size + internal;
    t.cost = ctx->tl_cost.as<float>();
    t.leafy = ctx->tl_leafy.as<uint32_t>();
    uint32_t* arrivals = ctx->lb_flags.as<uint32_t>();
    uint32_t* fault = t.leafy + internal;
    hipLaunchKernelGGL(k_clear_words, dim3(1), dim3(256), 0, s, fault, 1u);
    for (uint32_t r = 0; r < (rounds ? rounds : 1u); ++r) {
        hipLaunchKernelGGL(k_clear_words, dim3(blocks_for(internal, 256)), dim3(256), 0, s, arrivals,
                           (uint32_t)internal);
        hipLaunchKernelGGL(k_treelet_round, dim3(blocks_for(n, TL_BLOCK)), dim3(TL_BLOCK), 0, s, t, arrivals, fault);
    }
    if (stage_us) LUCIDA_HIP(hipEventRecord(ev[2], s), LUCIDA_VALIDATION);
    uint32_t* order = prim_indices ? prim_indices : ctx->tl_order.as<uint32_t>();
    hipLaunchKernelGGL(k_treelet_emit, dim3(blocks_for(2 * n - 1, 256)), dim3(256), 0, s, t,
                       ctx->lb_vals.as<uint32_t>(), nodes, order, fault);
    if (tris_sorted)
        hipLaunchKernelGGL(k_treelet_gather, dim3(blocks_for(n, 256)), dim3(256), 0, s, tris, order, (uint32_t)n,
                           tris_sorted);
    LUCIDA_HIP(hipGetLastError(), LUCIDA_SHADER);
    if (stage_us) LUCIDA_HIP(hipEventRecord(ev[3], s), LUCIDA_VALIDATION);
    // the root's emitted-internal count (0 when the root is one leaf)
    uint32_t root_inner = 0, faults = 0;
    LUCIDA_HIP(hipMemcpyAsync(&root_inner, t.inner, 4, hipMemcpyDeviceToHost, s), LUCIDA_READBACK);
    LUCIDA_HIP(hipMemcpyAsync(&faults, fault, 4, hipMemcpyDeviceToHost, s), LUCIDA_READBACK);
    LUCIDA_HIP(hipStreamSynchronize(s), LUCIDA_VALIDATION);
    if (faults || root_inner >= n)
        return set_error(LUCIDA_VALIDATION, "treelet BVH: inconsistent tree (parent links or layout)");
    *num_nodes = 1 + 2 * (size_t)root_inner;
    if (stage_us) {
        for (int k = 0; k < 3; ++k) {
            float ms = 0;
            LUCIDA_HIP(hipEventElapsedTime(&ms, ev[k], ev[k + 1]), LUCIDA_VALIDATION);
            stage_us[k] = ms * 1000.0f;
        }
        for (auto& e : ev) (void)hipEventDestroy(e);
    }
    return LUCIDA_OK;
}

}  // namespace lucida

using namespace lucida;

extern "C" {

lucida_status lucida_bvh_build_treelet(lucida_ctx* ctx, const GpuTriangle* tris, size_t n, const float* positions,
                                       size_t num_positions, uint32_t rounds, GpuBvhNode* nodes_out,
                                       size_t* num_nodes, uint32_t* prim_indices_out) {
    if (!ctx || !num_nodes || (n && (!tris || !positions || !nodes_out || !prim_indices_out)))
        return set_error(LUCIDA_INVALID_ARG, "lucida_bvh_build_treelet: null argument");
    clear_error();
    if (n == 0) {
        *num_nodes = 0;
        return LUCIDA_OK;
    }
    for (size_t i = 0; i < n; ++i)
        if (tris[i].i0 >= num_positions || tris[i].i1 >= num_positions || tris[i].i2 >= num_positions)
            return set_error(LUCIDA_INVALID_ARG, "lucida_bvh_build_treelet: triangle index out of range");
    hipStream_t s = ctx->stream;
    LUCIDA_HIP(ctx->lb_tris.ensure(n * sizeof(GpuTriangle)), LUCIDA_OOM);
    LUCIDA_HIP(ctx->lb_pos.ensure(num_positions * 16), LUCIDA_OOM);
    LUCIDA_HIP(ctx->lb_nodes.ensure((2 * n - 1) * sizeof(GpuBvhNode)), LUCIDA_OOM);
    LUCIDA_HIP(ctx->lb_prims.ensure(n * 4), LUCIDA_OOM);
    LUCIDA_HIP(hipMemcpyAsync(ctx->lb_tris.ptr, tris, n * sizeof(GpuTriangle), hipMemcpyHostToDevice, s),
               LUCIDA_VALIDATION);
    LUCIDA_HIP(hipMemcpyAsync(ctx->lb_pos.ptr, positions, num_positions * 16, hipMemcpyHostToDevice, s),
               LUCIDA_VALIDATION);
    lucida_status st = bvh_build_treelet_device(ctx, ctx->lb_tris.as<GpuTriangle>(), n, ctx->lb_pos.as<float>(),
                                                rounds, ctx->lb_nodes.as<GpuBvhNode>(), num_nodes,
                                                ctx->lb_prims.as<uint32_t>(), nullptr, nullptr);
    if (st != LUCIDA_OK) return st;
    LUCIDA_HIP(hipMemcpyAsync(nodes_out, ctx->lb_nodes.ptr, *num_nodes * sizeof(GpuBvhNode), hipMemcpyDeviceToHost, s),
               LUCIDA_READBACK);
    LUCIDA_HIP(hipMemcpyAsync(prim_indices_out, ctx->lb_prims.ptr, n * 4, hipMemcpyDeviceToHost, s), LUCIDA_READBACK);
    LUCIDA_HIP(hipStreamSynchronize(s), LUCIDA_VALIDATION);
    return LUCIDA_OK;
}

}  // extern "C"
