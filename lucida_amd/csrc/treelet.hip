// Treelet restructuring of the LBVH on the device. No counterpart in the
// reference, whose GPU path traces the LBVH as built (crates/gpu/src/lbvh.rs,
// relayout crates/core/src/lbvh.rs:315-406); the method is Karras and Aila,
// "Fast Parallel Construction of High-Quality Bounding Volume Hierarchies",
// HPG 2013.
//
// After the bit-exact LBVH (Morton sort, Karras hierarchy, fitted boxes), each
// round handles every internal node once its subtrees are final: nodes are
// bucketed by depth and one launch per level, deepest first, handles that
// level. A node of at least LUCIDA_TREELET_MIN_SIZE primitives gets a whole
// wave, which grows a treelet of up to 7 leaves under the node (always
// opening the leaf of largest surface area; the three levels under the
// node are fetched across lanes first), finds the treelet topology of least
// cost by dynamic programming over the 127 leaf subsets (subset boxes and
// counts in LDS, one subset size after another, each subset's partitions
// spread over a group of lanes), and rewrites the treelet's internal nodes in
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
    uint32_t icnt[SUBSETS];  // internal nodes the layout emits for the subset's best subtree
    uint32_t leaf[8];        // the treelet's leaves
    uint32_t inner[8];       // its internal nodes, the root first
    uint32_t stack_set[8], stack_slot[8];
};

// The 127 nonempty subsets of 7 leaves ordered by size, then mask; off[m] is
// where the subsets of m leaves start.
struct SubsetTable {
    uint8_t mask[SUBSETS - 1];
    uint8_t off[LEAVES + 2];
};
constexpr int popcount7(int v) { return (v & 1) + (v >> 1 & 1) + (v >> 2 & 1) + (v >> 3 & 1) + (v >> 4 & 1) + (v >> 5 & 1) + (v >> 6 & 1); }
constexpr SubsetTable make_subset_table() {
    SubsetTable t{};
    int n = 0;
    for (int m = 1; m <= LEAVES; ++m) {
        t.off[m] = (uint8_t)n;
        for (int v = 1; v < (int)SUBSETS; ++v)
            if (popcount7(v) == m) t.mask[n++] = (uint8_t)v;
    }
    t.off[LEAVES + 1] = (uint8_t)n;
    return t;
}
__constant__ SubsetTable c_subsets = make_subset_table();
// Lanes that share one subset's partitions, as a shift, per subset size: the
// subsets of one size and their 2^(m-1) - 1 partitions spread over the wave
// (21 x 1, 35 x 1, 35 x 1, 21 x 2, 7 x 8 and 1 x 64 lanes).
constexpr uint32_t GROUP_SHIFT[LEAVES + 1] = {0, 0, 0, 0, 0, 1, 3, 6};

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

// The bits of j placed, lowest first, at the set bits of mask.
__device__ inline uint32_t deposit(uint32_t j, uint32_t mask) {
    uint32_t out = 0;
    for (uint32_t b = 1; mask; b <<= 1) {
        const uint32_t low = mask & (0u - mask);
        if (j & b) out |= low;
        mask ^= low;
    }
    return out;
}

// The least-cost subtrees over all subsets of m leaves, from the costs of
// their proper subsets (already in LDS). A subset's partitions keep its
// lowest leaf on the left; partition j's left side is that leaf plus the
// bits of j deposited on the others, so ascending j is ascending mask. The
// split chosen is the least (cost, imbalance, mask), imbalance being the
// difference of the two sides' primitive counts, with costs compared to 20
// mantissa bits (cost_key): among costs equal up to their rounding (boxes
// that coincide, summed in different orders) the most even split, so ties do
// not build chains. The lanes sharing a subset take every G-th partition,
// keep their first least triple, and then agree on the least. A NaN cost is
// never chosen; with every cost NaN the split is the lowest leaf alone.
// A cost with its 4 lowest mantissa bits cleared: monotone in the cost
// (costs are not negative), so comparing keys is a total order.
__device__ inline float cost_key(float c) { return __uint_as_float(__float_as_uint(c) & 0xFFFFFFF0u); }

__device__ inline void solve_size(TreeletLds& L, uint32_t lane, int m, uint32_t full) {
    const uint32_t gs = GROUP_SHIFT[m];
    const uint32_t g = lane >> gs, r = lane & ((1u << gs) - 1u);
    const uint32_t first = c_subsets.off[m], nsub = c_subsets.off[m + 1] - first;
    const float inf = __builtin_huge_valf();
    float key = inf;
    uint32_t bp = 0xFFu, bim = 0xFFFFFFFFu, s = 0;
    if (g < nsub) {
        s = c_subsets.mask[first + g];
        if (s <= full) {
            const uint32_t low = s & (0u - s), rest = s ^ low;
            const uint32_t parts = (1u << (m - 1)) - 1u;
            for (uint32_t j = r; j < parts; j += 1u << gs) {
                const uint32_t pm = low | deposit(j, rest);
                const float cand = L.cost[pm] + L.cost[s ^ pm];
                if (!(cand == cand)) continue;  // NaN
                const float qc = cost_key(cand);
                if (qc <= key) {  // the counts are read only for a lower or equal key
                    const uint32_t np = L.count[pm], nq = L.count[s ^ pm];
                    const uint32_t im = np > nq ? np - nq : nq - np;
                    if (qc < key || im < bim) {
                        key = qc;
                        bim = im;
                        bp = pm;
                    }
                }
            }
        }
    }
    for (uint32_t o = 1; o < (1u << gs); o <<= 1) {
        const float ok = __shfl_xor(key, o, 64);
        const uint32_t oi = __shfl_xor(bim, o, 64);
        const uint32_t op = __shfl_xor(bp, o, 64);
        if (ok < key || (ok == key && (oi < bim || (oi == bim && op < bp)))) {
            key = ok;
            bim = oi;
            bp = op;
        }
    }
    if (r == 0 && g < nsub && s <= full) {
        float best = inf;
        if (bp == 0xFFu)
            bp = s & (0u - s);
        else
            best = L.cost[bp] + L.cost[s ^ bp];
        const float a = area(L.box[0][s], L.box[1][s], L.box[2][s], L.box[3][s], L.box[4][s], L.box[5][s]);
        const float inner = CI * a + best;
        const float leaf = CT * a * (float)L.count[s];
        const bool as_leaf = L.count[s] <= MAX_LEAF && leaf <= inner;
        L.cost[s] = as_leaf ? leaf : inner;
        L.part[s] = bp | (as_leaf ? LEAF_BIT : 0u);
        L.icnt[s] = as_leaf ? 0u : 1u + L.icnt[bp] + L.icnt[s ^ bp];
    }
}

// The whole wave rebuilds the treelet under node x (uniform).
__device__ void restructure(const TreeArrays& t, TreeletLds& L, uint32_t lane, uint32_t x) {
    wave_sync();  // the previous treelet's LDS reads are done
    // The three levels under x, fetched across lanes first: heap slot p < 14
    // (lane p) holds a node id and its area (-1 for a primitive, which is
    // never opened); slots 0 and 1 are x's children and slot p >= 2 is child
    // p & 1 of slot (p - 2) / 2.
    const uint2 c0 = t.child[x];
    uint32_t hid = lane == 0 ? c0.x : (lane == 1 ? c0.y : NONE);
    for (uint32_t lo = 2, hi = 6; lo < 14; lo = hi, hi = 2 * hi + 2) {
        const uint32_t par = __shfl(hid, lane >= 2 ? (lane - 2) >> 1 : 0u, 64);
        if (lane >= lo && lane < hi && par < t.leaf_base) {
            const uint2 c = t.child[par];
            hid = (lane & 1u) ? c.y : c.x;
        }
    }
    const bool hin = lane < 14 && hid < t.leaf_base;
    const float harea = hin ? node_area(t.box, hid) : -1.0f;
    const uint32_t hsize = hin ? t.size[hid] : 0u;
    // Grow: lane i < k holds leaf i (id, area, primitives, heap slot or NONE);
    // the leaf of largest area, then most primitives (the first such), is
    // replaced by its left child and its right child appended.
    uint32_t id = __shfl(hid, lane < 2 ? lane : 0u, 64);
    float id_area = __shfl(harea, lane < 2 ? lane : 0u, 64);
    uint32_t id_size = __shfl(hsize, lane < 2 ? lane : 0u, 64);
    uint32_t hpos = lane;
    if (lane >= 2) {
        id = NONE;
        id_area = -1.0f;
        id_size = 0;
        hpos = NONE;
    }
    uint32_t inner_id = lane == 0 ? x : NONE;
    int k = 2;
    while (k < LEAVES) {
        int best = -1;
        float best_a = -1.0f;
        uint32_t best_n = 0;
        for (int i = 0; i < k; ++i) {
            const float a = __shfl(id_area, i, 64);
            const uint32_t sz = __shfl(id_size, i, 64);
            if (a >= 0.0f && (a > best_a || (a == best_a && sz > best_n))) {
                best_a = a;
                best_n = sz;
                best = i;
            }
        }
        if (best < 0) break;
        const uint32_t e = __shfl(id, best, 64);
        const uint32_t ep = __shfl(hpos, best, 64);
        uint32_t cx, cy, px = NONE, py = NONE, nx, ny;
        float ax, ay;
        if (ep < 6) {  // its children are in the heap
            px = 2 * ep + 2;
            py = px + 1;
            cx = __shfl(hid, px, 64);
            cy = __shfl(hid, py, 64);
            ax = __shfl(harea, px, 64);
            ay = __shfl(harea, py, 64);
            nx = __shfl(hsize, px, 64);
            ny = __shfl(hsize, py, 64);
        } else {
            const uint2 ce = t.child[e];
            cx = ce.x;
            cy = ce.y;
            ax = cx < t.leaf_base ? node_area(t.box, cx) : -1.0f;
            ay = cy < t.leaf_base ? node_area(t.box, cy) : -1.0f;
            nx = cx < t.leaf_base ? t.size[cx] : 0u;
            ny = cy < t.leaf_base ? t.size[cy] : 0u;
        }
        if (lane == (uint32_t)(k - 1)) inner_id = e;
        if (lane == (uint32_t)best) {
            id = cx;
            id_area = ax;
            id_size = nx;
            hpos = px;
        }
        if (lane == (uint32_t)k) {
            id = cy;
            id_area = ay;
            id_size = ny;
            hpos = py;
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
        L.icnt[1u << lane] = inner_of(t, id);
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
    for (int m = 2; m <= k; ++m) {
        solve_size(L, lane, m, full);
        wave_sync();
    }
    // Rebuild in pre-order, left side first, reusing the treelet's internal
    // nodes in the order they were taken.
    if (lane == 0) {
        uint32_t sp = 1, next = 1;
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
            t.leafy[slot] = (pw & LEAF_BIT) ? 1u : 0u;
            t.inner[slot] = L.icnt[s];
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
    }
}

// Longest walk to the root a valid tree allows (its depth is far below this);
// a longer one means broken parent links, and the pass stops and reports it.
constexpr uint32_t MAX_CLIMB = 1u << 16;
constexpr uint32_t MAX_LEVELS = 128;  // depth limit of the level schedule

// One node's cost, leaf choice and emitted-internal count from its children's.
__device__ inline void combine(const TreeArrays& t, uint32_t x) {
    const uint2 c = t.child[x];
    const float a = node_area(t.box, x);
    const float split = t.cost[c.x] + t.cost[c.y];
    const float inner = CI * a + split;
    const uint32_t sz = t.size[x];
    const float leaf = CT * a * (float)sz;
    const bool as_leaf = sz <= MAX_LEAF && leaf <= inner;
    t.cost[x] = as_leaf ? leaf : inner;
    t.leafy[x] = as_leaf ? 1u : 0u;
    t.inner[x] = as_leaf ? 0u : 1u + inner_of(t, c.x) + inner_of(t, c.y);
}

// A round is scheduled by depth: every internal node's depth (its walk to the
// root) and class (treelet root or not) are counted per level, the nodes are
// bucketed into per-level lists, and one launch per level, deepest first,
// handles that level's nodes. A node's descendants are all deeper, so they
// are final when it is handled, and rebuilding a treelet touches only nodes
// below its root; launch boundaries order the levels, so no cross-wave
// synchronisation is needed inside a launch.
__global__ __launch_bounds__(256) void k_tl_depth(TreeArrays t, uint32_t* __restrict__ depth,
                                                  uint32_t* __restrict__ level_count, uint32_t* __restrict__ fault) {
    __shared__ uint32_t h[2 * MAX_LEVELS];
    for (int i = threadIdx.x; i < 2 * (int)MAX_LEVELS; i += 256) h[i] = 0;
    __syncthreads();
    const uint32_t x = blockIdx.x * 256 + threadIdx.x;
    if (x < t.n - 1) {
        uint32_t d = 0, cur = x;
        while (cur != 0 && d < MAX_LEVELS) {
            cur = t.parent[cur];
            ++d;
        }
        if (d >= MAX_LEVELS) {
            atomicOr(fault, 8u);
            d = MAX_LEVELS - 1;
        }
        const uint32_t key = 2 * d + (t.size[x] >= MIN_SIZE ? 1u : 0u);
        depth[x] = key;
        atomicAdd(&h[key], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 2 * (int)MAX_LEVELS; i += 256)
        if (h[i]) atomicAdd(&level_count[i], h[i]);
}

// Scatters node ids into the per-(level, class) lists; cursor[key] starts at
// that list's offset.
__global__ __launch_bounds__(256) void k_tl_bucket(uint32_t internal, const uint32_t* __restrict__ depth,
                                                   uint32_t* __restrict__ cursor, uint32_t* __restrict__ list) {
    __shared__ uint32_t h[2 * MAX_LEVELS], base[2 * MAX_LEVELS];
    for (int i = threadIdx.x; i < 2 * (int)MAX_LEVELS; i += 256) h[i] = 0;
    __syncthreads();
    const uint32_t x = blockIdx.x * 256 + threadIdx.x;
    uint32_t key = 0, local = 0;
    if (x < internal) {
        key = depth[x];
        local = atomicAdd(&h[key], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 2 * (int)MAX_LEVELS; i += 256) base[i] = h[i] ? atomicAdd(&cursor[i], h[i]) : 0u;
    __syncthreads();
    if (x < internal) list[base[key] + local] = x;
}

// One level: the first `waves` waves each rebuild the treelet of one node of
// `big` (nodes of at least MIN_SIZE primitives); the lanes of the remaining
// waves each combine one node of `small`.
__global__ __launch_bounds__(TL_BLOCK) void k_tl_level(TreeArrays t, const uint32_t* __restrict__ big,
                                                       uint32_t waves, const uint32_t* __restrict__ small,
                                                       uint32_t count) {
    __shared__ TreeletLds lds[TL_WAVES];
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t w = blockIdx.x * TL_WAVES + threadIdx.x / 64;
    if (w < waves) {
        restructure(t, lds[threadIdx.x / 64], lane, big[w]);
        return;
    }
    const uint32_t i = (w - waves) * 64 + lane;
    if (i < count) combine(t, small[i]);
}

__global__ __launch_bounds__(256) void k_tl_leaf_cost(TreeArrays t) {
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k < t.n) t.cost[t.leaf_base + k] = CT * node_area(t.box, t.leaf_base + k);
}

__global__ __launch_bounds__(256) void k_clear_words(uint32_t* __restrict__ p, uint32_t count) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < count) p[i] = 0u;
}

// Bucket cursors: the exclusive prefix sum of the per-(level, class) counts
// (one lane; 256 entries).
__global__ void k_tl_offsets(const uint32_t* __restrict__ count, uint32_t* __restrict__ cursor) {
    if (threadIdx.x != 0) return;
    uint32_t acc = 0;
    for (uint32_t k = 0; k < 2 * MAX_LEVELS; ++k) {
        cursor[k] = acc;
        acc += count[k];
    }
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
    if (stage_us) stage_us[0] = stage_us[1] = stage_us[2] = 0.0f;
    if (rounds > LUCIDA_TREELET_MAX_ROUNDS)
        return set_error(LUCIDA_INVALID_ARG, "treelet BVH: more than LUCIDA_TREELET_MAX_ROUNDS (64) rounds");
    if (n < 2)  // nothing to restructure: the LBVH's single leaf
        return lbvh_build_device(ctx, tris, n, positions, nodes, num_nodes, prim_indices, tris_sorted, nullptr);
    if (n > (1u << 30)) return set_error(LUCIDA_INVALID_ARG, "treelet BVH: more than 2^30 triangles");
    hipStream_t s = ctx->stream;
    StageEvents<4> ev;
    if (stage_us) {
        LUCIDA_HIP(ev.create(), LUCIDA_VALIDATION);
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
    uint32_t* fault = t.leafy + internal;
    LUCIDA_HIP(ctx->tl_list.ensure((internal + 4 * MAX_LEVELS) * 4), LUCIDA_OOM);
    uint32_t* list = ctx->tl_list.as<uint32_t>();
    uint32_t* level_count = list + internal;
    uint32_t* cursor = level_count + 2 * MAX_LEVELS;
    uint32_t* key = ctx->lb_flags.as<uint32_t>();  // per internal node: 2 x depth + treelet root
    hipLaunchKernelGGL(k_clear_words, dim3(1), dim3(256), 0, s, fault, 1u);
    hipLaunchKernelGGL(k_tl_leaf_cost, dim3(blocks_for(n, 256)), dim3(256), 0, s, t);
    for (uint32_t r = 0; r < (rounds ? rounds : 1u); ++r) {
        hipLaunchKernelGGL(k_clear_words, dim3(1), dim3(256), 0, s, level_count, 2 * MAX_LEVELS);
        hipLaunchKernelGGL(k_tl_depth, dim3(blocks_for(internal, 256)), dim3(256), 0, s, t, key, level_count, fault);
        uint32_t host[2 * MAX_LEVELS + 1];
        LUCIDA_HIP(hipMemcpyAsync(host, level_count, 2 * MAX_LEVELS * 4, hipMemcpyDeviceToHost, s), LUCIDA_READBACK);
        LUCIDA_HIP(hipMemcpyAsync(host + 2 * MAX_LEVELS, fault, 4, hipMemcpyDeviceToHost, s), LUCIDA_READBACK);
        LUCIDA_HIP(hipStreamSynchronize(s), LUCIDA_VALIDATION);
        if (host[2 * MAX_LEVELS]) return set_error(LUCIDA_VALIDATION, "treelet BVH: deeper than 128 levels");
        uint32_t off[2 * MAX_LEVELS], acc = 0;
        int deepest = 0;
        for (uint32_t k = 0; k < 2 * MAX_LEVELS; ++k) {
            off[k] = acc;
            acc += host[k];
            if (host[k]) deepest = (int)(k / 2);
        }
        // the device forms the same offsets as its bucket cursors, in stream order
        hipLaunchKernelGGL(k_tl_offsets, dim3(1), dim3(64), 0, s, level_count, cursor);
        hipLaunchKernelGGL(k_tl_bucket, dim3(blocks_for(internal, 256)), dim3(256), 0, s, (uint32_t)internal, key,
                           cursor, list);
        for (int d = deepest; d >= 0; --d) {
            const uint32_t nsmall = host[2 * d], nbig = host[2 * d + 1];
            if (!nsmall && !nbig) continue;
            const uint32_t waves = nbig + (nsmall + 63) / 64;
            hipLaunchKernelGGL(k_tl_level, dim3(blocks_for(waves, TL_WAVES)), dim3(TL_BLOCK), 0, s, t,
                               list + off[2 * d + 1], nbig, list + off[2 * d], nsmall);
        }
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

lucida_status lucida_bvh_build_treelet_device(lucida_ctx* ctx, const GpuTriangle* tris, size_t n,
                                              const float* positions, uint32_t rounds, GpuBvhNode* nodes,
                                              size_t* num_nodes, uint32_t* prim_indices, GpuTriangle* tris_sorted,
                                              float* stage_us) {
    if (!ctx || !num_nodes || (n && (!tris || !positions || !nodes)))
        return set_error(LUCIDA_INVALID_ARG, "lucida_bvh_build_treelet_device: null argument");
    clear_error();
    if (n == 0) {
        *num_nodes = 0;
        return LUCIDA_OK;
    }
    return bvh_build_treelet_device(ctx, tris, n, positions, rounds, nodes, num_nodes, prim_indices, tris_sorted,
                                    stage_us);
}

}  // extern "C"
