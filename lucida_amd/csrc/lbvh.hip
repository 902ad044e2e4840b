// GPU linear BVH for gfx950: prepare (padded triangle bounds + centroid
// bounds), 30-bit Morton codes, an LDS-staged LSD radix sort ranked with
// wave64 ballots, the Karras hierarchy, an atomic bottom-up fit, and a relayout
// into the reference's 32-byte GpuBvhNode tree, all on the device.
//
// Reference algorithm: crates/core/src/lbvh.rs (CPU twin) and
// shaders/lbvh/{prepare,radix,hierarchy,fit}.wgsl; host driver
// crates/gpu/src/lbvh.rs:369-665. Morton codes, the sorted order, the Karras
// topology, node boxes and the relayout match the CPU twin bit for bit; this
// file is compiled with -ffp-contract=off and correctly rounded division so
// the pad / centroid / normalise arithmetic is the twin's.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <vector>

#include "context.hpp"

namespace lucida {
namespace {

constexpr uint32_t NONE = 0xFFFFFFFFu;
constexpr uint32_t MAX_LEAF = 4;  // lbvh.rs:185 LBVH_MAX_LEAF

constexpr int PREP_BLOCK = 256;
constexpr int SORT_BLOCK = 256;
constexpr int SORT_WAVES = SORT_BLOCK / 64;
constexpr int SORT_ITEMS = 8;
constexpr int SORT_TILE = SORT_BLOCK * SORT_ITEMS;
constexpr int RADIX = 256;

__device__ inline float fmin3(float a, float b, float c) { return fminf(fminf(a, b), c); }

// ---------------------------------------------------------------------------
// prepare: triangle_bounds (bvh.rs:202, prepare.wgsl:48) and the centroid
// bounds (lbvh.rs:446-449). Each block reduces its centroids with wave
// reductions; the last block to finish folds the per-block partials, so the
// result needs no float atomics and is deterministic.
// ---------------------------------------------------------------------------

__device__ inline float wave_min(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ inline float wave_max(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

__global__ __launch_bounds__(PREP_BLOCK) void k_prepare(const GpuTriangle* __restrict__ tris,
                                                        const float4* __restrict__ pos, uint32_t n,
                                                        LucidaAabb* __restrict__ tri_bounds,
                                                        float* __restrict__ partial) {
    const uint32_t i = blockIdx.x * PREP_BLOCK + threadIdx.x;
    float c[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    if (i < n) {
        GpuTriangle t = tris[i];
        float4 p0 = pos[t.i0], p1 = pos[t.i1], p2 = pos[t.i2];
        float lo[3] = {fminf(fminf(fminf(INFINITY, p0.x), p1.x), p2.x), fminf(fminf(fminf(INFINITY, p0.y), p1.y), p2.y),
                       fminf(fminf(fminf(INFINITY, p0.z), p1.z), p2.z)};
        float hi[3] = {fmaxf(fmaxf(fmaxf(-INFINITY, p0.x), p1.x), p2.x),
                       fmaxf(fmaxf(fmaxf(-INFINITY, p0.y), p1.y), p2.y),
                       fmaxf(fmaxf(fmaxf(-INFINITY, p0.z), p1.z), p2.z)};
        if (!(lo[0] > hi[0])) {  // Aabb::pad skips an empty box (bvh.rs:100)
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                float scale = 1.0f + fmaxf(fabsf(lo[a]), fabsf(hi[a]));
                float p = 1.0e-6f * scale;
                lo[a] = lo[a] - p;
                hi[a] = hi[a] + p;
            }
        }
        LucidaAabb b;
        b.lo[0] = lo[0]; b.lo[1] = lo[1]; b.lo[2] = lo[2]; b._p0 = 0.0f;
        b.hi[0] = hi[0]; b.hi[1] = hi[1]; b.hi[2] = hi[2]; b._p1 = 0.0f;
        tri_bounds[i] = b;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float ca = (lo[a] + hi[a]) * 0.5f;
            c[a] = fminf(INFINITY, ca);
            c[3 + a] = fmaxf(-INFINITY, ca);
        }
    }
    __shared__ float red[PREP_BLOCK / 64][6];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        c[a] = wave_min(c[a]);
        c[3 + a] = wave_max(c[3 + a]);
    }
    if (lane == 0)
        for (int a = 0; a < 6; ++a) red[wave][a] = c[a];
    __syncthreads();
    if (threadIdx.x == 0) {
        float r[6];
        for (int a = 0; a < 6; ++a) r[a] = red[0][a];
        for (int w = 1; w < PREP_BLOCK / 64; ++w)
            for (int a = 0; a < 3; ++a) {
                r[a] = fminf(r[a], red[w][a]);
                r[3 + a] = fmaxf(r[3 + a], red[w][3 + a]);
            }
        for (int a = 0; a < 6; ++a) partial[blockIdx.x * 6 + a] = r[a];
    }
}

// Folds the per-block centroid bounds into the scene's (one block). A launch
// of its own: a last-block-done ticket would cost every block of k_prepare a
// device-scope fence and an atomic on one address.
// Also clears what the later stages count or mark, saving them launches:
// the root's parent (`root_parent`, when given) and the two-phase fit's list
// counters (`fit_counts`, eight words, when given).
__global__ __launch_bounds__(PREP_BLOCK) void k_fold_bounds(const float* __restrict__ partial, uint32_t blocks,
                                                            LucidaAabb* __restrict__ scene_box,
                                                            uint32_t* __restrict__ root_parent,
                                                            uint32_t* __restrict__ fit_counts) {
    if (root_parent && threadIdx.x == 0) *root_parent = 0xFFFFFFFFu;
    if (fit_counts && threadIdx.x < 8) fit_counts[threadIdx.x] = 0;
    __shared__ float red[PREP_BLOCK / 64][6];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    float r[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (uint32_t b = threadIdx.x; b < blocks; b += PREP_BLOCK) {
        for (int a = 0; a < 3; ++a) {
            r[a] = fminf(r[a], partial[b * 6 + a]);
            r[3 + a] = fmaxf(r[3 + a], partial[b * 6 + 3 + a]);
        }
    }
    for (int a = 0; a < 3; ++a) {
        r[a] = wave_min(r[a]);
        r[3 + a] = wave_max(r[3 + a]);
    }
    if (lane == 0)
        for (int a = 0; a < 6; ++a) red[wave][a] = r[a];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < PREP_BLOCK / 64; ++w)
            for (int a = 0; a < 3; ++a) {
                red[0][a] = fminf(red[0][a], red[w][a]);
                red[0][3 + a] = fmaxf(red[0][3 + a], red[w][3 + a]);
            }
        LucidaAabb sb;
        sb.lo[0] = red[0][0]; sb.lo[1] = red[0][1]; sb.lo[2] = red[0][2]; sb._p0 = 0.0f;
        sb.hi[0] = red[0][3]; sb.hi[1] = red[0][4]; sb.hi[2] = red[0][5]; sb._p1 = 0.0f;
        *scene_box = sb;
    }
}

// ---------------------------------------------------------------------------
// Morton codes: normalise (lbvh.rs:109) then morton3d (lbvh.rs:92).
// ---------------------------------------------------------------------------

__device__ inline uint32_t expand_bits(uint32_t v) {
    uint32_t x = v & 0x000003FFu;
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

// clamp(v * 1023, 0, 1023) then a saturating cast to u32 with NaN -> 0, as
// Rust's `as u32` does.
__device__ inline uint32_t quantise(float v) {
    float s = v * 1023.0f;
    if (s < 0.0f) s = 0.0f;
    if (s > 1023.0f) s = 1023.0f;
    return s >= 0.0f ? (uint32_t)s : 0u;
}

__global__ __launch_bounds__(PREP_BLOCK) void k_morton(const LucidaAabb* __restrict__ tri_bounds, uint32_t n,
                                                       const LucidaAabb* __restrict__ scene_box,
                                                       uint32_t* __restrict__ codes, uint32_t* __restrict__ vals) {
    const uint32_t i = blockIdx.x * PREP_BLOCK + threadIdx.x;
    if (i >= n) return;
    const LucidaAabb sb = *scene_box;
    const LucidaAabb b = tri_bounds[i];
    uint32_t q[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float ca = (b.lo[a] + b.hi[a]) * 0.5f;
        float e = sb.hi[a] - sb.lo[a];
        float nv = e > 0.0f ? (ca - sb.lo[a]) / e : 0.5f;
        q[a] = quantise(nv);
    }
    codes[i] = (expand_bits(q[0]) << 2) | (expand_bits(q[1]) << 1) | expand_bits(q[2]);
    vals[i] = i;
}

// ---------------------------------------------------------------------------
// LSD radix sort, 8-bit digits. Per pass: per-tile digit histograms (LDS
// atomics), one exclusive scan over the digit-major histogram, then a stable
// scatter whose in-tile ranks come from wave64 ballots: eight ballots give
// every lane the set of lanes holding its digit, a popcount below the lane
// gives its rank in the wave, and a per-digit prefix over the block's waves
// (in LDS) orders the waves. Tiles, rounds, waves and lanes are visited in
// input order, so equal keys keep their order.
// ---------------------------------------------------------------------------

// `count`, when given, is a device word that caps n (a queue whose length
// only the device knows; the grid covers n).
__global__ __launch_bounds__(SORT_BLOCK) void k_radix_hist(const uint32_t* __restrict__ keys, uint32_t n,
                                                           const uint32_t* __restrict__ count, uint32_t shift,
                                                           uint32_t* __restrict__ hist, uint32_t tiles) {
    __shared__ uint32_t h[RADIX];
    h[threadIdx.x] = 0;
    if (count) n = min(n, *count);
    __syncthreads();
    const uint32_t base = blockIdx.x * SORT_TILE;
    if (base < n) {
#pragma unroll
        for (int r = 0; r < SORT_ITEMS; ++r) {
            uint32_t idx = base + r * SORT_BLOCK + threadIdx.x;
            if (idx < n) atomicAdd(&h[(keys[idx] >> shift) & (RADIX - 1)], 1u);
        }
    }
    __syncthreads();
    hist[threadIdx.x * tiles + blockIdx.x] = h[threadIdx.x];
}

// Exclusive scan of the digit-major histogram, in two launches: per-segment
// sums (k_scan_reduce), then a scan of each segment from its carry, the sum
// of the segment sums before it, which each block adds up itself
// (k_scan_down). Every global access is coalesced; a
// segment is SCAN_ITEMS words per thread of a SCAN_BLOCK block.
constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_SEG = SCAN_BLOCK * SCAN_ITEMS;

// Inclusive scan across the block of one value per thread (wave64 shuffles,
// then the wave totals through LDS). `wsum` holds SCAN_BLOCK / 64 words.
__device__ inline uint32_t block_inclusive_scan(uint32_t v, uint32_t* wsum, uint32_t* total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    if (lane == 63) wsum[wave] = v;
    __syncthreads();
    uint32_t before = 0, all = 0;
#pragma unroll
    for (int w = 0; w < SCAN_BLOCK / 64; ++w) {
        const uint32_t s = wsum[w];
        if (w < wave) before += s;
        all += s;
    }
    __syncthreads();
    *total = all;
    return v + before;
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_reduce(const uint32_t* __restrict__ data, uint32_t count,
                                                            uint32_t* __restrict__ partials) {
    __shared__ uint32_t wsum[SCAN_BLOCK / 64];
    const uint32_t base = blockIdx.x * SCAN_SEG;
    uint32_t s = 0;
#pragma unroll
    for (int r = 0; r < SCAN_ITEMS; ++r) {
        const uint32_t idx = base + r * SCAN_BLOCK + threadIdx.x;
        if (idx < count) s += data[idx];
    }
    uint32_t total;
    (void)block_inclusive_scan(s, wsum, &total);
    if (threadIdx.x == 0) partials[blockIdx.x] = total;
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_down(uint32_t* __restrict__ data, uint32_t count,
                                                          const uint32_t* __restrict__ partials) {
    __shared__ uint32_t seg[SCAN_SEG];
    __shared__ uint32_t wsum[SCAN_BLOCK / 64];
    const uint32_t base = blockIdx.x * SCAN_SEG;
#pragma unroll
    for (int r = 0; r < SCAN_ITEMS; ++r) {
        const uint32_t idx = base + r * SCAN_BLOCK + threadIdx.x;
        seg[r * SCAN_BLOCK + threadIdx.x] = idx < count ? data[idx] : 0u;
    }
    __syncthreads();
    // thread t owns segment words [t * SCAN_ITEMS, (t + 1) * SCAN_ITEMS)
    uint32_t local[SCAN_ITEMS];
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        local[i] = seg[threadIdx.x * SCAN_ITEMS + i];
        s += local[i];
    }
    // this segment's carry: the sums of the segments before it
    uint32_t before = 0;
    for (uint32_t b = threadIdx.x; b < blockIdx.x; b += SCAN_BLOCK) before += partials[b];
    uint32_t carry, total;
    (void)block_inclusive_scan(before, wsum, &carry);
    uint32_t run = carry + block_inclusive_scan(s, wsum, &total) - s;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        seg[threadIdx.x * SCAN_ITEMS + i] = run;
        run += local[i];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < SCAN_ITEMS; ++r) {
        const uint32_t idx = base + r * SCAN_BLOCK + threadIdx.x;
        if (idx < count) data[idx] = seg[r * SCAN_BLOCK + threadIdx.x];
    }
}

// The scatter stages its tile in LDS: items are ranked within the tile by
// digit (ballots within a wave, per-digit prefixes over the block's waves and
// rounds), placed in LDS in digit order, and then written out so that
// consecutive threads write consecutive addresses of one digit's run (an
// item at tile position i of digit d goes to the digit's global offset plus
// i minus the digit's first tile position). Without the staging, the 256
// lanes of a round write to up to 256 different runs.
__global__ __launch_bounds__(SORT_BLOCK) void k_radix_scatter(const uint32_t* __restrict__ keys,
                                                              const uint32_t* __restrict__ vals, uint32_t n,
                                                              const uint32_t* __restrict__ count, uint32_t shift,
                                                              const uint32_t* __restrict__ offsets, uint32_t tiles,
                                                              uint32_t* __restrict__ keys_out,
                                                              uint32_t* __restrict__ vals_out) {
    __shared__ uint32_t s_key[SORT_TILE];
    __shared__ uint32_t s_val[SORT_TILE];
    __shared__ uint32_t wave_cnt[SORT_WAVES][RADIX];
    __shared__ uint32_t running[RADIX];   // per digit: items of the tile ranked so far
    __shared__ uint32_t out_base[RADIX];  // per digit: global offset minus first tile position
    __shared__ uint32_t wsum[SORT_BLOCK / 64];
    if (count) n = min(n, *count);
    const uint32_t base = blockIdx.x * SORT_TILE;
    if (base >= n) return;  // the whole block: nothing of this tile is queued
    const uint32_t in_tile = min((uint32_t)SORT_TILE, n - base);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    running[threadIdx.x] = 0;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    uint32_t key[SORT_ITEMS], val[SORT_ITEMS], pos[SORT_ITEMS];
#pragma unroll
    for (int r = 0; r < SORT_ITEMS; ++r) {
#pragma unroll
        for (int w = 0; w < SORT_WAVES; ++w) wave_cnt[w][threadIdx.x] = 0;
        const uint32_t idx = base + r * SORT_BLOCK + threadIdx.x;
        const bool valid = idx < n;
        key[r] = valid ? keys[idx] : 0u;
        val[r] = valid ? vals[idx] : 0u;
        const uint32_t digit = (key[r] >> shift) & (RADIX - 1);
        uint64_t same = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (digit >> b) & 1u;
            const uint64_t m = __ballot(bit);
            same &= bit ? m : ~m;
        }
        const uint32_t rank = __popcll(same & lt_mask);
        __syncthreads();  // zeroing of wave_cnt done
        if (valid && rank == 0) wave_cnt[wave][digit] = (uint32_t)__popcll(same);
        __syncthreads();
        {
            uint32_t acc = running[threadIdx.x];
#pragma unroll
            for (int w = 0; w < SORT_WAVES; ++w) {
                const uint32_t t = wave_cnt[w][threadIdx.x];
                wave_cnt[w][threadIdx.x] = acc;
                acc += t;
            }
            running[threadIdx.x] = acc;
        }
        __syncthreads();
        pos[r] = valid ? wave_cnt[wave][digit] + rank : 0xFFFFFFFFu;  // rank among the tile's items of its digit
        __syncthreads();  // wave_cnt is read above before the next round clears it
    }
    // first tile position of each digit, and where its run goes
    const uint32_t cnt = running[threadIdx.x];
    uint32_t total;
    const uint32_t first = block_inclusive_scan(cnt, wsum, &total) - cnt;
    running[threadIdx.x] = first;
    out_base[threadIdx.x] = offsets[threadIdx.x * tiles + blockIdx.x] - first;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < SORT_ITEMS; ++r) {
        if (pos[r] != 0xFFFFFFFFu) {
            const uint32_t at = running[(key[r] >> shift) & (RADIX - 1)] + pos[r];
            s_key[at] = key[r];
            s_val[at] = val[r];
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < in_tile; i += SORT_BLOCK) {
        const uint32_t k = s_key[i];
        const uint32_t dst = out_base[(k >> shift) & (RADIX - 1)] + i;
        if (keys_out) keys_out[dst] = k;
        vals_out[dst] = s_val[i];
    }
}

// ---------------------------------------------------------------------------
// Karras hierarchy (lbvh.rs:211-292, hierarchy.wgsl:46-117): one thread per
// internal node; leaves are ids n-1 .. 2n-2.
// ---------------------------------------------------------------------------

__device__ inline int lcp(const uint32_t* __restrict__ codes, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    uint32_t a = codes[i], b = codes[j];
    if (a == b) {
        uint32_t x = (uint32_t)i ^ (uint32_t)j;
        return 32 + (x == 0 ? 32 : __clz(x));
    }
    return __clz(a ^ b);
}

// With `lists` (the two-phase fit), each node is also classified for it: a
// node whose subtree holds more than TREELET leaves is a "top" node (listed,
// with a slot), and each child of a top node with at most TREELET leaves is a
// treelet root (listed); a whole tree of at most TREELET leaves is one
// treelet. The children's subtree sizes follow from the split. counts[0] =
// treelets, counts[1] = top nodes (zeroed before the launch).
constexpr uint32_t TREELET = 1024;
constexpr uint32_t TOP_LDS = 3072;  // top slots held in LDS (48 bytes each)

struct FitLists {
    uint32_t* counts;
    uint32_t* treelets;
    uint32_t* top;
    uint32_t* slot_of;
};

__global__ __launch_bounds__(256) void k_karras(const uint32_t* __restrict__ codes, uint32_t n,
                                                uint2* __restrict__ karras, uint2* __restrict__ ranges,
                                                uint32_t* __restrict__ parent, FitLists lists) {
    const uint32_t gi = blockIdx.x * 256 + threadIdx.x;
    if (n < 2 || gi >= n - 1) return;
    const int N = (int)n, i = (int)gi;
    int dl = lcp(codes, N, i, i + 1) - lcp(codes, N, i, i - 1);
    int d = (dl > 0) - (dl < 0);
    int delta_min = lcp(codes, N, i, i - d);
    int l_max = 2;
    while (lcp(codes, N, i, i + l_max * d) > delta_min) l_max *= 2;
    int l = 0;
    for (int t = l_max / 2; t >= 1; t /= 2)
        if (lcp(codes, N, i, i + (l + t) * d) > delta_min) l += t;
    int j = i + l * d;
    int first = i < j ? i : j, last = i < j ? j : i;
    int delta_node = lcp(codes, N, i, j);
    int s = 0, t = l;
    while (t > 1) {
        t = (t + 1) / 2;
        if (lcp(codes, N, i, i + (s + t) * d) > delta_node) s += t;
    }
    int split = i + s * d + (d < 0 ? d : 0);
    const uint32_t leaf_base = n - 1;
    uint32_t left = split == first ? leaf_base + (uint32_t)split : (uint32_t)split;
    uint32_t right = split + 1 == last ? leaf_base + (uint32_t)(split + 1) : (uint32_t)(split + 1);
    karras[gi] = make_uint2(left, right);
    ranges[gi] = make_uint2((uint32_t)first, (uint32_t)last);
    parent[left] = gi;
    parent[right] = gi;
    if (lists.counts) {
        const uint32_t size = (uint32_t)(last - first + 1);
        if (size <= TREELET) {
            if (gi == 0) lists.treelets[atomicAdd(&lists.counts[0], 1u)] = 0;  // the whole tree is one treelet
            return;
        }
        // one atomic for both lists: the top count in the high word, treelets in the low
        const bool tl = (uint32_t)(split - first + 1) <= TREELET, tr = (uint32_t)(last - split) <= TREELET;
        const unsigned long long old = atomicAdd(reinterpret_cast<unsigned long long*>(lists.counts),
                                                 (1ull << 32) | (unsigned long long)((tl ? 1u : 0u) + (tr ? 1u : 0u)));
        const uint32_t slot = (uint32_t)(old >> 32), t0 = (uint32_t)old;
        lists.top[slot] = gi;
        lists.slot_of[gi] = slot;
        if (tl) lists.treelets[t0] = left;
        if (tr) lists.treelets[t0 + (tl ? 1u : 0u)] = right;
    }
}

// ---------------------------------------------------------------------------
// Atomic bottom-up fit (the CPU twin's visit counting, lbvh.rs:490-525): one
// thread per leaf climbs; at each internal node the first arrival stops and the
// second, which now sees both children, writes the union, the subtree size and
// the count of internal nodes the relayout will emit under it. Agent-scope
// acquire/release orders the child boxes across XCDs.
// ---------------------------------------------------------------------------

// Plain loads: they follow the acquiring atomic, which invalidated this CU's
// caches at agent scope.
__device__ inline void box_load(const LucidaAabb* p, float* lo, float* hi) {
    float4 a = reinterpret_cast<const float4*>(p)[0];
    float4 b = reinterpret_cast<const float4*>(p)[1];
    lo[0] = a.x; lo[1] = a.y; lo[2] = a.z;
    hi[0] = b.x; hi[1] = b.y; hi[2] = b.z;
}
__device__ inline void box_store(LucidaAabb* p, const float* lo, const float* hi) {
    reinterpret_cast<float4*>(p)[0] = make_float4(lo[0], lo[1], lo[2], 0.0f);
    reinterpret_cast<float4*>(p)[1] = make_float4(hi[0], hi[1], hi[2], 0.0f);
}

__global__ __launch_bounds__(256) void k_fit(const LucidaAabb* __restrict__ tri_bounds,
                                             const uint32_t* __restrict__ order, uint32_t n,
                                             const uint2* __restrict__ karras, const uint32_t* __restrict__ parent,
                                             LucidaAabb* node_box, uint32_t* subtree, uint32_t* inner,
                                             uint32_t* flags) {
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    const uint32_t leaf_base = n - 1;
    float lo[3], hi[3];
    {
        LucidaAabb b = tri_bounds[order[k]];
        for (int a = 0; a < 3; ++a) { lo[a] = b.lo[a]; hi[a] = b.hi[a]; }
    }
    box_store(&node_box[leaf_base + k], lo, hi);
    uint32_t node = parent[leaf_base + k];
    while (node != NONE) {
        // Release this child's box, count the arrival, and acquire the sibling's
        // box only when this thread is the second arrival.
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        uint32_t old = __hip_atomic_fetch_add(&flags[node], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (old == 0) return;  // first arrival: the sibling's thread finishes this node
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        const uint2 c = karras[node];
        float llo[3], lhi[3], rlo[3], rhi[3];
        box_load(&node_box[c.x], llo, lhi);
        box_load(&node_box[c.y], rlo, rhi);
        for (int a = 0; a < 3; ++a) {
            lo[a] = fminf(llo[a], rlo[a]);
            hi[a] = fmaxf(lhi[a], rhi[a]);
        }
        uint32_t sl = c.x >= leaf_base ? 1u : __hip_atomic_load(&subtree[c.x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        uint32_t sr = c.y >= leaf_base ? 1u : __hip_atomic_load(&subtree[c.y], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        uint32_t il = c.x >= leaf_base ? 0u : __hip_atomic_load(&inner[c.x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        uint32_t ir = c.y >= leaf_base ? 0u : __hip_atomic_load(&inner[c.y], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        uint32_t size = sl + sr;
        box_store(&node_box[node], lo, hi);
        __hip_atomic_store(&subtree[node], size, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&inner[node], size > MAX_LEAF ? 1u + il + ir : 0u, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
        node = parent[node];
    }
}

// ---------------------------------------------------------------------------
// Level-synchronous bottom-up fit: the same unions as the atomic fit, ordered
// by depth instead of by arrival. Each internal node's depth comes from walking
// its parents; nodes are bucketed by depth and one launch per level, deepest
// first, unions the children. Launch boundaries order the levels, so no
// agent-scope fences (and their L2 write-back / invalidate on every XCD) are
// needed.
// ---------------------------------------------------------------------------

constexpr int MAX_LEVELS = 128;

__global__ __launch_bounds__(256) void k_leaf_boxes(const LucidaAabb* __restrict__ tri_bounds,
                                                    const uint32_t* __restrict__ order, uint32_t n,
                                                    LucidaAabb* __restrict__ node_box) {
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k < n) node_box[n - 1 + k] = tri_bounds[order[k]];
}

// Depth of every internal node and a per-level histogram (LDS per block, one
// atomic per level per block). depth >= MAX_LEVELS raises the overflow flag.
__global__ __launch_bounds__(256) void k_depth(uint32_t n, const uint32_t* __restrict__ parent,
                                               uint32_t* __restrict__ depth, uint32_t* __restrict__ level_count,
                                               uint32_t* __restrict__ overflow) {
    __shared__ uint32_t h[MAX_LEVELS];
    for (int i = threadIdx.x; i < MAX_LEVELS; i += 256) h[i] = 0;
    __syncthreads();
    const uint32_t x = blockIdx.x * 256 + threadIdx.x;
    if (x < n - 1) {
        uint32_t d = 0, cur = x;
        while (cur != 0 && d < MAX_LEVELS) {
            cur = parent[cur];
            ++d;
        }
        if (d >= MAX_LEVELS) {
            atomicOr(overflow, 1u);
            d = MAX_LEVELS - 1;
        }
        depth[x] = d;
        atomicAdd(&h[d], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < MAX_LEVELS; i += 256)
        if (h[i]) atomicAdd(&level_count[i], h[i]);
}

// Bucket cursors: the exclusive prefix sum of the per-level counts (one lane;
// MAX_LEVELS entries).
__global__ void k_level_offsets(const uint32_t* __restrict__ count, uint32_t* __restrict__ cursor) {
    if (threadIdx.x != 0) return;
    uint32_t acc = 0;
    for (int l = 0; l < MAX_LEVELS; ++l) {
        cursor[l] = acc;
        acc += count[l];
    }
}

// Scatters node ids into per-level lists: cursor[d] starts at level d's offset.
__global__ __launch_bounds__(256) void k_bucket(uint32_t n, const uint32_t* __restrict__ depth,
                                                uint32_t* __restrict__ cursor, uint32_t* __restrict__ list) {
    __shared__ uint32_t h[MAX_LEVELS], base[MAX_LEVELS];
    for (int i = threadIdx.x; i < MAX_LEVELS; i += 256) h[i] = 0;
    __syncthreads();
    const uint32_t x = blockIdx.x * 256 + threadIdx.x;
    uint32_t d = 0, local = 0;
    if (x < n - 1) {
        d = depth[x];
        local = atomicAdd(&h[d], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < MAX_LEVELS; i += 256) base[i] = h[i] ? atomicAdd(&cursor[i], h[i]) : 0u;
    __syncthreads();
    if (x < n - 1) list[base[d] + local] = x;
}

__global__ __launch_bounds__(256) void k_fit_level(const uint32_t* __restrict__ list, uint32_t count, uint32_t n,
                                                   const uint2* __restrict__ karras, LucidaAabb* node_box,
                                                   uint32_t* subtree, uint32_t* inner) {
    const uint32_t t = blockIdx.x * 256 + threadIdx.x;
    if (t >= count) return;
    const uint32_t node = list[t];
    const uint32_t leaf_base = n - 1;
    const uint2 c = karras[node];
    float llo[3], lhi[3], rlo[3], rhi[3], lo[3], hi[3];
    box_load(&node_box[c.x], llo, lhi);
    box_load(&node_box[c.y], rlo, rhi);
    for (int a = 0; a < 3; ++a) {
        lo[a] = fminf(llo[a], rlo[a]);
        hi[a] = fmaxf(lhi[a], rhi[a]);
    }
    const uint32_t sl = c.x >= leaf_base ? 1u : subtree[c.x];
    const uint32_t sr = c.y >= leaf_base ? 1u : subtree[c.y];
    const uint32_t il = c.x >= leaf_base ? 0u : inner[c.x];
    const uint32_t ir = c.y >= leaf_base ? 0u : inner[c.y];
    const uint32_t size = sl + sr;
    box_store(&node_box[node], lo, hi);
    subtree[node] = size;
    inner[node] = size > MAX_LEAF ? 1u + il + ir : 0u;
}

// ---------------------------------------------------------------------------
// Two-phase fit: the same unions again, nearly all of them inside one
// workgroup with their state in LDS. A "treelet" is a subtree of at most
// TREELET leaves whose parent's subtree is larger; its leaves are
// consecutive sorted primitives (a Karras node's range is its subtree's
// leaves) and so are the indices of its internal nodes (a node is one end of
// its range), so one workgroup fits a treelet by the atomic bottom-up walk
// with boxes and arrival counts in LDS indexed from the treelet's first leaf.
// The nodes above the treelets ("top" nodes, subtrees of more than TREELET
// leaves: about 2n / TREELET) are fitted afterwards by one workgroup that
// keeps their topology and arrival counts in LDS (device memory past
// TOP_LDS of them) and climbs from the top nodes whose children are both
// treelet roots. Subtree sizes come from the ranges, so both lists are known
// before any box is.
// ---------------------------------------------------------------------------

__device__ inline uint32_t range_size(const uint2* __restrict__ ranges, uint32_t leaf_base, uint32_t node) {
    return node >= leaf_base ? 1u : ranges[node].y - ranges[node].x + 1u;
}

// One treelet per iteration of the block's loop (the grid strides over the
// treelet list). The treelet's topology is read once, coalesced, into LDS:
// for the node at index f + k (k < m) its parent's and children's slots
// (children tagged as leaves by LEAF_TAG); slot k of the leaf arrays is leaf
// f + k. Then each leaf's thread climbs with LDS reads only, arrival bits
// flipped by atomic XOR: the second arrival at a node sees its bit set.
constexpr uint32_t LEAF_TAG = 0x8000u;

__global__ __launch_bounds__(TREELET) void k_fit_treelets(const LucidaAabb* __restrict__ tri_bounds,
                                                          const uint32_t* __restrict__ order, uint32_t n,
                                                          const uint2* __restrict__ karras,
                                                          const uint2* __restrict__ ranges,
                                                          const uint32_t* __restrict__ parent,
                                                          const uint32_t* __restrict__ treelets,
                                                          const uint32_t* __restrict__ counts,
                                                          LucidaAabb* __restrict__ node_box,
                                                          uint32_t* __restrict__ subtree,
                                                          uint32_t* __restrict__ inner) {
    __shared__ float leaf_lo[TREELET][3], leaf_hi[TREELET][3];
    __shared__ float node_lo[TREELET][3], node_hi[TREELET][3];
    __shared__ uint32_t size_emit[TREELET];  // subtree size (low 16 bits), emitted internal count (high 16)
    __shared__ uint32_t kids[TREELET];       // child slots, low 16 bits left, high 16 right
    __shared__ uint32_t up[TREELET];         // parent slot of internal node f + k
    __shared__ uint32_t arrived[TREELET / 32];
    const uint32_t leaf_base = n - 1;
    const uint32_t count = counts[0];
    const uint32_t t = threadIdx.x;
    for (uint32_t j = blockIdx.x; j < count; j += gridDim.x) {
        const uint32_t root = treelets[j];
        uint32_t f, m;
        if (root >= leaf_base) {
            f = root - leaf_base;
            m = 1;
        } else {
            f = ranges[root].x;
            m = ranges[root].y - f + 1u;
        }
        if (t < TREELET / 32) arrived[t] = 0;
        float lo[3], hi[3];
        uint32_t leaf_up = 0;
        if (t < m) {
            const LucidaAabb b = tri_bounds[order[f + t]];
            for (int a = 0; a < 3; ++a) {
                lo[a] = b.lo[a];
                hi[a] = b.hi[a];
                leaf_lo[t][a] = lo[a];
                leaf_hi[t][a] = hi[a];
            }
            node_box[leaf_base + f + t] = b;
            leaf_up = parent[leaf_base + f + t] - f;
            const uint32_t i = f + t;  // an internal index inside the range (one of them may lie outside the treelet: never visited)
            if (i < leaf_base) {
                const uint2 c = karras[i];
                const uint32_t l = c.x >= leaf_base ? ((c.x - leaf_base - f) | LEAF_TAG) : c.x - f;
                const uint32_t r = c.y >= leaf_base ? ((c.y - leaf_base - f) | LEAF_TAG) : c.y - f;
                kids[t] = (l & 0xFFFFu) | (r << 16);
                up[t] = parent[i] - f;
            }
        }
        __syncthreads();
        if (t < m && root < leaf_base) {
            uint32_t slot = leaf_up;  // the node being arrived at
            for (;;) {
                // this child's LDS entries are written: release them to the sibling's thread
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                const uint32_t bit = 1u << (slot & 31u);
                if ((atomicXor(&arrived[slot >> 5], bit) & bit) == 0u) break;  // first arrival: the sibling goes on
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                const uint32_t k2 = kids[slot];
                const uint32_t ch[2] = {k2 & 0xFFFFu, k2 >> 16};
                float clo[2][3], chi[2][3];
                uint32_t csize[2], cemit[2];
                for (int q = 0; q < 2; ++q) {
                    const bool leaf = (ch[q] & LEAF_TAG) != 0u;
                    const uint32_t cs = ch[q] & ~LEAF_TAG;
                    for (int a = 0; a < 3; ++a) {
                        clo[q][a] = leaf ? leaf_lo[cs][a] : node_lo[cs][a];
                        chi[q][a] = leaf ? leaf_hi[cs][a] : node_hi[cs][a];
                    }
                    csize[q] = leaf ? 1u : size_emit[cs] & 0xFFFFu;
                    cemit[q] = leaf ? 0u : size_emit[cs] >> 16;
                }
                for (int a = 0; a < 3; ++a) {
                    lo[a] = fminf(clo[0][a], clo[1][a]);
                    hi[a] = fmaxf(chi[0][a], chi[1][a]);
                }
                const uint32_t size = csize[0] + csize[1];
                const uint32_t emit = size > MAX_LEAF ? 1u + cemit[0] + cemit[1] : 0u;
                for (int a = 0; a < 3; ++a) {
                    node_lo[slot][a] = lo[a];
                    node_hi[slot][a] = hi[a];
                }
                size_emit[slot] = size | (emit << 16);
                const uint32_t node = f + slot;
                box_store(&node_box[node], lo, hi);
                subtree[node] = size;
                inner[node] = emit;
                if (node == root) break;
                slot = up[slot];
            }
        }
        __syncthreads();  // LDS is reused by the next treelet
    }
}

// The top nodes, by one workgroup, with their state in LDS (device memory
// past TOP_LDS slots; gfx950 gives a workgroup up to 160 KB of LDS). Per
// top slot: a record (box, subtree size, emitted-internal count), the slots
// of its top children (NONE for a child that is a treelet root, final since
// k_fit_treelets), its parent's slot and an arrival count that starts at the
// number of treelet-root children (also kept in the second byte, which marks
// the nodes ready at the start); the count reaching 2 makes the node ready.
// A node with one treelet-root child holds that child's record until it is
// fitted; one with two is fitted while the state is read in. Unions take the
// left child first, as every other fit does. The climbs read and write LDS
// only, ordered by workgroup-scope release / acquire; the records go to the
// node arrays at the end.
struct TopRec {
    float lo[3], hi[3];
    uint32_t size, emit;
};

__device__ inline TopRec top_union(const TopRec& l, const TopRec& r) {
    TopRec u;
    for (int a = 0; a < 3; ++a) {
        u.lo[a] = fminf(l.lo[a], r.lo[a]);
        u.hi[a] = fmaxf(l.hi[a], r.hi[a]);
    }
    u.size = l.size + r.size;
    u.emit = u.size > MAX_LEAF ? 1u + l.emit + r.emit : 0u;
    return u;
}

__global__ __launch_bounds__(1024) void k_fit_top(uint32_t n, const uint2* __restrict__ karras,
                                                  const uint2* __restrict__ ranges,
                                                  const uint32_t* __restrict__ parent,
                                                  const uint32_t* __restrict__ top,
                                                  const uint32_t* __restrict__ slot_of,
                                                  const uint32_t* __restrict__ counts, uint32_t* __restrict__ spill,
                                                  LucidaAabb* node_box, uint32_t* subtree, uint32_t* inner) {
    __shared__ TopRec rec_s[TOP_LDS];
    __shared__ uint32_t kid0_s[TOP_LDS], kid1_s[TOP_LDS], up_s[TOP_LDS], arr_s[TOP_LDS];
    const uint32_t leaf_base = n - 1;
    const uint32_t T = counts[1];
    TopRec* rec_g = reinterpret_cast<TopRec*>(spill);
    uint32_t* kid0_g = spill + 8 * (size_t)T;
    uint32_t* kid1_g = kid0_g + T;
    uint32_t* up_g = kid1_g + T;
    uint32_t* arr_g = up_g + T;
    auto rec = [&](uint32_t s) -> TopRec& { return s < TOP_LDS ? rec_s[s] : rec_g[s]; };
    auto at = [](uint32_t* lds, uint32_t* g, uint32_t s) -> uint32_t* { return s < TOP_LDS ? lds + s : g + s; };
    // the record of a child that is a treelet root (or a leaf)
    auto final_rec = [&](uint32_t x) {
        TopRec r;
        box_load(&node_box[x], r.lo, r.hi);
        r.size = x >= leaf_base ? 1u : subtree[x];
        r.emit = x >= leaf_base ? 0u : inner[x];
        return r;
    };
    for (uint32_t s = threadIdx.x; s < T; s += blockDim.x) {
        const uint32_t node = top[s];
        const uint2 c = karras[node];
        const uint32_t k0 = range_size(ranges, leaf_base, c.x) > TREELET ? slot_of[c.x] : NONE;
        const uint32_t k1 = range_size(ranges, leaf_base, c.y) > TREELET ? slot_of[c.y] : NONE;
        if (k0 == NONE && k1 == NONE) rec(s) = top_union(final_rec(c.x), final_rec(c.y));
        else if (k0 == NONE) rec(s) = final_rec(c.x);
        else if (k1 == NONE) rec(s) = final_rec(c.y);
        const uint32_t ready = (k0 == NONE ? 1u : 0u) + (k1 == NONE ? 1u : 0u);
        *at(kid0_s, kid0_g, s) = k0;
        *at(kid1_s, kid1_g, s) = k1;
        *at(up_s, up_g, s) = node == 0 ? NONE : slot_of[parent[node]];
        *at(arr_s, arr_g, s) = ready | (ready << 8);  // arrivals (low byte) and the count at the start
    }
    __syncthreads();
    for (uint32_t s0 = threadIdx.x; s0 < T; s0 += blockDim.x) {
        if ((*at(arr_s, arr_g, s0) >> 8) != 2u) continue;  // not ready at the start: a top child climbs into it
        uint32_t s = s0;  // fitted
        for (;;) {
            const uint32_t ps = *at(up_s, up_g, s);
            if (ps == NONE) break;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            if ((atomicAdd(at(arr_s, arr_g, ps), 1u) & 0xFFu) != 1u) break;  // the sibling finishes the parent
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            s = ps;
            const uint32_t k0 = *at(kid0_s, kid0_g, s), k1 = *at(kid1_s, kid1_g, s);
            const TopRec own = rec(s);  // a treelet-root child's record, if it has one
            TopRec l = own, r = own;
            if (k0 != NONE) l = rec(k0);
            if (k1 != NONE) r = rec(k1);
            rec(s) = top_union(l, r);
        }
    }
    __syncthreads();
    for (uint32_t s = threadIdx.x; s < T; s += blockDim.x) {
        const uint32_t node = top[s];
        const TopRec& r = rec(s);
        box_store(&node_box[node], r.lo, r.hi);
        subtree[node] = r.size;
        inner[node] = r.emit;
    }
}

// ---------------------------------------------------------------------------
// Relayout (lbvh.rs:315-406) on the device. The reference walks the tree
// depth first, left child first, collapsing subtrees of <= 4 primitives into
// leaves and appending each emitted internal node's two children as an
// adjacent pair. So an emitted internal node X with pre-order rank R(X) among
// emitted internal nodes has its children at 1 + 2 R(X), and
// R(child) = R(X) + 1 (+ the emitted internal count of the left sibling's
// subtree, for a right child). Leaves list primitives in sorted order, so
// prim_indices is the sorted order itself and a leaf starts at its range's
// first sorted position.
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_rank(uint32_t n, const uint2* __restrict__ karras,
                                              const uint32_t* __restrict__ parent,
                                              const uint32_t* __restrict__ subtree,
                                              const uint32_t* __restrict__ inner, uint32_t* __restrict__ rank) {
    const uint32_t x = blockIdx.x * 256 + threadIdx.x;
    if (x >= n - 1) return;
    if (subtree[x] <= MAX_LEAF) return;  // not an emitted internal node
    uint32_t r = 0, cur = x;
    while (cur != 0) {
        uint32_t p = parent[cur];
        uint2 c = karras[p];
        r += 1;
        if (cur == c.y) r += c.x >= n - 1 ? 0u : inner[c.x];
        cur = p;
    }
    rank[x] = r;
}

__global__ __launch_bounds__(256) void k_emit(uint32_t n, const uint2* __restrict__ karras,
                                              const uint2* __restrict__ ranges, const uint32_t* __restrict__ parent,
                                              const uint32_t* __restrict__ subtree,
                                              const uint32_t* __restrict__ rank,
                                              const LucidaAabb* __restrict__ node_box, GpuBvhNode* __restrict__ out) {
    const uint32_t id = blockIdx.x * 256 + threadIdx.x;
    const uint32_t leaf_base = n - 1;
    if (id >= 2 * n - 1) return;
    uint32_t slot;
    if (id == 0) {
        slot = 0;
    } else {
        uint32_t p = parent[id];
        if (subtree[p] <= MAX_LEAF) return;  // inside a collapsed leaf
        slot = 1 + 2 * rank[p] + (karras[p].y == id ? 1u : 0u);
    }
    const bool is_leaf = id >= leaf_base;
    const uint32_t size = is_leaf ? 1u : subtree[id];
    GpuBvhNode nd;
    const LucidaAabb b = node_box[id];
    nd.bounds_min[0] = b.lo[0]; nd.bounds_min[1] = b.lo[1]; nd.bounds_min[2] = b.lo[2];
    nd.bounds_max[0] = b.hi[0]; nd.bounds_max[1] = b.hi[1]; nd.bounds_max[2] = b.hi[2];
    if (is_leaf || size <= MAX_LEAF) {
        nd.left_first = is_leaf ? id - leaf_base : ranges[id].x;
        nd.count = size;
    } else {
        nd.left_first = 1 + 2 * rank[id];
        nd.count = 0;
    }
    out[slot] = nd;
}

__global__ __launch_bounds__(256) void k_gather_tris(const GpuTriangle* __restrict__ tris,
                                                     const uint32_t* __restrict__ order, uint32_t n,
                                                     GpuTriangle* __restrict__ out) {
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k < n) out[k] = tris[order[k]];
}

__global__ void k_single_leaf(const LucidaAabb* __restrict__ tri_bounds, GpuBvhNode* out, uint32_t* prims) {
    LucidaAabb b = tri_bounds[0];
    GpuBvhNode nd;
    for (int a = 0; a < 3; ++a) {
        nd.bounds_min[a] = b.lo[a];
        nd.bounds_max[a] = b.hi[a];
    }
    nd.left_first = 0;
    nd.count = 1;
    out[0] = nd;
    if (prims) prims[0] = 0;
}

inline uint32_t blocks_for(size_t n, int block) { return (uint32_t)((n + block - 1) / block); }

}  // namespace

// Exclusive prefix sum of `count` device words in place; `partials` holds
// ceil(count / 2048) words of scratch.
lucida_status scan_u32_device(lucida_ctx* ctx, uint32_t* data, uint32_t count, uint32_t* partials) {
    if (count == 0) return LUCIDA_OK;
    const uint32_t segs = blocks_for(count, SCAN_SEG);
    hipLaunchKernelGGL(k_scan_reduce, dim3(segs), dim3(SCAN_BLOCK), 0, ctx->stream, data, count, partials);
    hipLaunchKernelGGL(k_scan_down, dim3(segs), dim3(SCAN_BLOCK), 0, ctx->stream, data, count, partials);
    LUCIDA_HIP(hipGetLastError(), LUCIDA_SHADER);
    return LUCIDA_OK;
}

// Words of histogram + scan scratch a sort of n keys uses.
static size_t sort_scratch_words(size_t n) {
    const size_t hist_words = blocks_for(n, SORT_TILE) * (size_t)RADIX;
    return hist_words + blocks_for(hist_words, SCAN_SEG);
}

// Radix-sorts device arrays keys/vals in place over bits [0, bits), using
// tmp arrays of the same length and `work` (sort_scratch_words(n) words,
// every word written before it is read, so it needs no clearing).
static lucida_status sort_device_with(lucida_ctx* ctx, uint32_t* keys, uint32_t* vals, uint32_t* tk, uint32_t* tv,
                                      size_t n, uint32_t bits, uint32_t* work) {
    const uint32_t tiles = blocks_for(n, SORT_TILE);
    const uint32_t hist_words = tiles * RADIX;
    const uint32_t segs = blocks_for(hist_words, SCAN_SEG);
    uint32_t* hist = work;
    uint32_t* partials = hist + hist_words;
    uint32_t *src_k = keys, *src_v = vals, *dst_k = tk, *dst_v = tv;
    const uint32_t passes = (bits + 7) / 8;
    for (uint32_t pass = 0; pass < passes; ++pass) {
        const uint32_t shift = pass * 8;
        hipLaunchKernelGGL(k_radix_hist, dim3(tiles), dim3(SORT_BLOCK), 0, ctx->stream, src_k, (uint32_t)n,
                           nullptr, shift, hist, tiles);
        hipLaunchKernelGGL(k_scan_reduce, dim3(segs), dim3(SCAN_BLOCK), 0, ctx->stream, hist, hist_words, partials);
            hipLaunchKernelGGL(k_scan_down, dim3(segs), dim3(SCAN_BLOCK), 0, ctx->stream, hist, hist_words, partials);
        hipLaunchKernelGGL(k_radix_scatter, dim3(tiles), dim3(SORT_BLOCK), 0, ctx->stream, src_k, src_v,
                           (uint32_t)n, nullptr, shift, hist, tiles, dst_k, dst_v);
        std::swap(src_k, dst_k);
        std::swap(src_v, dst_v);
    }
    if (src_k != keys) {  // odd number of passes: copy back
        LUCIDA_HIP(hipMemcpyAsync(keys, src_k, n * 4, hipMemcpyDeviceToDevice, ctx->stream), LUCIDA_VALIDATION);
        LUCIDA_HIP(hipMemcpyAsync(vals, src_v, n * 4, hipMemcpyDeviceToDevice, ctx->stream), LUCIDA_VALIDATION);
    }
    LUCIDA_HIP(hipGetLastError(), LUCIDA_SHADER);
    return LUCIDA_OK;
}

// The same sort with its scratch in the context's (growing) histogram buffer.
static lucida_status sort_device(lucida_ctx* ctx, uint32_t* keys, uint32_t* vals, uint32_t* tk, uint32_t* tv,
                                 size_t n, uint32_t bits) {
    if (n == 0) return LUCIDA_OK;
    if (n > 0xFFFFFFFFull / 2) return set_error(LUCIDA_INVALID_ARG, "sort: more than 2^31 keys");
    LUCIDA_HIP(ctx->lb_hist.ensure(sort_scratch_words(n) * 4), LUCIDA_OOM);
    return sort_device_with(ctx, keys, vals, tk, tv, n, bits, ctx->lb_hist.as<uint32_t>());
}

// ---------------------------------------------------------------------------
// Sort scratch that is never allocated: a file-scope device array, one copy
// per GPU, for callers that cannot allocate (code under graph capture, tight
// serving loops). 8 MiB + 4 KiB covers LUCIDA_SORT_NOALLOC_MAX_KEYS keys.
// Host code never passes `scratch` itself (its host-side name is not a
// device address): each device's address comes from hipGetSymbolAddress,
// made once per device with that device current, and is kept in
// g_scratch_ptr.
// ---------------------------------------------------------------------------

constexpr size_t SORT_NOALLOC_WORDS =
    (LUCIDA_SORT_NOALLOC_MAX_KEYS / SORT_TILE) * RADIX + (LUCIDA_SORT_NOALLOC_MAX_KEYS / SORT_TILE) * RADIX / SCAN_SEG;
static_assert(LUCIDA_SORT_NOALLOC_MAX_KEYS % SORT_TILE == 0, "max keys is a whole number of tiles");
static_assert((LUCIDA_SORT_NOALLOC_MAX_KEYS / SORT_TILE) * RADIX % SCAN_SEG == 0, "whole scan segments");

static __device__ __align__(256) uint32_t scratch[SORT_NOALLOC_WORDS];

constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mutex;
static uint32_t* g_scratch_ptr[MAX_DEVICES] = {};

// Resolves this device's copy of `scratch` (the current device must be
// ctx->device). Called from lucida_ctx_create; later contexts on the same
// device reuse the address.
lucida_status sort_scratch_init(lucida_ctx* ctx) {
    if (ctx->device < 0 || ctx->device >= MAX_DEVICES)
        return set_error(LUCIDA_NO_DEVICE, "sort scratch: device index beyond the supported 64");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[ctx->device]) {
        uint32_t* ptr = nullptr;
        LUCIDA_HIP(hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)), LUCIDA_NO_DEVICE);
        g_scratch_ptr[ctx->device] = ptr;
    }
    return LUCIDA_OK;
}

size_t sort_counted_scratch_words(size_t capacity) {
    const size_t hist_words = blocks_for(capacity, SORT_TILE) * (size_t)RADIX;
    return hist_words + blocks_for(hist_words, SCAN_SEG);
}

// Stable sort of (keys, vals)[0, *count) by key bits [0, bits) on `stream`,
// bits a multiple of 8 (one pass per 8 bits). The count stays on the device:
// the grid covers `capacity`, and tiles past the count write empty
// histograms and return. Passes alternate between (keys, vals) and (tk, tv):
// after an even number the result is in keys/vals, after an odd number in
// tk/tv; *sorted_vals (when given) receives the array holding the sorted
// values. Every launch is a kernel (no memset or copy), so the sort may be
// captured in a graph.
lucida_status sort_pairs_counted(hipStream_t stream, uint32_t* keys, uint32_t* vals, uint32_t* tk, uint32_t* tv,
                                 size_t capacity, const uint32_t* count, uint32_t bits, uint32_t* work,
                                 uint32_t** sorted_vals) {
    if (sorted_vals) *sorted_vals = vals;
    if (capacity == 0) return LUCIDA_OK;
    if (bits == 0 || bits % 8 != 0 || bits > 32) return set_error(LUCIDA_INVALID_ARG, "sort: bits must be 8, 16, 24 or 32");
    if (capacity > 0xFFFFFFFFull / 2) return set_error(LUCIDA_INVALID_ARG, "sort: more than 2^31 keys");
    const uint32_t tiles = blocks_for(capacity, SORT_TILE);
    const uint32_t hist_words = tiles * RADIX;
    const uint32_t segs = blocks_for(hist_words, SCAN_SEG);
    uint32_t* hist = work;
    uint32_t* partials = work + hist_words;
    uint32_t *src_k = keys, *src_v = vals, *dst_k = tk, *dst_v = tv;
    for (uint32_t shift = 0; shift < bits; shift += 8) {
        hipLaunchKernelGGL(k_radix_hist, dim3(tiles), dim3(SORT_BLOCK), 0, stream, src_k, (uint32_t)capacity, count,
                           shift, hist, tiles);
        hipLaunchKernelGGL(k_scan_reduce, dim3(segs), dim3(SCAN_BLOCK), 0, stream, hist, hist_words, partials);
        hipLaunchKernelGGL(k_scan_down, dim3(segs), dim3(SCAN_BLOCK), 0, stream, hist, hist_words, partials);
        // the last pass needs only the values
        hipLaunchKernelGGL(k_radix_scatter, dim3(tiles), dim3(SORT_BLOCK), 0, stream, src_k, src_v,
                           (uint32_t)capacity, count, shift, hist, tiles, shift + 8 < bits ? dst_k : nullptr, dst_v);
        std::swap(src_k, dst_k);
        std::swap(src_v, dst_v);
    }
    if (sorted_vals) *sorted_vals = src_v;
    LUCIDA_HIP(hipGetLastError(), LUCIDA_SHADER);
    return LUCIDA_OK;
}

// Level-synchronous fit (see k_depth / k_bucket / k_fit_level): one readback
// of the per-level histogram, then one launch per level, deepest first.
static lucida_status fit_levels(lucida_ctx* ctx, size_t n) {
    hipStream_t s = ctx->stream;
    const size_t internal = n - 1;
    LUCIDA_HIP(ctx->lb_level.ensure(internal * 4 + (3 * MAX_LEVELS + 8) * 4), LUCIDA_OOM);
    uint32_t* list = ctx->lb_level.as<uint32_t>();
    uint32_t* level_count = list + internal;
    uint32_t* cursor = level_count + MAX_LEVELS;
    uint32_t* overflow = cursor + MAX_LEVELS;
    uint32_t* depth = ctx->lb_flags.as<uint32_t>();
    uint32_t* subtree = ctx->lb_subtree.as<uint32_t>();
    LUCIDA_HIP(hipMemsetAsync(level_count, 0, (2 * MAX_LEVELS + 8) * 4, s), LUCIDA_VALIDATION);
    hipLaunchKernelGGL(k_leaf_boxes, dim3(blocks_for(n, 256)), dim3(256), 0, s, ctx->lb_bounds.as<LucidaAabb>(),
                       ctx->lb_vals.as<uint32_t>(), (uint32_t)n, ctx->lb_nodebox.as<LucidaAabb>());
    hipLaunchKernelGGL(k_depth, dim3(blocks_for(internal, 256)), dim3(256), 0, s, (uint32_t)n,
                       ctx->lb_parent.as<uint32_t>(), depth, level_count, overflow);
    uint32_t host[MAX_LEVELS + 1];
    LUCIDA_HIP(hipMemcpyAsync(host, level_count, MAX_LEVELS * 4, hipMemcpyDeviceToHost, s), LUCIDA_READBACK);
    LUCIDA_HIP(hipMemcpyAsync(host + MAX_LEVELS, overflow, 4, hipMemcpyDeviceToHost, s), LUCIDA_READBACK);
    LUCIDA_HIP(hipStreamSynchronize(s), LUCIDA_VALIDATION);
    if (host[MAX_LEVELS]) return set_error(LUCIDA_VALIDATION, "LBVH deeper than 128 levels");
    uint32_t off[MAX_LEVELS], acc = 0;
    int deepest = 0;
    for (int l = 0; l < MAX_LEVELS; ++l) {
        off[l] = acc;
        acc += host[l];
        if (host[l]) deepest = l;
    }
    // the device forms the same offsets as its bucket cursors, in stream order
    hipLaunchKernelGGL(k_level_offsets, dim3(1), dim3(64), 0, s, level_count, cursor);
    hipLaunchKernelGGL(k_bucket, dim3(blocks_for(internal, 256)), dim3(256), 0, s, (uint32_t)n, depth, cursor, list);
    for (int l = deepest; l >= 0; --l) {
        if (!host[l]) continue;
        hipLaunchKernelGGL(k_fit_level, dim3(blocks_for(host[l], 256)), dim3(256), 0, s, list + off[l], host[l],
                           (uint32_t)n, ctx->lb_karras.as<uint2>(), ctx->lb_nodebox.as<LucidaAabb>(), subtree,
                           subtree + internal);
    }
    LUCIDA_HIP(hipGetLastError(), LUCIDA_SHADER);
    return LUCIDA_OK;
}

// The lists of the two-phase fit, in lb_level (k_karras fills them).
static FitLists fit_lists(lucida_ctx* ctx, size_t n) {
    const size_t internal = n - 1;
    FitLists l;
    l.counts = ctx->lb_level.as<uint32_t>();
    l.treelets = l.counts + 8;
    l.top = l.treelets + internal + 1;
    l.slot_of = l.top + internal;
    return l;
}

// Two-phase fit: k_fit_treelets fits the treelets k_karras listed,
// k_fit_top the nodes above them.
static lucida_status fit_treelets(lucida_ctx* ctx, size_t n) {
    hipStream_t s = ctx->stream;
    const size_t internal = n - 1;
    const FitLists l = fit_lists(ctx, n);
    uint32_t* spill = l.slot_of + internal;  // 12 words per top slot
    uint32_t* subtree = ctx->lb_subtree.as<uint32_t>();
    // at most 2n / TREELET + 1 treelets in a balanced tree; blocks stride over the rest
    const uint32_t grid = std::min<uint32_t>(blocks_for(2 * n, TREELET) + 1, 4096);
    hipLaunchKernelGGL(k_fit_treelets, dim3(grid), dim3(TREELET), 0, s, ctx->lb_bounds.as<LucidaAabb>(),
                       ctx->lb_vals.as<uint32_t>(), (uint32_t)n, ctx->lb_karras.as<uint2>(),
                       ctx->lb_rank.as<uint2>(), ctx->lb_parent.as<uint32_t>(), l.treelets, l.counts,
                       ctx->lb_nodebox.as<LucidaAabb>(), subtree, subtree + internal);
    hipLaunchKernelGGL(k_fit_top, dim3(1), dim3(1024), 0, s, (uint32_t)n, ctx->lb_karras.as<uint2>(),
                       ctx->lb_rank.as<uint2>(), ctx->lb_parent.as<uint32_t>(), l.top, l.slot_of, l.counts, spill,
                       ctx->lb_nodebox.as<LucidaAabb>(), subtree, subtree + internal);
    LUCIDA_HIP(hipGetLastError(), LUCIDA_SHADER);
    return LUCIDA_OK;
}

// The device stages of the build, up to and including the fit. Results stay in
// the context's scratch (lb_codes = sorted codes, lb_vals = order, ...).
static lucida_status build_stages(lucida_ctx* ctx, const GpuTriangle* tris, size_t n, const float* positions,
                                  float* stage_us, bool keep_unsorted_codes) {
    hipStream_t s = ctx->stream;
    const uint32_t nb = blocks_for(n, PREP_BLOCK);
    const size_t nodes = 2 * n - 1;
    const size_t internal = n > 1 ? n - 1 : 1;
    LUCIDA_HIP(ctx->lb_bounds.ensure(n * sizeof(LucidaAabb)), LUCIDA_OOM);
    LUCIDA_HIP(ctx->lb_partial.ensure((size_t)nb * 6 * 4), LUCIDA_OOM);
    LUCIDA_HIP(ctx->lb_misc.ensure(256), LUCIDA_OOM);
    LUCIDA_HIP(ctx->lb_codes.ensure(n * 4 * (keep_unsorted_codes ? 2 : 1)), LUCIDA_OOM);
    LUCIDA_HIP(ctx->lb_vals.ensure(n * 4), LUCIDA_OOM);
    LUCIDA_HIP(ctx->lb_tmpk.ensure(n * 4), LUCIDA_OOM);
    LUCIDA_HIP(ctx->lb_tmpv.ensure(n * 4), LUCIDA_OOM);
    LUCIDA_HIP(ctx->lb_karras.ensure(internal * 8), LUCIDA_OOM);
    LUCIDA_HIP(ctx->lb_rank.ensure(internal * 8), LUCIDA_OOM);  // ranges, then ranks
    LUCIDA_HIP(ctx->lb_parent.ensure(nodes * 4), LUCIDA_OOM);
    LUCIDA_HIP(ctx->lb_nodebox.ensure(nodes * sizeof(LucidaAabb)), LUCIDA_OOM);
    LUCIDA_HIP(ctx->lb_subtree.ensure(internal * 4 * 3), LUCIDA_OOM);  // subtree, inner, rank
    LUCIDA_HIP(ctx->lb_flags.ensure(internal * 4), LUCIDA_OOM);

    uint32_t* misc = ctx->lb_misc.as<uint32_t>();  // [8..16) scene box
    LucidaAabb* scene_box = reinterpret_cast<LucidaAabb*>(misc + 8);
    StageEvents<5> ev;
    const bool timed = stage_us != nullptr;
    if (timed) LUCIDA_HIP(ev.create(), LUCIDA_VALIDATION);
    if (timed) LUCIDA_HIP(hipEventRecord(ev[0], s), LUCIDA_VALIDATION);

    hipLaunchKernelGGL(k_prepare, dim3(nb), dim3(PREP_BLOCK), 0, s, tris, reinterpret_cast<const float4*>(positions),
                       (uint32_t)n, ctx->lb_bounds.as<LucidaAabb>(), ctx->lb_partial.as<float>());
    // the two-phase fit's lists (k_karras fills them)
    const bool lists_on = n >= 2 && ctx->lbvh_fit == 0;
    if (lists_on) LUCIDA_HIP(ctx->lb_level.ensure((15 * (n - 1) + 16) * 4), LUCIDA_OOM);
    const FitLists lists = lists_on ? fit_lists(ctx, n) : FitLists{nullptr, nullptr, nullptr, nullptr};
    hipLaunchKernelGGL(k_fold_bounds, dim3(1), dim3(PREP_BLOCK), 0, s, ctx->lb_partial.as<float>(), (uint32_t)nb,
                       scene_box, n >= 2 ? ctx->lb_parent.as<uint32_t>() : nullptr, lists.counts);
    hipLaunchKernelGGL(k_morton, dim3(nb), dim3(PREP_BLOCK), 0, s, ctx->lb_bounds.as<LucidaAabb>(), (uint32_t)n,
                       scene_box, ctx->lb_codes.as<uint32_t>(), ctx->lb_vals.as<uint32_t>());
    if (keep_unsorted_codes)
        LUCIDA_HIP(hipMemcpyAsync(ctx->lb_codes.as<uint32_t>() + n, ctx->lb_codes.as<uint32_t>(), n * 4,
                                  hipMemcpyDeviceToDevice, s),
                   LUCIDA_VALIDATION);
    if (timed) LUCIDA_HIP(hipEventRecord(ev[1], s), LUCIDA_VALIDATION);
    lucida_status st = sort_device(ctx, ctx->lb_codes.as<uint32_t>(), ctx->lb_vals.as<uint32_t>(),
                                   ctx->lb_tmpk.as<uint32_t>(), ctx->lb_tmpv.as<uint32_t>(), n, 30);
    if (st != LUCIDA_OK) return st;
    if (timed) LUCIDA_HIP(hipEventRecord(ev[2], s), LUCIDA_VALIDATION);
    if (n >= 2) {
        // the root's parent and the list counters were cleared by k_fold_bounds
        hipLaunchKernelGGL(k_karras, dim3(blocks_for(n - 1, 256)), dim3(256), 0, s, ctx->lb_codes.as<uint32_t>(),
                           (uint32_t)n, ctx->lb_karras.as<uint2>(), ctx->lb_rank.as<uint2>(),
                           ctx->lb_parent.as<uint32_t>(), lists);
    }
    if (timed) LUCIDA_HIP(hipEventRecord(ev[3], s), LUCIDA_VALIDATION);
    if (n >= 2) {
        uint32_t* subtree = ctx->lb_subtree.as<uint32_t>();
        if (ctx->lbvh_fit == 0) {
            lucida_status fs = fit_treelets(ctx, n);
            if (fs != LUCIDA_OK) return fs;
        } else if (ctx->lbvh_fit == 1) {
            LUCIDA_HIP(hipMemsetAsync(ctx->lb_flags.ptr, 0, (n - 1) * 4, s), LUCIDA_VALIDATION);
            hipLaunchKernelGGL(k_fit, dim3(blocks_for(n, 256)), dim3(256), 0, s, ctx->lb_bounds.as<LucidaAabb>(),
                               ctx->lb_vals.as<uint32_t>(), (uint32_t)n, ctx->lb_karras.as<uint2>(),
                               ctx->lb_parent.as<uint32_t>(), ctx->lb_nodebox.as<LucidaAabb>(), subtree,
                               subtree + internal, ctx->lb_flags.as<uint32_t>());
        } else {
            lucida_status fs = fit_levels(ctx, n);
            if (fs != LUCIDA_OK) return fs;
        }
    }
    if (timed) LUCIDA_HIP(hipEventRecord(ev[4], s), LUCIDA_VALIDATION);
    LUCIDA_HIP(hipGetLastError(), LUCIDA_SHADER);
    if (timed) {
        LUCIDA_HIP(hipEventSynchronize(ev[4]), LUCIDA_VALIDATION);
        for (int k = 0; k < 4; ++k) LUCIDA_HIP(hipEventElapsedTime(&stage_us[k], ev[k], ev[k + 1]), LUCIDA_VALIDATION);
        for (int k = 0; k < 4; ++k) stage_us[k] *= 1000.0f;
    }
    return LUCIDA_OK;
}

// The build up to the fitted tree, for passes that lay it out themselves:
// the Karras children (lb_karras), parents (lb_parent), boxes of every node
// (lb_nodebox), subtree sizes (lb_subtree) and the sorted order (lb_vals).
lucida_status lbvh_fit_device(lucida_ctx* ctx, const GpuTriangle* tris, size_t n, const float* positions) {
    return build_stages(ctx, tris, n, positions, nullptr, false);
}

// Relayout into GpuBvhNode on the device. The node count is 1 + 2 x the
// root's emitted-internal count, read back by the caller (`inner` word 0)
// once everything is queued.
static lucida_status relayout_device(lucida_ctx* ctx, size_t n, GpuBvhNode* nodes_out, uint32_t* prims_out) {
    hipStream_t s = ctx->stream;
    const size_t internal = n - 1;
    uint32_t* subtree = ctx->lb_subtree.as<uint32_t>();
    uint32_t* inner = subtree + internal;
    uint32_t* rank = subtree + 2 * internal;
    hipLaunchKernelGGL(k_rank, dim3(blocks_for(internal, 256)), dim3(256), 0, s, (uint32_t)n,
                       ctx->lb_karras.as<uint2>(), ctx->lb_parent.as<uint32_t>(), subtree, inner, rank);
    hipLaunchKernelGGL(k_emit, dim3(blocks_for(2 * n - 1, 256)), dim3(256), 0, s, (uint32_t)n,
                       ctx->lb_karras.as<uint2>(), ctx->lb_rank.as<uint2>(), ctx->lb_parent.as<uint32_t>(), subtree,
                       rank, ctx->lb_nodebox.as<LucidaAabb>(), nodes_out);
    if (prims_out)
        LUCIDA_HIP(hipMemcpyAsync(prims_out, ctx->lb_vals.ptr, n * 4, hipMemcpyDeviceToDevice, s), LUCIDA_VALIDATION);
    return LUCIDA_OK;
}

lucida_status lbvh_build_device(lucida_ctx* ctx, const GpuTriangle* tris, size_t n, const float* positions,
                                GpuBvhNode* nodes, size_t* num_nodes, uint32_t* prim_indices,
                                GpuTriangle* tris_sorted, float* stage_us) {
    if (n == 0) {
        *num_nodes = 0;
        return LUCIDA_OK;
    }
    if (n > (1u << 30)) return set_error(LUCIDA_INVALID_ARG, "LBVH: more than 2^30 triangles");
    float us[4] = {0, 0, 0, 0};
    lucida_status st = build_stages(ctx, tris, n, positions, stage_us ? us : nullptr, false);
    if (st != LUCIDA_OK) return st;
    StageEvents<2> ev;
    hipEvent_t& e0 = ev[0];
    hipEvent_t& e1 = ev[1];
    if (stage_us) {
        LUCIDA_HIP(ev.create(), LUCIDA_VALIDATION);
        LUCIDA_HIP(hipEventRecord(e0, ctx->stream), LUCIDA_VALIDATION);
    }
    if (n == 1) {
        hipLaunchKernelGGL(k_single_leaf, dim3(1), dim3(1), 0, ctx->stream, ctx->lb_bounds.as<LucidaAabb>(), nodes,
                           prim_indices);
        *num_nodes = 1;
    } else {
        st = relayout_device(ctx, n, nodes, prim_indices);
        if (st != LUCIDA_OK) return st;
    }
    if (tris_sorted)
        hipLaunchKernelGGL(k_gather_tris, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, tris,
                           ctx->lb_vals.as<uint32_t>(), (uint32_t)n, tris_sorted);
    LUCIDA_HIP(hipGetLastError(), LUCIDA_SHADER);
    if (stage_us) LUCIDA_HIP(hipEventRecord(e1, ctx->stream), LUCIDA_VALIDATION);
    if (n > 1) {
        // the root's emitted-internal count (0 when the root is a collapsed leaf)
        uint32_t root_inner = 0;
        LUCIDA_HIP(hipMemcpyAsync(&root_inner, ctx->lb_subtree.as<uint32_t>() + (n - 1), 4, hipMemcpyDeviceToHost,
                                  ctx->stream),
                   LUCIDA_READBACK);
        LUCIDA_HIP(hipStreamSynchronize(ctx->stream), LUCIDA_VALIDATION);
        *num_nodes = 1 + 2 * (size_t)root_inner;
    }
    if (stage_us) {
        LUCIDA_HIP(hipEventSynchronize(e1), LUCIDA_VALIDATION);
        float ms = 0;
        LUCIDA_HIP(hipEventElapsedTime(&ms, e0, e1), LUCIDA_VALIDATION);
        for (int k = 0; k < 4; ++k) stage_us[k] = us[k];
        stage_us[4] = ms * 1000.0f;
    }
    return LUCIDA_OK;
}

}  // namespace lucida

using namespace lucida;

extern "C" {

lucida_status lucida_ctx_set_lbvh_fit(lucida_ctx* ctx, int mode) {
    if (!ctx || mode < 0 || mode > 2)
        return set_error(LUCIDA_INVALID_ARG, "lucida_ctx_set_lbvh_fit: mode is 0, 1 or 2");
    ctx->lbvh_fit = mode;
    return LUCIDA_OK;
}

lucida_status lucida_sort_u32_device(lucida_ctx* ctx, uint32_t* keys, uint32_t* vals, uint32_t* tk, uint32_t* tv,
                                     size_t n) {
    if (!ctx || (n && (!keys || !vals || !tk || !tv)))
        return set_error(LUCIDA_INVALID_ARG, "lucida_sort_u32_device: null argument");
    clear_error();
    return sort_device(ctx, keys, vals, tk, tv, n, 32);
}

lucida_status lucida_sort_u32_device_noalloc(lucida_ctx* ctx, uint32_t* keys, uint32_t* vals, uint32_t* tk,
                                             uint32_t* tv, size_t n) {
    if (!ctx || (n && (!keys || !vals || !tk || !tv)))
        return set_error(LUCIDA_INVALID_ARG, "lucida_sort_u32_device_noalloc: null argument");
    if (n > LUCIDA_SORT_NOALLOC_MAX_KEYS)
        return set_error(LUCIDA_INVALID_ARG, "lucida_sort_u32_device_noalloc: more than LUCIDA_SORT_NOALLOC_MAX_KEYS "
                                             "(2^24) keys; use lucida_sort_u32_device");
    clear_error();
    if (n == 0) return LUCIDA_OK;
    uint32_t* ptr = ctx->device >= 0 && ctx->device < MAX_DEVICES ? g_scratch_ptr[ctx->device] : nullptr;
    if (!ptr || sort_scratch_words(n) > SORT_NOALLOC_WORDS)
        return set_error(LUCIDA_VALIDATION, "lucida_sort_u32_device_noalloc: the device's sort scratch is not set up");
    // 32 key bits: four passes, so the result lands back in keys/vals and no
    // copy is queued.
    return sort_device_with(ctx, keys, vals, tk, tv, n, 32, ptr);
}

lucida_status lucida_sort_u32(lucida_ctx* ctx, const uint32_t* keys, const uint32_t* vals, size_t n,
                              uint32_t* keys_out, uint32_t* vals_out) {
    if (!ctx || (n && (!keys || !vals || !keys_out || !vals_out)))
        return set_error(LUCIDA_INVALID_ARG, "lucida_sort_u32: null argument");
    clear_error();
    if (n == 0) return LUCIDA_OK;
    DevBuf buf;
    LUCIDA_HIP(buf.ensure(n * 16), LUCIDA_OOM);
    uint32_t* d = buf.as<uint32_t>();
    lucida_status st = LUCIDA_OK;
    if (hipMemcpyAsync(d, keys, n * 4, hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
        hipMemcpyAsync(d + n, vals, n * 4, hipMemcpyHostToDevice, ctx->stream) != hipSuccess) {
        buf.release();
        return set_error(LUCIDA_VALIDATION, "lucida_sort_u32: upload failed");
    }
    st = sort_device(ctx, d, d + n, d + 2 * n, d + 3 * n, n, 32);
    if (st == LUCIDA_OK) {
        if (hipMemcpyAsync(keys_out, d, n * 4, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipMemcpyAsync(vals_out, d + n, n * 4, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess)
            st = set_error(LUCIDA_READBACK, "lucida_sort_u32: readback failed");
    }
    buf.release();
    return st;
}

lucida_status lucida_lbvh_build_parts(lucida_ctx* ctx, const GpuTriangle* tris, size_t n, const float* positions,
                                      size_t num_positions, lucida_lbvh_parts* out) {
    if (!ctx || !tris || !positions || !out || n == 0)
        return set_error(LUCIDA_INVALID_ARG, "lucida_lbvh_build_parts: empty geometry has no tree");
    clear_error();
    for (size_t i = 0; i < n; ++i)
        if (tris[i].i0 >= num_positions || tris[i].i1 >= num_positions || tris[i].i2 >= num_positions)
            return set_error(LUCIDA_INVALID_ARG, "lucida_lbvh_build_parts: triangle index out of range");
    LUCIDA_HIP(ctx->lb_tris.ensure(n * sizeof(GpuTriangle)), LUCIDA_OOM);
    LUCIDA_HIP(ctx->lb_pos.ensure(num_positions * 16), LUCIDA_OOM);
    hipStream_t s = ctx->stream;
    LUCIDA_HIP(hipMemcpyAsync(ctx->lb_tris.ptr, tris, n * sizeof(GpuTriangle), hipMemcpyHostToDevice, s),
               LUCIDA_VALIDATION);
    LUCIDA_HIP(hipMemcpyAsync(ctx->lb_pos.ptr, positions, num_positions * 16, hipMemcpyHostToDevice, s),
               LUCIDA_VALIDATION);
    lucida_status st =
        build_stages(ctx, ctx->lb_tris.as<GpuTriangle>(), n, ctx->lb_pos.as<float>(), nullptr, true);
    if (st != LUCIDA_OK) return st;
    const size_t internal = n - 1;
    uint32_t* misc = ctx->lb_misc.as<uint32_t>();
    if (out->codes)
        LUCIDA_HIP(hipMemcpyAsync(out->codes, ctx->lb_codes.as<uint32_t>() + n, n * 4, hipMemcpyDeviceToHost, s),
                   LUCIDA_READBACK);
    if (out->order) LUCIDA_HIP(hipMemcpyAsync(out->order, ctx->lb_vals.ptr, n * 4, hipMemcpyDeviceToHost, s), LUCIDA_READBACK);
    if (out->scene_bounds)
        LUCIDA_HIP(hipMemcpyAsync(out->scene_bounds, misc + 8, 32, hipMemcpyDeviceToHost, s), LUCIDA_READBACK);
    if (n >= 2) {
        if (out->karras)
            LUCIDA_HIP(hipMemcpyAsync(out->karras, ctx->lb_karras.ptr, internal * 8, hipMemcpyDeviceToHost, s),
                       LUCIDA_READBACK);
        if (out->parent)
            LUCIDA_HIP(hipMemcpyAsync(out->parent, ctx->lb_parent.ptr, (2 * n - 1) * 4, hipMemcpyDeviceToHost, s),
                       LUCIDA_READBACK);
        if (out->node_bounds)
            LUCIDA_HIP(hipMemcpyAsync(out->node_bounds, ctx->lb_nodebox.ptr, (2 * n - 1) * sizeof(LucidaAabb),
                                      hipMemcpyDeviceToHost, s),
                       LUCIDA_READBACK);
        if (out->subtree_size)
            LUCIDA_HIP(hipMemcpyAsync(out->subtree_size, ctx->lb_subtree.ptr, internal * 4, hipMemcpyDeviceToHost, s),
                       LUCIDA_READBACK);
    } else {
        if (out->parent) out->parent[0] = 0xFFFFFFFFu;
        if (out->node_bounds)
            LUCIDA_HIP(hipMemcpyAsync(out->node_bounds, ctx->lb_bounds.ptr, sizeof(LucidaAabb), hipMemcpyDeviceToHost, s),
                       LUCIDA_READBACK);
    }
    LUCIDA_HIP(hipStreamSynchronize(s), LUCIDA_VALIDATION);
    return LUCIDA_OK;
}

lucida_status lucida_lbvh_build_device(lucida_ctx* ctx, const GpuTriangle* tris, size_t n, const float* positions,
                                       GpuBvhNode* nodes, size_t* num_nodes, uint32_t* prim_indices,
                                       GpuTriangle* tris_sorted, float* stage_us) {
    if (!ctx || !num_nodes || (n && (!tris || !positions || !nodes)))
        return set_error(LUCIDA_INVALID_ARG, "lucida_lbvh_build_device: null argument");
    clear_error();
    return lbvh_build_device(ctx, tris, n, positions, nodes, num_nodes, prim_indices, tris_sorted, stage_us);
}

lucida_status lucida_lbvh_build(lucida_ctx* ctx, const GpuTriangle* tris, size_t n, const float* positions,
                                size_t num_positions, GpuBvhNode* nodes_out, size_t* num_nodes,
                                uint32_t* prim_indices_out) {
    if (!ctx || !num_nodes || (n && (!tris || !positions || !nodes_out || !prim_indices_out)))
        return set_error(LUCIDA_INVALID_ARG, "lucida_lbvh_build: null argument");
    clear_error();
    if (n == 0) {
        *num_nodes = 0;
        return LUCIDA_OK;
    }
    for (size_t i = 0; i < n; ++i)
        if (tris[i].i0 >= num_positions || tris[i].i1 >= num_positions || tris[i].i2 >= num_positions)
            return set_error(LUCIDA_INVALID_ARG, "lucida_lbvh_build: triangle index out of range");
    hipStream_t s = ctx->stream;
    LUCIDA_HIP(ctx->lb_tris.ensure(n * sizeof(GpuTriangle)), LUCIDA_OOM);
    LUCIDA_HIP(ctx->lb_pos.ensure(num_positions * 16), LUCIDA_OOM);
    LUCIDA_HIP(ctx->lb_nodes.ensure((2 * n - 1) * sizeof(GpuBvhNode)), LUCIDA_OOM);
    LUCIDA_HIP(ctx->lb_prims.ensure(n * 4), LUCIDA_OOM);
    LUCIDA_HIP(hipMemcpyAsync(ctx->lb_tris.ptr, tris, n * sizeof(GpuTriangle), hipMemcpyHostToDevice, s),
               LUCIDA_VALIDATION);
    LUCIDA_HIP(hipMemcpyAsync(ctx->lb_pos.ptr, positions, num_positions * 16, hipMemcpyHostToDevice, s),
               LUCIDA_VALIDATION);
    lucida_status st = lbvh_build_device(ctx, ctx->lb_tris.as<GpuTriangle>(), n, ctx->lb_pos.as<float>(),
                                         ctx->lb_nodes.as<GpuBvhNode>(), num_nodes, ctx->lb_prims.as<uint32_t>(),
                                         nullptr, nullptr);
    if (st != LUCIDA_OK) return st;
    LUCIDA_HIP(hipMemcpyAsync(nodes_out, ctx->lb_nodes.ptr, *num_nodes * sizeof(GpuBvhNode), hipMemcpyDeviceToHost, s),
               LUCIDA_READBACK);
    LUCIDA_HIP(hipMemcpyAsync(prim_indices_out, ctx->lb_prims.ptr, n * 4, hipMemcpyDeviceToHost, s), LUCIDA_READBACK);
    LUCIDA_HIP(hipStreamSynchronize(s), LUCIDA_VALIDATION);
    return LUCIDA_OK;
}

}  // extern "C"
