// Collapse of the binary BVH into the 4-wide BVH (widebvh.hpp), on the device.
//
// No reference counterpart: the reference traverses its binary GpuBvhNode
// tree directly (bvh.wgsl:137). The binary tree stays the scene's BVH of
// record (bit-exact with the CPU twin); the wide tree is derived from it and
// holds exactly the same leaves, so a ray meets the same triangles.
//
//   1. parents of every node (from the internal nodes' child indices);
//   2. depth parity of every internal node (walk to the root);
//   3. an exclusive scan over "internal and even depth" numbers the wide
//      nodes in the binary tree's preorder (the root is wide node 0);
//   4. each wide node gathers up to four children: a leaf child of its binary
//      node as a leaf, an internal child's two children (a leaf, or the wide
//      node of an even-depth internal node).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "context.hpp"
#include "widebvh.hpp"

namespace lucida {
namespace {

constexpr uint32_t NO_PARENT = 0xFFFFFFFFu;

__global__ void k_fill_u32(uint32_t* a, size_t n, uint32_t v) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) a[i] = v;
}

__global__ void k_parents(const GpuBvhNode* __restrict__ nodes, uint32_t n, uint32_t* __restrict__ parent) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const GpuBvhNode nd = nodes[i];
        if (nd.count == 0) {
            parent[nd.left_first] = i;
            parent[nd.left_first + 1] = i;
        }
    }
}

// flag[i] = 1 for internal nodes at even depth (root: depth 0).
__global__ void k_even_internal(const GpuBvhNode* __restrict__ nodes, uint32_t n, const uint32_t* __restrict__ parent,
                                uint32_t* __restrict__ flag) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        uint32_t f = 0;
        if (nodes[i].count == 0) {
            uint32_t depth = 0, x = i;
            while (parent[x] != NO_PARENT && depth <= n) {
                x = parent[x];
                ++depth;
            }
            f = (depth & 1u) ? 0u : 1u;
        }
        flag[i] = f;
    }
}

// Up to four children of a wide node, before encoding.
struct Children {
    float lo[4][3], hi[4][3];
    uint32_t ref[4];
    int n;
};

__device__ inline void add_child(Children& c, const GpuBvhNode& nd, uint32_t ref) {
    for (int a = 0; a < 3; ++a) {
        c.lo[c.n][a] = nd.bounds_min[a];
        c.hi[c.n][a] = nd.bounds_max[a];
    }
    c.ref[c.n] = ref;
    ++c.n;
}

__device__ inline uint32_t leaf_ref(const GpuBvhNode& c) { return WIDE_LEAF | (c.left_first << 5) | c.count; }

// Smallest e >= -126 with 255 * 2^e >= ext (ext >= 0, finite).
__device__ inline int scale_exponent(double ext) {
    if (!(ext > 0.0)) return -126;
    int e = ilogb(ext) - 8;
    if (e < -126) e = -126;
    while (ldexp(255.0, e) < ext) ++e;
    while (e > -126 && ldexp(255.0, e - 1) >= ext) --e;
    return e;
}

// Encodes the children relative to their union box: per axis the origin is
// the union's minimum and the scale 2^e the smallest power of two with
// 255 * 2^e covering the extent; each child plane is rounded outward
// (floor for minima, ceil for maxima, in exact double arithmetic), so
// origin + q * 2^e lies outside the child's box on every side. Returns false
// for a non-finite box or a scale beyond WIDE_MAX_EXP.
__device__ inline bool encode(const Children& c, WideNode& w) {
    bool ok = true;
    for (int k = 0; k < 4; ++k) {
        w.child[k] = WIDE_EMPTY;
        w.qlo_x[k] = w.qlo_y[k] = w.qlo_z[k] = 0;
        w.qhi_x[k] = w.qhi_y[k] = w.qhi_z[k] = 0;
    }
    w._pad0 = 0;
    w._pad1[0] = w._pad1[1] = 0u;
    for (int a = 0; a < 3; ++a) {
        float lo = c.lo[0][a], hi = c.hi[0][a];
        for (int k = 1; k < c.n; ++k) {
            lo = fminf(lo, c.lo[k][a]);
            hi = fmaxf(hi, c.hi[k][a]);
        }
        ok = ok && isfinite(lo) && isfinite(hi) && lo <= hi;
        const int e = ok ? scale_exponent((double)hi - (double)lo) : 0;
        ok = ok && e <= WIDE_MAX_EXP;
        w.origin[a] = lo;
        w.exp[a] = (int8_t)(ok ? e : 0);
        const double unit = ldexp(1.0, ok ? e : 0);
        for (int k = 0; k < c.n; ++k) {
            double ql = ok ? floor(((double)c.lo[k][a] - (double)lo) / unit) : 0.0;
            double qh = ok ? ceil(((double)c.hi[k][a] - (double)lo) / unit) : 255.0;
            ok = ok && ql >= 0.0 && qh <= 255.0 && ql <= qh;
            ql = fmin(fmax(ql, 0.0), 255.0);
            qh = fmin(fmax(qh, 0.0), 255.0);
            uint8_t* L = a == 0 ? w.qlo_x : a == 1 ? w.qlo_y : w.qlo_z;
            uint8_t* H = a == 0 ? w.qhi_x : a == 1 ? w.qhi_y : w.qhi_z;
            L[k] = (uint8_t)ql;
            H[k] = (uint8_t)qh;
        }
    }
    for (int k = 0; k < c.n; ++k) w.child[k] = c.ref[k];
    return ok;
}

__global__ void k_build_wide(const GpuBvhNode* __restrict__ nodes, uint32_t n, const uint32_t* __restrict__ flag,
                             const uint32_t* __restrict__ index, WideNode* __restrict__ wide,
                             uint32_t* __restrict__ bad) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (!flag[i]) continue;
        const GpuBvhNode x = nodes[i];
        Children c;
        c.n = 0;
        for (uint32_t j = x.left_first; j < x.left_first + 2; ++j) {
            const GpuBvhNode cn = nodes[j];
            if (cn.count > 0) {
                add_child(c, cn, leaf_ref(cn));
            } else {
                for (uint32_t g = cn.left_first; g < cn.left_first + 2; ++g) {
                    const GpuBvhNode gn = nodes[g];
                    add_child(c, gn, gn.count > 0 ? leaf_ref(gn) : index[g]);
                }
            }
        }
        WideNode w;
        if (!encode(c, w)) atomicOr(bad, 1u);
        wide[index[i]] = w;
    }
}

// A tree that is a single leaf: one wide node holding it.
__global__ void k_wide_single_leaf(const GpuBvhNode* __restrict__ nodes, WideNode* __restrict__ wide,
                                   uint32_t* __restrict__ bad) {
    Children c;
    c.n = 0;
    add_child(c, nodes[0], leaf_ref(nodes[0]));
    WideNode w;
    if (!encode(c, w)) atomicOr(bad, 1u);
    wide[0] = w;
}

}  // namespace

lucida_status build_wide_bvh(lucida_ctx* ctx, lucida_scene* scene) {
    if (scene->wide) {
        (void)hipFree(scene->wide);
        scene->wide = nullptr;
        scene->num_wide = 0;
    }
    const size_t n = scene->num_bvh_nodes;
    if (n == 0 || scene->num_instances) return LUCIDA_OK;
    if (n >= (1u << 28)) return set_error(LUCIDA_INVALID_ARG, "wide BVH: too many nodes");
    hipStream_t s = ctx->stream;
    const GpuBvhNode* nodes = scene->bvh_nodes;
    GpuBvhNode root;
    LUCIDA_HIP(hipMemcpyAsync(&root, nodes, sizeof root, hipMemcpyDeviceToHost, s), LUCIDA_READBACK);
    LUCIDA_HIP(hipStreamSynchronize(s), LUCIDA_READBACK);
    uint32_t* parent = nullptr;
    uint32_t* flag = nullptr;
    uint32_t* index = nullptr;
    uint32_t* partials = nullptr;  // scan partials, then the encoder's failure flag in its last word
    lucida_status st = LUCIDA_OK;
    const size_t segs = (n + 2047) / 2048;
    if (hipMalloc(&parent, n * 4) != hipSuccess || hipMalloc(&flag, n * 4) != hipSuccess ||
        hipMalloc(&index, n * 4) != hipSuccess || hipMalloc(&partials, segs * 4 + 8) != hipSuccess)
        st = set_error(LUCIDA_OOM, "wide BVH: scratch allocation failed");
    uint32_t* bad = partials ? partials + segs + 1 : nullptr;
    if (st == LUCIDA_OK) hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, bad, (size_t)1, 0u);
    if (st == LUCIDA_OK && root.count > 0) {
        if (hipMalloc(&scene->wide, sizeof(WideNode)) != hipSuccess) {
            st = set_error(LUCIDA_OOM, "wide BVH: node allocation failed");
        } else {
            hipLaunchKernelGGL(k_wide_single_leaf, dim3(1), dim3(1), 0, s, nodes, (WideNode*)scene->wide, bad);
            scene->num_wide = 1;
        }
    }
    const bool single = root.count > 0;
    const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
    uint32_t last_flag = 0, last_index = 0;
    if (st == LUCIDA_OK && !single) {
        hipLaunchKernelGGL(k_fill_u32, dim3(blocks), dim3(256), 0, s, parent, n, NO_PARENT);
        hipLaunchKernelGGL(k_parents, dim3(blocks), dim3(256), 0, s, nodes, (uint32_t)n, parent);
        hipLaunchKernelGGL(k_even_internal, dim3(blocks), dim3(256), 0, s, nodes, (uint32_t)n, parent, flag);
        if (hipMemcpyAsync(index, flag, n * 4, hipMemcpyDeviceToDevice, s) != hipSuccess)
            st = set_error(LUCIDA_VALIDATION, "wide BVH: copy failed");
    }
    if (st == LUCIDA_OK && !single) st = scan_u32_device(ctx, index, (uint32_t)n, partials);
    if (st == LUCIDA_OK && !single &&
        (hipMemcpyAsync(&last_flag, flag + n - 1, 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
         hipMemcpyAsync(&last_index, index + n - 1, 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
         hipStreamSynchronize(s) != hipSuccess))
        st = set_error(LUCIDA_READBACK, "wide BVH: count readback failed");
    const size_t total = (size_t)last_index + last_flag;
    if (st == LUCIDA_OK && !single &&
        (total == 0 || hipMalloc(&scene->wide, total * sizeof(WideNode)) != hipSuccess))
        st = set_error(LUCIDA_OOM, "wide BVH: node allocation failed");
    if (st == LUCIDA_OK && !single) {
        hipLaunchKernelGGL(k_build_wide, dim3(blocks), dim3(256), 0, s, nodes, (uint32_t)n, flag, index,
                           (WideNode*)scene->wide, bad);
        scene->num_wide = total;
    }
    uint32_t h_bad = 0;
    if (st == LUCIDA_OK && (hipGetLastError() != hipSuccess ||
                            hipMemcpyAsync(&h_bad, bad, 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
                            hipStreamSynchronize(s) != hipSuccess))
        st = set_error(LUCIDA_SHADER, "wide BVH: build kernel failed");
    (void)hipStreamSynchronize(s);
    for (void* p : {(void*)parent, (void*)flag, (void*)index, (void*)partials})
        if (p) (void)hipFree(p);
    // A box the 8-bit encoding cannot hold (non-finite, or wider than
    // 255 * 2^WIDE_MAX_EXP): the scene keeps only its binary tree.
    if ((st != LUCIDA_OK || h_bad) && scene->wide) {
        (void)hipFree(scene->wide);
        scene->wide = nullptr;
        scene->num_wide = 0;
    }
    return st;
}

}  // namespace lucida
