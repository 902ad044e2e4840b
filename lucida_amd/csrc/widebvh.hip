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

__device__ inline void put_child(WideNode& w, int k, const GpuBvhNode& c, uint32_t ref) {
    w.lo_x[k] = c.bounds_min[0];
    w.lo_y[k] = c.bounds_min[1];
    w.lo_z[k] = c.bounds_min[2];
    w.hi_x[k] = c.bounds_max[0];
    w.hi_y[k] = c.bounds_max[1];
    w.hi_z[k] = c.bounds_max[2];
    w.child[k] = ref;
}

__device__ inline uint32_t leaf_ref(const GpuBvhNode& c) { return WIDE_LEAF | (c.left_first << 5) | c.count; }

__global__ void k_build_wide(const GpuBvhNode* __restrict__ nodes, uint32_t n, const uint32_t* __restrict__ flag,
                             const uint32_t* __restrict__ index, WideNode* __restrict__ wide) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (!flag[i]) continue;
        const GpuBvhNode x = nodes[i];
        WideNode w;
        for (int k = 0; k < 4; ++k) {
            w.lo_x[k] = w.lo_y[k] = w.lo_z[k] = 0.0f;
            w.hi_x[k] = w.hi_y[k] = w.hi_z[k] = 0.0f;
            w.child[k] = WIDE_EMPTY;
            w._pad[k] = 0u;
        }
        int k = 0;
        for (uint32_t c = x.left_first; c < x.left_first + 2; ++c) {
            const GpuBvhNode cn = nodes[c];
            if (cn.count > 0) {
                put_child(w, k++, cn, leaf_ref(cn));
            } else {
                for (uint32_t g = cn.left_first; g < cn.left_first + 2; ++g) {
                    const GpuBvhNode gn = nodes[g];
                    put_child(w, k++, gn, gn.count > 0 ? leaf_ref(gn) : index[g]);
                }
            }
        }
        wide[index[i]] = w;
    }
}

// A tree that is a single leaf: one wide node holding it.
__global__ void k_wide_single_leaf(const GpuBvhNode* __restrict__ nodes, WideNode* __restrict__ wide) {
    WideNode w;
    for (int k = 0; k < 4; ++k) {
        w.lo_x[k] = w.lo_y[k] = w.lo_z[k] = 0.0f;
        w.hi_x[k] = w.hi_y[k] = w.hi_z[k] = 0.0f;
        w.child[k] = WIDE_EMPTY;
        w._pad[k] = 0u;
    }
    put_child(w, 0, nodes[0], leaf_ref(nodes[0]));
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
    if (root.count > 0) {
        LUCIDA_HIP(hipMalloc(&scene->wide, sizeof(WideNode)), LUCIDA_OOM);
        hipLaunchKernelGGL(k_wide_single_leaf, dim3(1), dim3(1), 0, s, nodes, (WideNode*)scene->wide);
        LUCIDA_HIP(hipGetLastError(), LUCIDA_SHADER);
        scene->num_wide = 1;
        return LUCIDA_OK;
    }
    uint32_t* parent = nullptr;
    uint32_t* flag = nullptr;
    uint32_t* index = nullptr;
    uint32_t* partials = nullptr;
    lucida_status st = LUCIDA_OK;
    const size_t segs = (n + 2047) / 2048;
    if (hipMalloc(&parent, n * 4) != hipSuccess || hipMalloc(&flag, n * 4) != hipSuccess ||
        hipMalloc(&index, n * 4) != hipSuccess || hipMalloc(&partials, segs * 4 + 4) != hipSuccess)
        st = set_error(LUCIDA_OOM, "wide BVH: scratch allocation failed");
    const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
    uint32_t last_flag = 0, last_index = 0;
    if (st == LUCIDA_OK) {
        hipLaunchKernelGGL(k_fill_u32, dim3(blocks), dim3(256), 0, s, parent, n, NO_PARENT);
        hipLaunchKernelGGL(k_parents, dim3(blocks), dim3(256), 0, s, nodes, (uint32_t)n, parent);
        hipLaunchKernelGGL(k_even_internal, dim3(blocks), dim3(256), 0, s, nodes, (uint32_t)n, parent, flag);
        if (hipMemcpyAsync(index, flag, n * 4, hipMemcpyDeviceToDevice, s) != hipSuccess)
            st = set_error(LUCIDA_VALIDATION, "wide BVH: copy failed");
    }
    if (st == LUCIDA_OK) st = scan_u32_device(ctx, index, (uint32_t)n, partials);
    if (st == LUCIDA_OK &&
        (hipMemcpyAsync(&last_flag, flag + n - 1, 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
         hipMemcpyAsync(&last_index, index + n - 1, 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
         hipStreamSynchronize(s) != hipSuccess))
        st = set_error(LUCIDA_READBACK, "wide BVH: count readback failed");
    const size_t total = (size_t)last_index + last_flag;
    if (st == LUCIDA_OK && (total == 0 || hipMalloc(&scene->wide, total * sizeof(WideNode)) != hipSuccess))
        st = set_error(LUCIDA_OOM, "wide BVH: node allocation failed");
    if (st == LUCIDA_OK) {
        hipLaunchKernelGGL(k_build_wide, dim3(blocks), dim3(256), 0, s, nodes, (uint32_t)n, flag, index,
                           (WideNode*)scene->wide);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
            st = set_error(LUCIDA_SHADER, "wide BVH: build kernel failed");
        else
            scene->num_wide = total;
    }
    (void)hipStreamSynchronize(s);
    for (void* p : {(void*)parent, (void*)flag, (void*)index, (void*)partials})
        if (p) (void)hipFree(p);
    if (st != LUCIDA_OK && scene->wide) {
        (void)hipFree(scene->wide);
        scene->wide = nullptr;
        scene->num_wide = 0;
    }
    return st;
}

}  // namespace lucida
