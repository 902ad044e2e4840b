// Two-level scenes built on the device (Scene::build_instances, scene.rs:252).
//
// One LBVH per mesh (the BLASes) and one LBVH over the instances' world
// bounds (the TLAS), with the reference's flattened layout: every BLAS's nodes
// first, rebased so child indices are absolute in the shared node array and
// leaves index the shared triangle array; the TLAS last, its leaves indexing
// the primitive array where the instances follow the analytic primitives, in
// TLAS leaf order. The reference builds both levels with binned SAH on the
// host (bvh.rs Bvh::build / build_over_bounds); here both are the device LBVH,
// so a 4,096-instance TLAS and its BLAS take a few launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "context.hpp"

namespace lucida {
namespace {

// Copies n nodes, adding `internal_add` to internal nodes' child index and
// `leaf_add` to leaves' first primitive.
__global__ void k_rebase_nodes(const GpuBvhNode* __restrict__ in, size_t n, uint32_t internal_add, uint32_t leaf_add,
                               GpuBvhNode* __restrict__ out) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        GpuBvhNode nd = in[i];
        nd.left_first += nd.count == 0 ? internal_add : leaf_add;
        out[i] = nd;
    }
}

struct Mat34 {
    double m[3][4];
};

// Inverse of an affine [R | t]: [R^-1 | -R^-1 t], in double precision.
static bool invert_affine(const float* f, Mat34* out) {
    double a[3][3], t[3];
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) a[r][c] = f[4 * r + c];
        t[r] = f[4 * r + 3];
    }
    const double c00 = a[1][1] * a[2][2] - a[1][2] * a[2][1];
    const double c01 = a[1][2] * a[2][0] - a[1][0] * a[2][2];
    const double c02 = a[1][0] * a[2][1] - a[1][1] * a[2][0];
    const double det = a[0][0] * c00 + a[0][1] * c01 + a[0][2] * c02;
    if (!(std::fabs(det) > 1e-30) || !std::isfinite(det)) return false;
    const double inv_det = 1.0 / det;
    double r[3][3];
    r[0][0] = c00 * inv_det;
    r[1][0] = c01 * inv_det;
    r[2][0] = c02 * inv_det;
    r[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) * inv_det;
    r[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) * inv_det;
    r[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) * inv_det;
    r[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) * inv_det;
    r[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) * inv_det;
    r[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) * inv_det;
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) out->m[i][j] = r[i][j];
        out->m[i][3] = -(r[i][0] * t[0] + r[i][1] * t[1] + r[i][2] * t[2]);
    }
    return true;
}

// Instance::world_bounds (instance.rs:100): all eight corners of the object
// box through the forward transform, then Aabb::pad (bvh.rs:99).
static void world_bounds(const float* f, const GpuBvhNode& object, float lo[3], float hi[3]) {
    for (int k = 0; k < 3; ++k) {
        lo[k] = INFINITY;
        hi[k] = -INFINITY;
    }
    for (int c = 0; c < 8; ++c) {
        const double p[3] = {(c & 1) ? object.bounds_max[0] : object.bounds_min[0],
                             (c & 2) ? object.bounds_max[1] : object.bounds_min[1],
                             (c & 4) ? object.bounds_max[2] : object.bounds_min[2]};
        for (int k = 0; k < 3; ++k) {
            const float w = (float)(f[4 * k] * p[0] + f[4 * k + 1] * p[1] + f[4 * k + 2] * p[2] + f[4 * k + 3]);
            lo[k] = std::min(lo[k], w);
            hi[k] = std::max(hi[k], w);
        }
    }
    for (int k = 0; k < 3; ++k) {
        const float p = 1.0e-6f * (1.0f + std::max(std::fabs(lo[k]), std::fabs(hi[k])));
        lo[k] -= p;
        hi[k] += p;
    }
}

struct DevTmp {
    std::vector<void*> ptrs;
    ~DevTmp() {
        for (void* p : ptrs) (void)hipFree(p);
    }
    template <class T>
    T* alloc(size_t count, hipError_t* e) {
        void* p = nullptr;
        *e = hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T));
        if (*e == hipSuccess) ptrs.push_back(p);
        return reinterpret_cast<T*>(p);
    }
};

static lucida_status build_two_level(lucida_ctx* ctx, lucida_scene* s, const lucida_scene_desc* d,
                                     const lucida_mesh_range* meshes, size_t num_meshes,
                                     const lucida_instance_placement* placements, size_t num_placements) {
    hipStream_t st = ctx->stream;
    DevTmp tmp;
    hipError_t e = hipSuccess;
    // Both levels are LBVHs, restructured by treelets when build_lbvh is 2.
    const bool treelet = d->build_lbvh == 2;
    auto build_tree = [&](const GpuTriangle* t, size_t count, const float* pos, GpuBvhNode* out, size_t* nn,
                          uint32_t* order, GpuTriangle* sorted) {
        return treelet ? bvh_build_treelet_device(ctx, t, count, pos, LUCIDA_TREELET_ROUNDS, out, nn, order, sorted,
                                                  nullptr)
                       : lbvh_build_device(ctx, t, count, pos, out, nn, order, sorted, nullptr);
    };
    // --- BLASes -----------------------------------------------------------------
    size_t max_tris = 0;
    for (size_t m = 0; m < num_meshes; ++m) max_tris = std::max<size_t>(max_tris, meshes[m].triangle_count);
    GpuBvhNode* blas_nodes = tmp.alloc<GpuBvhNode>(2 * max_tris, &e);
    GpuTriangle* blas_sorted = tmp.alloc<GpuTriangle>(max_tris, &e);
    if (e != hipSuccess) return set_error(LUCIDA_OOM, "instancing: BLAS scratch allocation failed");
    std::vector<uint32_t> node_offset(num_meshes);
    std::vector<size_t> node_count(num_meshes);
    std::vector<GpuBvhNode> mesh_root(num_meshes);
    size_t total_blas = 0;
    for (size_t m = 0; m < num_meshes; ++m) total_blas += 2 * (size_t)meshes[m].triangle_count;
    GpuBvhNode* all_blas = tmp.alloc<GpuBvhNode>(total_blas, &e);
    if (e != hipSuccess) return set_error(LUCIDA_OOM, "instancing: node allocation failed");
    size_t off = 0;
    for (size_t m = 0; m < num_meshes; ++m) {
        const uint32_t first = meshes[m].first_triangle, count = meshes[m].triangle_count;
        size_t nn = 0;
        lucida_status r = build_tree(s->triangles + first, count, s->positions, blas_nodes, &nn, nullptr, blas_sorted);
        if (r != LUCIDA_OK) return r;
        // compact_primitives (bvh.rs:517): the mesh's slice takes leaf order
        LUCIDA_HIP(hipMemcpyAsync(s->triangles + first, blas_sorted, (size_t)count * sizeof(GpuTriangle),
                                  hipMemcpyDeviceToDevice, st),
                   LUCIDA_VALIDATION);
        hipLaunchKernelGGL(k_rebase_nodes, dim3((unsigned)std::min<size_t>((nn + 255) / 256, 4096)), dim3(256), 0, st,
                           blas_nodes, nn, (uint32_t)off, first, all_blas + off);
        LUCIDA_HIP(hipMemcpyAsync(&mesh_root[m], blas_nodes, sizeof(GpuBvhNode), hipMemcpyDeviceToHost, st),
                   LUCIDA_READBACK);
        LUCIDA_HIP(hipStreamSynchronize(st), LUCIDA_VALIDATION);  // blas_nodes is reused by the next mesh
        node_offset[m] = (uint32_t)off;
        node_count[m] = nn;
        off += nn;
    }
    const size_t blas_total = off;

    // --- TLAS over the instances' world bounds -----------------------------------
    const size_t n = num_placements;
    std::vector<float> box_pos(8 * n);
    std::vector<GpuTriangle> box_tris(n);
    std::vector<GpuInstance> inst(n);
    for (size_t i = 0; i < n; ++i) {
        const lucida_instance_placement& pl = placements[i];
        Mat34 w2o;
        if (!invert_affine(pl.object_to_world, &w2o))
            return set_error(LUCIDA_INVALID_ARG, "instancing: a placement's transform is singular");
        float lo[3], hi[3];
        world_bounds(pl.object_to_world, mesh_root[pl.mesh], lo, hi);
        for (int k = 0; k < 3; ++k) {
            box_pos[8 * i + k] = lo[k];
            box_pos[8 * i + 4 + k] = hi[k];
        }
        box_pos[8 * i + 3] = box_pos[8 * i + 7] = 0.0f;
        box_tris[i] = GpuTriangle{(uint32_t)(2 * i), (uint32_t)(2 * i + 1), (uint32_t)(2 * i), 0u};
        GpuInstance& g = inst[i];
        std::memset(&g, 0, sizeof g);
        for (int c = 0; c < 3; ++c) {
            g.inv_row0[c] = (float)w2o.m[0][c];
            g.inv_row1[c] = (float)w2o.m[1][c];
            g.inv_row2[c] = (float)w2o.m[2][c];
            g.inv_translation[c] = (float)w2o.m[c][3];
        }
        g.blas_root = node_offset[pl.mesh];
        g.kind = LUCIDA_PRIM_KIND_INSTANCE;
        g.material = pl.material;
    }
    float* d_pos = tmp.alloc<float>(8 * n, &e);
    GpuTriangle* d_tris = tmp.alloc<GpuTriangle>(n, &e);
    GpuTriangle* d_sorted = tmp.alloc<GpuTriangle>(n, &e);
    GpuBvhNode* tlas = tmp.alloc<GpuBvhNode>(2 * n, &e);
    uint32_t* order = tmp.alloc<uint32_t>(n, &e);
    if (e != hipSuccess) return set_error(LUCIDA_OOM, "instancing: TLAS scratch allocation failed");
    LUCIDA_HIP(hipMemcpyAsync(d_pos, box_pos.data(), box_pos.size() * 4, hipMemcpyHostToDevice, st), LUCIDA_VALIDATION);
    LUCIDA_HIP(hipMemcpyAsync(d_tris, box_tris.data(), n * sizeof(GpuTriangle), hipMemcpyHostToDevice, st),
               LUCIDA_VALIDATION);
    size_t tn = 0;
    lucida_status r = build_tree(d_tris, n, d_pos, tlas, &tn, order, d_sorted);
    if (r != LUCIDA_OK) return r;
    std::vector<uint32_t> h_order(n);
    LUCIDA_HIP(hipMemcpyAsync(h_order.data(), order, n * 4, hipMemcpyDeviceToHost, st), LUCIDA_READBACK);
    LUCIDA_HIP(hipStreamSynchronize(st), LUCIDA_READBACK);

    // --- assemble ------------------------------------------------------------------
    const size_t total_nodes = blas_total + tn;
    if (total_nodes >= (1u << 28)) return set_error(LUCIDA_INVALID_ARG, "instancing: too many BVH nodes");
    GpuBvhNode* nodes = nullptr;
    LUCIDA_HIP(hipMalloc(&nodes, total_nodes * sizeof(GpuBvhNode)), LUCIDA_OOM);
    LUCIDA_HIP(hipMemcpyAsync(nodes, all_blas, blas_total * sizeof(GpuBvhNode), hipMemcpyDeviceToDevice, st),
               LUCIDA_VALIDATION);
    const uint32_t prim_base = (uint32_t)d->num_primitives;
    hipLaunchKernelGGL(k_rebase_nodes, dim3((unsigned)std::min<size_t>((tn + 255) / 256, 4096)), dim3(256), 0, st,
                       tlas, tn, (uint32_t)blas_total, prim_base, nodes + blas_total);
    // instances in TLAS leaf order after the analytic primitives
    std::vector<GpuPrimitive> prims(d->primitives, d->primitives + d->num_primitives);
    for (size_t k = 0; k < n; ++k) {
        GpuPrimitive p;
        std::memcpy(&p, &inst[h_order[k]], sizeof p);
        prims.push_back(p);
    }
    GpuPrimitive* dprims = nullptr;
    LUCIDA_HIP(hipMalloc(&dprims, prims.size() * sizeof(GpuPrimitive)), LUCIDA_OOM);
    LUCIDA_HIP(hipMemcpyAsync(dprims, prims.data(), prims.size() * sizeof(GpuPrimitive), hipMemcpyHostToDevice, st),
               LUCIDA_VALIDATION);
    LUCIDA_HIP(hipGetLastError(), LUCIDA_SHADER);
    LUCIDA_HIP(hipStreamSynchronize(st), LUCIDA_VALIDATION);
    if (s->bvh_nodes) (void)hipFree(s->bvh_nodes);
    if (s->primitives) (void)hipFree(s->primitives);
    s->bvh_nodes = nodes;
    s->num_bvh_nodes = total_nodes;
    s->primitives = dprims;
    s->num_instances = n;
    s->tlas_root = (uint32_t)blas_total;
    const lucida_status es = build_tri_edges(ctx, s);
    return es != LUCIDA_OK ? es : build_prim_planes(ctx, s);
}

}  // namespace
}  // namespace lucida

using namespace lucida;

extern "C" {

lucida_status lucida_scene_create_instanced(lucida_ctx* ctx, const lucida_scene_desc* desc,
                                            const lucida_mesh_range* meshes, size_t num_meshes,
                                            const lucida_instance_placement* placements, size_t num_placements,
                                            lucida_scene** out) {
    if (!ctx || !desc || !out || (num_meshes && !meshes) || (num_placements && !placements))
        return set_error(LUCIDA_INVALID_ARG, "lucida_scene_create_instanced: null argument");
    if (desc->num_bvh_nodes || desc->num_instances)
        return set_error(LUCIDA_INVALID_ARG, "lucida_scene_create_instanced: desc already holds a BVH or instances");
    if (num_meshes == 0 || num_placements == 0)
        return set_error(LUCIDA_INVALID_ARG, "lucida_scene_create_instanced: no meshes or no placements");
    if (num_placements > (1u << 24)) return set_error(LUCIDA_INVALID_ARG, "instancing: more than 2^24 placements");
    for (size_t m = 0; m < num_meshes; ++m)
        if (meshes[m].triangle_count == 0 ||
            (uint64_t)meshes[m].first_triangle + meshes[m].triangle_count > desc->num_triangles)
            return set_error(LUCIDA_INVALID_ARG, "instancing: a mesh range is empty or outside the triangles");
    for (size_t i = 0; i < num_placements; ++i) {
        if (placements[i].mesh >= num_meshes) return set_error(LUCIDA_INVALID_ARG, "instancing: mesh index out of range");
        if (placements[i].material != 0xFFFFFFFFu && placements[i].material >= desc->num_materials)
            return set_error(LUCIDA_INVALID_ARG, "instancing: material override out of range");
        for (int k = 0; k < 12; ++k)
            if (!std::isfinite(placements[i].object_to_world[k]))
                return set_error(LUCIDA_INVALID_ARG, "instancing: non-finite transform");
    }
    lucida_scene_desc base = *desc;
    base.build_lbvh = 0;
    lucida_scene* s = nullptr;
    lucida_status st = lucida_scene_create(ctx, &base, &s);
    if (st != LUCIDA_OK) return st;
    st = build_two_level(ctx, s, desc, meshes, num_meshes, placements, num_placements);
    if (st != LUCIDA_OK) {
        lucida_scene_destroy(s);
        return st;
    }
    *out = s;
    return LUCIDA_OK;
}

lucida_status lucida_scene_counts(lucida_scene* s, size_t* num_triangles, size_t* num_bvh_nodes,
                                  size_t* num_instances, uint32_t* tlas_root) {
    if (!s) return set_error(LUCIDA_INVALID_ARG, "lucida_scene_counts: null scene");
    if (num_triangles) *num_triangles = s->num_triangles;
    if (num_bvh_nodes) *num_bvh_nodes = s->num_bvh_nodes;
    if (num_instances) *num_instances = s->num_instances;
    if (tlas_root) *tlas_root = s->tlas_root;
    return LUCIDA_OK;
}

lucida_status lucida_scene_download_instances(lucida_scene* s, GpuInstance* out) {
    if (!s || (s->num_instances && !out)) return set_error(LUCIDA_INVALID_ARG, "lucida_scene_download_instances: bad argument");
    if (!s->num_instances) return LUCIDA_OK;
    LUCIDA_HIP(hipMemcpyAsync(out, s->primitives + s->num_primitives, s->num_instances * sizeof(GpuInstance),
                              hipMemcpyDeviceToHost, s->ctx->stream),
               LUCIDA_READBACK);
    LUCIDA_HIP(hipStreamSynchronize(s->ctx->stream), LUCIDA_READBACK);
    return LUCIDA_OK;
}

}  // extern "C"
