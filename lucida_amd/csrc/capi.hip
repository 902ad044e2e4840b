// Context, device memory, scene residency and films behind the C ABI.
//
// Stand-ins for the reference's harness plumbing: Gpu::new / check_errors
// (crates/gpu/src/lib.rs:144-215) and the per-render buffer uploads of
// wavefront::render_native_on_inner (crates/gpu/src/wavefront.rs:215-296).
// Unlike the reference, a scene is uploaded once and stays resident.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "context.hpp"

namespace lucida {

static thread_local std::string g_last_error;

// Streaming copy, 16 bytes per lane per iteration.
__global__ void k_copy16(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

lucida_status set_error(lucida_status code, const char* msg) {
    g_last_error = msg ? msg : "";
    return code;
}
void clear_error() { g_last_error.clear(); }

lucida_status lbvh_build_device(lucida_ctx* ctx, const GpuTriangle* tris, size_t n, const float* positions,
                                GpuBvhNode* nodes, size_t* num_nodes, uint32_t* prim_indices,
                                GpuTriangle* tris_sorted, float* stage_us);

// Walks the tree at `root` (at most n nodes, so a cycle ends the walk) and
// checks that children stay inside the array and leaves inside [lo, hi).
static bool tree_is_sound(const GpuBvhNode* nodes, size_t n, uint32_t root, size_t lo, size_t hi) {
    if (root >= n) return false;
    std::vector<uint32_t> stack{root};
    size_t visited = 0;
    while (!stack.empty()) {
        if (++visited > n) return false;
        const GpuBvhNode& nd = nodes[stack.back()];
        stack.pop_back();
        if (nd.count == 0) {
            if ((size_t)nd.left_first + 1 >= n) return false;
            stack.push_back(nd.left_first);
            stack.push_back(nd.left_first + 1);
        } else if (nd.left_first < lo || (size_t)nd.left_first + nd.count > hi) {
            return false;
        }
    }
    return true;
}

template <class T>
static lucida_status upload(lucida_ctx* ctx, T** dst, const T* src, size_t count, size_t min_count = 1) {
    size_t bytes = (count > min_count ? count : min_count) * sizeof(T);
    LUCIDA_HIP(hipMalloc(reinterpret_cast<void**>(dst), bytes), LUCIDA_OOM);
    LUCIDA_HIP(hipMemsetAsync(*dst, 0, bytes, ctx->stream), LUCIDA_VALIDATION);
    if (count && src)
        LUCIDA_HIP(hipMemcpyAsync(*dst, src, count * sizeof(T), hipMemcpyHostToDevice, ctx->stream), LUCIDA_VALIDATION);
    return LUCIDA_OK;
}

static void scene_free(lucida_scene* s) {
    void* ptrs[] = {s->materials, s->primitives, s->positions,  s->vertex_attrs,    s->triangles,   s->tri_edges,
                    s->prim_w,    s->bvh_nodes,  s->lights,     s->env_pixels,      s->env_conditional, s->env_marginal};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
}

}  // namespace lucida

using namespace lucida;

extern "C" {

const char* lucida_last_error(void) { return g_last_error.c_str(); }

lucida_status lucida_ctx_create(int device, lucida_ctx** out) {
    if (!out) return set_error(LUCIDA_INVALID_ARG, "lucida_ctx_create: null out");
    clear_error();
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
        return set_error(LUCIDA_NO_ADAPTER, "no HIP device is visible");
    if (device < 0 || device >= count) return set_error(LUCIDA_NO_DEVICE, "device index out of range");
    LUCIDA_HIP(hipSetDevice(device), LUCIDA_NO_DEVICE);
    hipDeviceProp_t prop;
    LUCIDA_HIP(hipGetDeviceProperties(&prop, device), LUCIDA_NO_DEVICE);
    auto* ctx = new lucida_ctx();
    ctx->device = device;
    ctx->cu_count = prop.multiProcessorCount;
    ctx->name = prop.gcnArchName;
    bool ok = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreate(&ctx->ev0) == hipSuccess && hipEventCreate(&ctx->ev1) == hipSuccess &&
              hipEventCreateWithFlags(&ctx->ev_resolved, hipEventDisableTiming) == hipSuccess;
    for (int k = 0; k < 2 && ok; ++k) {  // every stream at the default priority
        WfPipe& w = ctx->wf[k];
        ok = (k == 0 ? (w.s = ctx->stream, true) : hipStreamCreateWithFlags(&w.s, hipStreamNonBlocking) == hipSuccess) &&
             hipStreamCreateWithFlags(&w.s2, hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&w.ev_shaded, hipEventDisableTiming) == hipSuccess &&
             hipEventCreateWithFlags(&w.ev_connected, hipEventDisableTiming) == hipSuccess &&
             hipEventCreateWithFlags(&w.ev_sorted, hipEventDisableTiming) == hipSuccess &&
             hipEventCreateWithFlags(&w.ev_shadow_sorted, hipEventDisableTiming) == hipSuccess;
    }
    if (!ok) {
        delete ctx;
        return set_error(LUCIDA_NO_DEVICE, "could not create the context's streams or events");
    }
    if (sort_scratch_init(ctx) != LUCIDA_OK) {
        std::string msg = lucida_last_error();
        lucida_ctx_destroy(ctx);
        return set_error(LUCIDA_NO_DEVICE, msg.c_str());
    }
    *out = ctx;
    return LUCIDA_OK;
}

void lucida_ctx_destroy(lucida_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (WfPipe& w : ctx->wf) {
        if (w.s) (void)hipStreamSynchronize(w.s);
        if (w.s2) (void)hipStreamSynchronize(w.s2);
    }
    DevBuf* bufs[] = {&ctx->lb_bounds, &ctx->lb_partial, &ctx->lb_codes, &ctx->lb_vals,   &ctx->lb_tmpk,
                      &ctx->lb_tmpv,   &ctx->lb_hist,    &ctx->lb_karras, &ctx->lb_parent, &ctx->lb_nodebox,
                      &ctx->lb_subtree, &ctx->lb_flags,  &ctx->lb_misc,  &ctx->lb_rank,   &ctx->lb_nodes,
                      &ctx->lb_tris,   &ctx->lb_pos,     &ctx->lb_prims, &ctx->lb_level, &ctx->ggx_table, &ctx->dn_buf,
                      &ctx->conv_buf, &ctx->mk_spill, &ctx->mk_state};
    for (DevBuf* b : bufs) b->release();
    for (WfPipe& w : ctx->wf) {
        DevBuf* wb[] = {&w.planes, &w.hits, &w.shadow, &w.queues, &w.counters, &w.uniforms, &w.spill};
        for (DevBuf* b : wb) b->release();
        hipEvent_t evs[] = {w.ev_shaded, w.ev_connected, w.ev_sorted, w.ev_shadow_sorted};
        for (hipEvent_t e : evs)
            if (e) (void)hipEventDestroy(e);
        if (w.s2) (void)hipStreamDestroy(w.s2);
        if (w.s && w.s != ctx->stream) (void)hipStreamDestroy(w.s);
    }
    (void)hipEventDestroy(ctx->ev0);
    (void)hipEventDestroy(ctx->ev1);
    (void)hipEventDestroy(ctx->ev_resolved);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

lucida_status lucida_ctx_synchronize(lucida_ctx* ctx) {
    if (!ctx) return set_error(LUCIDA_INVALID_ARG, "lucida_ctx_synchronize: null ctx");
    for (WfPipe& w : ctx->wf) {
        LUCIDA_HIP(hipStreamSynchronize(w.s2), LUCIDA_VALIDATION);
        LUCIDA_HIP(hipStreamSynchronize(w.s), LUCIDA_VALIDATION);
    }
    LUCIDA_HIP(hipGetLastError(), LUCIDA_VALIDATION);
    return LUCIDA_OK;
}

lucida_status lucida_ctx_info(lucida_ctx* ctx, char* name, size_t name_len, int* compute_units) {
    if (!ctx) return set_error(LUCIDA_INVALID_ARG, "lucida_ctx_info: null ctx");
    if (name && name_len) {
        std::strncpy(name, ctx->name.c_str(), name_len - 1);
        name[name_len - 1] = 0;
    }
    if (compute_units) *compute_units = ctx->cu_count;
    return LUCIDA_OK;
}

lucida_status lucida_device_alloc(lucida_ctx* ctx, size_t bytes, void** dptr) {
    if (!ctx || !dptr) return set_error(LUCIDA_INVALID_ARG, "lucida_device_alloc: null argument");
    LUCIDA_HIP(hipMalloc(dptr, bytes ? bytes : 4), LUCIDA_OOM);
    return LUCIDA_OK;
}

lucida_status lucida_device_free(lucida_ctx* ctx, void* dptr) {
    if (!ctx) return set_error(LUCIDA_INVALID_ARG, "lucida_device_free: null ctx");
    if (dptr) {
        LUCIDA_HIP(hipStreamSynchronize(ctx->stream), LUCIDA_VALIDATION);
        LUCIDA_HIP(hipFree(dptr), LUCIDA_VALIDATION);
    }
    return LUCIDA_OK;
}

lucida_status lucida_copy_to_device(lucida_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return set_error(LUCIDA_INVALID_ARG, "lucida_copy_to_device: null argument");
    if (!bytes) return LUCIDA_OK;
    LUCIDA_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream), LUCIDA_VALIDATION);
    LUCIDA_HIP(hipStreamSynchronize(ctx->stream), LUCIDA_VALIDATION);
    return LUCIDA_OK;
}

lucida_status lucida_copy_to_host(lucida_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return set_error(LUCIDA_INVALID_ARG, "lucida_copy_to_host: null argument");
    if (!bytes) return LUCIDA_OK;
    LUCIDA_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream), LUCIDA_READBACK);
    LUCIDA_HIP(hipStreamSynchronize(ctx->stream), LUCIDA_READBACK);
    return LUCIDA_OK;
}

lucida_status lucida_copy_device(lucida_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return set_error(LUCIDA_INVALID_ARG, "lucida_copy_device: null argument");
    if (((uintptr_t)dst | (uintptr_t)src | bytes) & 15u)
        return set_error(LUCIDA_INVALID_ARG, "lucida_copy_device: pointers and size must be multiples of 16");
    if (!bytes) return LUCIDA_OK;
    const size_t n = bytes / 16;
    const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_copy16, dim3(blocks), dim3(256), 0, ctx->stream, reinterpret_cast<const float4*>(src),
                       reinterpret_cast<float4*>(dst), n);
    LUCIDA_HIP(hipGetLastError(), LUCIDA_SHADER);
    LUCIDA_HIP(hipStreamSynchronize(ctx->stream), LUCIDA_VALIDATION);
    return LUCIDA_OK;
}

lucida_status lucida_scene_create(lucida_ctx* ctx, const lucida_scene_desc* d, lucida_scene** out) {
    if (!ctx || !d || !out) return set_error(LUCIDA_INVALID_ARG, "lucida_scene_create: null argument");
    clear_error();
    if (d->num_materials == 0) return set_error(LUCIDA_INVALID_ARG, "scene has no materials");
    for (size_t i = 0; i < d->num_triangles; ++i) {
        const GpuTriangle& t = d->triangles[i];
        if (t.i0 >= d->num_positions || t.i1 >= d->num_positions || t.i2 >= d->num_positions ||
            t.material >= d->num_materials)
            return set_error(LUCIDA_INVALID_ARG, "triangle index or material out of range");
    }
    for (size_t i = 0; i < d->num_primitives; ++i)
        if (d->primitives[i].material >= d->num_materials)
            return set_error(LUCIDA_INVALID_ARG, "primitive material out of range");
    for (size_t i = 0; i < d->num_lights; ++i)
        if (d->lights[i].material >= d->num_materials) return set_error(LUCIDA_INVALID_ARG, "light material out of range");
    for (size_t i = 0; i < d->num_bvh_nodes; ++i)
        if (d->bvh_nodes[i].count > 31) return set_error(LUCIDA_INVALID_ARG, "BVH leaf holds more than 31 primitives");
    if (d->num_instances == 0) {
        for (size_t i = 0; i < d->num_bvh_nodes; ++i) {
            const GpuBvhNode& nd = d->bvh_nodes[i];
            if (nd.count == 0 ? (size_t)nd.left_first + 1 >= d->num_bvh_nodes
                              : (size_t)nd.left_first + nd.count > d->num_triangles)
                return set_error(LUCIDA_INVALID_ARG, "BVH node references out of range");
        }
    } else {
        // Two-level layout (Scene::build_instances, scene.rs:252): the TLAS at
        // tlas_root indexes the instances that follow the primitives; each
        // instance's BLAS indexes the triangles.
        if (!d->instances || d->num_bvh_nodes == 0)
            return set_error(LUCIDA_INVALID_ARG, "instanced scene without instances or BVH nodes");
        if (!tree_is_sound(d->bvh_nodes, d->num_bvh_nodes, d->tlas_root, d->num_primitives,
                           d->num_primitives + d->num_instances))
            return set_error(LUCIDA_INVALID_ARG, "TLAS references out of range");
        std::vector<uint32_t> roots;
        for (size_t i = 0; i < d->num_instances; ++i) {
            const GpuInstance& in = d->instances[i];
            if (in.kind != LUCIDA_PRIM_KIND_INSTANCE || in.blas_root >= d->num_bvh_nodes ||
                (in.material != 0xFFFFFFFFu && in.material >= d->num_materials))
                return set_error(LUCIDA_INVALID_ARG, "instance record out of range");
            roots.push_back(in.blas_root);
        }
        std::sort(roots.begin(), roots.end());
        roots.erase(std::unique(roots.begin(), roots.end()), roots.end());
        for (uint32_t r : roots)
            if (!tree_is_sound(d->bvh_nodes, d->num_bvh_nodes, r, 0, d->num_triangles))
                return set_error(LUCIDA_INVALID_ARG, "BLAS references out of range");
    }
    auto* s = new lucida_scene();
    s->ctx = ctx;
    lucida_status st = LUCIDA_OK;
#define UP(dst, src, cnt)                                  \
    if (st == LUCIDA_OK) st = upload(ctx, &s->dst, src, cnt);
    UP(materials, d->materials, d->num_materials);
    // Instances follow the analytic primitives in one array, as uploaded by
    // SceneBlob::primitives_upload (gpu_layout.rs:743).
    std::vector<GpuPrimitive> prims(d->primitives, d->primitives + d->num_primitives);
    for (size_t i = 0; i < d->num_instances; ++i) {
        GpuPrimitive pr;
        std::memcpy(&pr, &d->instances[i], sizeof pr);
        prims.push_back(pr);
    }
    UP(primitives, prims.data(), prims.size());
    UP(positions, d->positions, d->num_positions * 4);
    UP(vertex_attrs, d->vertex_attrs, d->num_positions);
    UP(triangles, d->triangles, d->num_triangles);
    UP(lights, d->lights, d->num_lights);
    size_t env_px = (size_t)d->env_width * d->env_height;
    UP(env_pixels, d->env_pixels, env_px ? env_px * 4 : 0);
    UP(env_conditional, d->env_conditional, env_px ? (size_t)(d->env_width + 1) * d->env_height : 0);
    UP(env_marginal, d->env_marginal, env_px ? (size_t)d->env_height + 1 : 0);
#undef UP
    s->num_materials = d->num_materials;
    s->num_primitives = d->num_primitives;
    s->num_instances = d->num_instances;
    s->num_positions = d->num_positions;
    s->num_triangles = d->num_triangles;
    s->num_lights = d->num_lights;
    s->tlas_root = d->tlas_root;
    s->env_width = d->env_width;
    s->env_height = d->env_height;
    s->env_total_weight = d->env_total_weight;
    if (st == LUCIDA_OK) {
        if (d->num_bvh_nodes) {
            st = upload(ctx, &s->bvh_nodes, d->bvh_nodes, d->num_bvh_nodes);
            s->num_bvh_nodes = d->num_bvh_nodes;
        } else if (d->build_lbvh && d->num_triangles) {
            // LBVH on the device (restructured by treelets when build_lbvh is
            // 2), then the triangles permuted into leaf order
            // (Bvh::compact_primitives, bvh.rs:517) so leaves index them directly.
            size_t n = d->num_triangles;
            GpuTriangle* sorted = nullptr;
            if (hipMalloc(&s->bvh_nodes, (2 * n - 1) * sizeof(GpuBvhNode)) != hipSuccess ||
                hipMalloc(&sorted, n * sizeof(GpuTriangle)) != hipSuccess) {
                st = set_error(LUCIDA_OOM, "scene BVH allocation failed");
            } else {
                size_t num_nodes = 0;
                st = d->build_lbvh == 2
                         ? bvh_build_treelet_device(ctx, s->triangles, n, s->positions, LUCIDA_TREELET_ROUNDS,
                                                    s->bvh_nodes, &num_nodes, nullptr, sorted, nullptr)
                         : lbvh_build_device(ctx, s->triangles, n, s->positions, s->bvh_nodes, &num_nodes, nullptr,
                                             sorted, nullptr);
                if (st == LUCIDA_OK) {
                    (void)hipStreamSynchronize(ctx->stream);
                    (void)hipFree(s->triangles);
                    s->triangles = sorted;
                    sorted = nullptr;
                    s->num_bvh_nodes = num_nodes;
                }
            }
            if (sorted) (void)hipFree(sorted);
        }
    }
    if (st == LUCIDA_OK && s->num_triangles) st = build_tri_edges(ctx, s);
    if (st == LUCIDA_OK) st = build_prim_planes(ctx, s);
    if (st == LUCIDA_OK && hipStreamSynchronize(ctx->stream) != hipSuccess)
        st = set_error(LUCIDA_VALIDATION, "scene upload failed");
    if (st != LUCIDA_OK) {
        scene_free(s);
        delete s;
        return st;
    }
    *out = s;
    return LUCIDA_OK;
}

void lucida_scene_destroy(lucida_scene* s) {
    if (!s) return;
    (void)hipStreamSynchronize(s->ctx->stream);
    scene_free(s);
    delete s;
}

lucida_status lucida_scene_num_bvh_nodes(lucida_scene* s, size_t* out) {
    if (!s || !out) return set_error(LUCIDA_INVALID_ARG, "lucida_scene_num_bvh_nodes: null argument");
    *out = s->num_bvh_nodes;
    return LUCIDA_OK;
}

lucida_status lucida_scene_download_bvh(lucida_scene* s, GpuBvhNode* nodes, GpuTriangle* tris) {
    if (!s) return set_error(LUCIDA_INVALID_ARG, "lucida_scene_download_bvh: null scene");
    hipStream_t st = s->ctx->stream;
    if (nodes && s->num_bvh_nodes)
        LUCIDA_HIP(hipMemcpyAsync(nodes, s->bvh_nodes, s->num_bvh_nodes * sizeof(GpuBvhNode), hipMemcpyDeviceToHost, st),
                   LUCIDA_READBACK);
    if (tris && s->num_triangles)
        LUCIDA_HIP(hipMemcpyAsync(tris, s->triangles, s->num_triangles * sizeof(GpuTriangle), hipMemcpyDeviceToHost, st),
                   LUCIDA_READBACK);
    LUCIDA_HIP(hipStreamSynchronize(st), LUCIDA_READBACK);
    return LUCIDA_OK;
}

lucida_status lucida_film_create(lucida_ctx* ctx, uint32_t width, uint32_t height, lucida_film** out) {
    if (!ctx || !out || width == 0 || height == 0) return set_error(LUCIDA_INVALID_ARG, "lucida_film_create: bad argument");
    auto* f = new lucida_film();
    f->ctx = ctx;
    f->width = width;
    f->height = height;
    size_t bytes = (size_t)width * height * sizeof(GpuAccum);
    if (hipMalloc(&f->accum, bytes) != hipSuccess) {
        delete f;
        return set_error(LUCIDA_OOM, "film allocation failed");
    }
    if (hipMemsetAsync(f->accum, 0, bytes, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess) {
        (void)hipFree(f->accum);
        delete f;
        return set_error(LUCIDA_VALIDATION, "film clear failed");
    }
    *out = f;
    return LUCIDA_OK;
}

void lucida_film_destroy(lucida_film* f) {
    if (!f) return;
    (void)hipStreamSynchronize(f->ctx->stream);
    (void)hipFree(f->accum);
    delete f;
}

lucida_status lucida_film_clear(lucida_film* f) {
    if (!f) return set_error(LUCIDA_INVALID_ARG, "lucida_film_clear: null film");
    LUCIDA_HIP(hipMemsetAsync(f->accum, 0, (size_t)f->width * f->height * sizeof(GpuAccum), f->ctx->stream),
               LUCIDA_VALIDATION);
    LUCIDA_HIP(hipStreamSynchronize(f->ctx->stream), LUCIDA_VALIDATION);
    return LUCIDA_OK;
}

lucida_status lucida_film_download(lucida_film* f, GpuAccum* host_out) {
    if (!f || !host_out) return set_error(LUCIDA_INVALID_ARG, "lucida_film_download: null argument");
    LUCIDA_HIP(hipMemcpyAsync(host_out, f->accum, (size_t)f->width * f->height * sizeof(GpuAccum),
                              hipMemcpyDeviceToHost, f->ctx->stream),
               LUCIDA_READBACK);
    LUCIDA_HIP(hipStreamSynchronize(f->ctx->stream), LUCIDA_READBACK);
    return LUCIDA_OK;
}

lucida_status lucida_film_upload(lucida_film* f, const GpuAccum* host_in) {
    if (!f || !host_in) return set_error(LUCIDA_INVALID_ARG, "lucida_film_upload: null argument");
    LUCIDA_HIP(hipMemcpyAsync(f->accum, host_in, (size_t)f->width * f->height * sizeof(GpuAccum),
                              hipMemcpyHostToDevice, f->ctx->stream),
               LUCIDA_VALIDATION);
    LUCIDA_HIP(hipStreamSynchronize(f->ctx->stream), LUCIDA_VALIDATION);
    return LUCIDA_OK;
}

void* lucida_film_device_ptr(lucida_film* f) { return f ? f->accum : nullptr; }

}  // extern "C"
