// Context, scene and film objects behind the C ABI, and the HIP error plumbing.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>
#include <vector>

#include "../../include/lucida.h"
#include "status.hpp"

namespace lucida {

// Evaluates a HIP call; on failure records the message and returns `code`
// from the enclosing function.
#define LUCIDA_HIP(expr, code)                                                                 \
    do {                                                                                       \
        hipError_t e__ = (expr);                                                               \
        if (e__ != hipSuccess) {                                                               \
            char b__[512];                                                                     \
            snprintf(b__, sizeof b__, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e__), \
                     __FILE__, __LINE__);                                                      \
            return ::lucida::set_error(code, b__);                                             \
        }                                                                                      \
    } while (0)

// A device buffer owned by a context, grown on demand and reused.
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t want) {
        if (want <= bytes) return hipSuccess;
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
        hipError_t e = hipMalloc(&ptr, want);
        if (e == hipSuccess) bytes = want;
        return e;
    }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    template <class T>
    T* as() const { return reinterpret_cast<T*>(ptr); }
};

// The timing events of one call, destroyed on every return path.
template <int N>
struct StageEvents {
    hipEvent_t ev[N] = {};
    hipError_t create() {
        for (auto& e : ev) {
            const hipError_t r = hipEventCreate(&e);
            if (r != hipSuccess) return r;
        }
        return hipSuccess;
    }
    hipEvent_t& operator[](int k) { return ev[k]; }
    StageEvents() = default;
    StageEvents(const StageEvents&) = delete;
    StageEvents& operator=(const StageEvents&) = delete;
    ~StageEvents() {
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
    }
};

// One wavefront pipeline: a stream for GENERATE, EXTEND, SHADE, the
// extension sort and RESOLVE, a second one (default priority, like the
// first) for the shadow sort and CONNECT, the events ordering them, and the
// device buffers of its batch (path state, hits, shadow rays, queues,
// counters, uniforms, stack spill). A render runs its batches on the
// context's two pipelines in turn, so one pipeline's latency-bound SHADE and
// sorts overlap the other's walks.
struct WfPipe {
    hipStream_t s = nullptr, s2 = nullptr;
    hipEvent_t ev_shaded = nullptr, ev_connected = nullptr;
    // the two queue sorts, one per stream, joined before the walks start
    hipEvent_t ev_sorted = nullptr, ev_shadow_sorted = nullptr;
    DevBuf planes, hits, shadow, queues, counters, uniforms, spill;
    size_t capacity = 0;
};

}  // namespace lucida

struct lucida_ctx {
    int device = 0;
    int cu_count = 0;
    std::string name;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // Wavefront pipelines; wf[0].s is `stream`. Batched renders use wf[0].
    lucida::WfPipe wf[2];
    // A render's latest RESOLVE (batches accumulate into the film in order).
    hipEvent_t ev_resolved = nullptr;
    // LBVH scratch, reused between builds.
    lucida::DevBuf lb_bounds, lb_partial, lb_codes, lb_vals, lb_tmpk, lb_tmpv, lb_hist, lb_karras, lb_parent,
        lb_nodebox, lb_subtree, lb_flags, lb_misc, lb_rank, lb_nodes, lb_tris, lb_pos, lb_prims, lb_level;
    // Treelet restructuring: per-node costs, leaf choices, the leaf order.
    lucida::DevBuf tl_cost, tl_leafy, tl_order, tl_list;
    int lbvh_fit = 0;  // 0: two-phase fit (block-local, then the top), 1: atomic bottom-up, 2: level-synchronous
    // GGX directional-albedo table (energy compensation), 32 x 32 floats.
    lucida::DevBuf ggx_table;
    bool ggx_ready = false;
    // Denoiser planes.
    lucida::DevBuf dn_buf;
    // Convergence meter partials and result.
    lucida::DevBuf conv_buf;
    // Megakernel renders: traversal stack spill columns; counters and the
    // batches' uniforms.
    lucida::DevBuf mk_spill, mk_state;
};

struct lucida_scene {
    lucida_ctx* ctx = nullptr;
    GpuMaterial* materials = nullptr;
    GpuPrimitive* primitives = nullptr;  // analytic primitives, then instances
    float* positions = nullptr;
    GpuVertexAttr* vertex_attrs = nullptr;
    GpuTriangle* triangles = nullptr;
    float4* tri_edges = nullptr;  // per triangle: p0, p1 - p0, p2 - p0 (the intersection's own operands)
    float4* prim_w = nullptr;     // per analytic primitive a quad's w = n / |n|^2 (quad_w), then their boxes
    GpuBvhNode* bvh_nodes = nullptr;
    GpuLight* lights = nullptr;
    float* env_pixels = nullptr;
    float* env_conditional = nullptr;
    float* env_marginal = nullptr;
    size_t num_materials = 0, num_primitives = 0, num_instances = 0, num_positions = 0, num_triangles = 0,
           num_bvh_nodes = 0, num_lights = 0;
    uint32_t tlas_root = 0, env_width = 0, env_height = 0;
    float env_total_weight = 0.0f;
};

struct lucida_film {
    lucida_ctx* ctx = nullptr;
    uint32_t width = 0, height = 0;
    GpuAccum* accum = nullptr;
};

namespace lucida {
// Ensures the context's GGX energy table is resident (computed on the device).
lucida_status ensure_ggx_table(lucida_ctx* ctx);
// Device LBVH build (lbvh.hip): nodes in GpuBvhNode layout, optional prim
// order and triangles permuted into leaf order.
lucida_status lbvh_build_device(lucida_ctx* ctx, const GpuTriangle* tris, size_t n, const float* positions,
                                GpuBvhNode* nodes, size_t* num_nodes, uint32_t* prim_indices,
                                GpuTriangle* tris_sorted, float* stage_us);
// The LBVH build up to the fitted tree (lbvh.hip): Karras children, parents,
// every node's box, subtree sizes and the sorted order in the context's lb_*
// buffers; n >= 2.
lucida_status lbvh_fit_device(lucida_ctx* ctx, const GpuTriangle* tris, size_t n, const float* positions);
// The LBVH restructured by treelets (treelet.hip), laid out like
// lbvh_build_device's output; stage_us (optional) receives the fit, the
// restructuring rounds and the layout in microseconds.
lucida_status bvh_build_treelet_device(lucida_ctx* ctx, const GpuTriangle* tris, size_t n, const float* positions,
                                       uint32_t rounds, GpuBvhNode* nodes, size_t* num_nodes,
                                       uint32_t* prim_indices, GpuTriangle* tris_sorted, float* stage_us);
// Resolves the current device's copy of the sort's fixed scratch array
// (lbvh.hip), once per device; lucida_ctx_create calls it with ctx->device
// current.
lucida_status sort_scratch_init(lucida_ctx* ctx);
// Exclusive prefix sum of device words in place (lbvh.hip); `partials` holds
// ceil(count / 2048) words.
lucida_status scan_u32_device(lucida_ctx* ctx, uint32_t* data, uint32_t count, uint32_t* partials);
// Stable radix sort of (keys, vals)[0, *count) by the low `bits` (a multiple
// of 8) key bits on `stream` (lbvh.hip); the count is a device word at most
// `capacity`; tk/tv are temporaries of `capacity` words, `scratch` holds
// sort_counted_scratch_words(capacity) words, and *sorted_vals receives vals
// or tv, whichever holds the result (tv after an odd number of 8-bit passes).
size_t sort_counted_scratch_words(size_t capacity);
lucida_status sort_pairs_counted(hipStream_t stream, uint32_t* keys, uint32_t* vals, uint32_t* tk, uint32_t* tv,
                                 size_t capacity, const uint32_t* count, uint32_t bits, uint32_t* scratch,
                                 uint32_t** sorted_vals);
// Fills scene->prim_w (quad planes and the primitives' box) from its analytic
// primitives (none: a no-op).
lucida_status build_prim_planes(lucida_ctx* ctx, lucida_scene* scene);
// Fills scene->tri_edges from its (final, leaf-ordered) triangles.
lucida_status build_tri_edges(lucida_ctx* ctx, lucida_scene* scene);
// A-trous filter (atrous.hip) without the final wait: reserve the planes for
// the largest film first, then queue filters on the context's stream.
lucida_status atrous_reserve(lucida_ctx* ctx, size_t pixels);
void atrous_enqueue(lucida_ctx* ctx, const GpuAccum* in, uint32_t w, uint32_t h, const lucida_denoise_params& params,
                    GpuAccum* out);
}  // namespace lucida
