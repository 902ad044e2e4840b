/*
 * lucida.h — the C-ABI drop-in boundary of the MI355X-native Lucida hot path.
 *
 * Two halves:
 *
 *  1. The byte layout of every buffer that crosses the boundary. Each struct is
 *     byte-identical to its `#[repr(C)]` twin in the reference's
 *     crates/core/src/gpu_layout.rs, and every offset that file asserts is
 *     asserted again below (gpu_layout.rs:993-1166).
 *
 *  2. The entry points. Each replaces one call of the reference's `pt-gpu` crate
 *     (crates/gpu/src/lib.rs, wavefront.rs, lbvh.rs, denoise.rs); the comment on
 *     each names the call it stands in for. Errors are status codes mirroring
 *     `GpuError` (crates/gpu/src/lib.rs:100-112) plus a thread-local message,
 *     standing in for the error's String.
 *
 * Plain pointers and sizes only: no torch, HIP or C++ types in the signatures.
 * Pointers documented as "host" are ordinary host memory; pointers documented
 * as "device" are device addresses on the context's GPU.
 */
#ifndef LUCIDA_H
#define LUCIDA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Byte layout (crates/core/src/gpu_layout.rs)                              */
/* ------------------------------------------------------------------------ */

/* gpu_layout.rs:46 — principled material, 80 B. */
typedef struct GpuMaterial {
    float base_color[3]; /* 0 */
    float metallic;      /* 12 */
    float emissive[3];   /* 16 */
    float roughness;     /* 28 */
    float ior;           /* 32 */
    float transmission;  /* 36 */
    float _pad0;         /* 40 */
    float _pad1;         /* 44 */
    float eta[3];        /* 48 */
    float _pad2;         /* 60 */
    float k[3];          /* 64 */
    float _pad3;         /* 76, size 80 */
} GpuMaterial;

#define LUCIDA_PRIM_KIND_SPHERE 0u
#define LUCIDA_PRIM_KIND_QUAD 1u
#define LUCIDA_PRIM_KIND_INSTANCE 2u

/* gpu_layout.rs:195 — analytic sphere or parallelogram, 64 B. */
typedef struct GpuPrimitive {
    float position[3]; /* 0 */
    float radius;      /* 12 */
    float edge_u[3];   /* 16 */
    uint32_t kind;     /* 28 */
    float edge_v[3];   /* 32 */
    uint32_t material; /* 44 */
    float normal[3];   /* 48 */
    float _pad0;       /* 60, size 64 */
} GpuPrimitive;

/* gpu_layout.rs:256 — per-vertex shading attributes, 32 B. */
typedef struct GpuVertexAttr {
    float normal[3]; /* 0 */
    float _pad0;     /* 12 */
    float uv[2];     /* 16 */
    float _pad1[2];  /* 24, size 32 */
} GpuVertexAttr;

/* gpu_layout.rs:270 — indexed triangle, 16 B. */
typedef struct GpuTriangle {
    uint32_t i0, i1, i2; /* 0, 4, 8 */
    uint32_t material;   /* 12, size 16 */
} GpuTriangle;

/* gpu_layout.rs:298 — binary BVH node, 32 B. count==0: internal, left child at
 * left_first, right child at left_first+1. count>0: leaf over
 * prim_indices[left_first .. left_first+count). */
typedef struct GpuBvhNode {
    float bounds_min[3]; /* 0 */
    uint32_t left_first; /* 12 */
    float bounds_max[3]; /* 16 */
    uint32_t count;      /* 28, size 32 */
} GpuBvhNode;

/* gpu_layout.rs:345 — one pixel of the accumulation buffer, 64 B. Sums, not
 * means: a Film pixel is radiance / samples. */
typedef struct GpuAccum {
    float radiance[3];    /* 0 */
    float samples;        /* 12 */
    float albedo[3];      /* 16 */
    float depth;          /* 28 */
    float normal[3];      /* 32 */
    float traversal;      /* 44 */
    float radiance_sq[3]; /* 48 */
    float _pad0;          /* 60, size 64 */
} GpuAccum;

/* gpu_layout.rs:395 — per-dispatch uniforms, 160 B. */
typedef struct GpuUniforms {
    float cam_origin[3];      /* 0 */
    float lens_radius;        /* 12 */
    float cam_upper_left[3];  /* 16 */
    float focus_distance;     /* 28 */
    float cam_horizontal[3];  /* 32 */
    float _pad1;              /* 44 */
    float cam_vertical[3];    /* 48 */
    float _pad2;              /* 60 */
    uint32_t width;           /* 64 */
    uint32_t height;          /* 68 */
    uint32_t sample_offset;   /* 72 */
    uint32_t samples_per_launch; /* 76 */
    uint32_t max_depth;       /* 80 */
    uint32_t frame_seed;      /* 84 */
    uint32_t num_primitives;  /* 88 */
    uint32_t num_lights;      /* 92 */
    float background[3];      /* 96 */
    uint32_t num_triangles;   /* 108 */
    uint32_t num_bvh_nodes;   /* 112 */
    uint32_t render_mode;     /* 116 */
    uint32_t sampling_mode;   /* 120 */
    uint32_t env_width;       /* 124 */
    uint32_t env_height;      /* 128 */
    float env_total_weight;   /* 132 */
    uint32_t sampler_kind;    /* 136 */
    uint32_t num_instances;   /* 140 */
    uint32_t tlas_root;       /* 144 */
    uint32_t _pad3;           /* 148 */
    uint32_t _pad4;           /* 152 */
    uint32_t _pad5;           /* 156, size 160 */
} GpuUniforms;

/* gpu_layout.rs:496 — instance overlaying GpuPrimitive, 64 B. */
typedef struct GpuInstance {
    float inv_row0[3];        /* 0 */
    uint32_t blas_root;       /* 12 */
    float inv_row1[3];        /* 16 */
    uint32_t kind;            /* 28 */
    float inv_row2[3];        /* 32 */
    uint32_t material;        /* 44 */
    float inv_translation[3]; /* 48 */
    float _pad0;              /* 60, size 64 */
} GpuInstance;

#define LUCIDA_LIGHT_KIND_QUAD 0u
#define LUCIDA_LIGHT_KIND_TRIANGLE 1u

/* gpu_layout.rs:556 — one flattened area light, 64 B. */
typedef struct GpuLight {
    float origin[3];   /* 0 */
    float area;        /* 12 */
    float edge_u[3];   /* 16 */
    uint32_t kind;     /* 28 */
    float edge_v[3];   /* 32 */
    uint32_t material; /* 44 */
    float normal[3];   /* 48 */
    float _pad0;       /* 60, size 64 */
} GpuLight;

/* gpu_layout.rs:583 — the reference's inter-stage path record, 112 B. Kept for
 * layout compatibility; this build streams the same fields as separate
 * 16-byte planes instead (see lucida_amd/csrc/wavefront.hip). */
typedef struct GpuPathState {
    float origin[3];        /* 0 */
    uint32_t pixel;         /* 12 */
    float direction[3];     /* 16 */
    uint32_t rng_state;     /* 28 */
    float throughput[3];    /* 32 */
    float prev_bsdf_pdf;    /* 44 */
    float radiance[3];      /* 48 */
    uint32_t depth;         /* 60 */
    float prev_position[3]; /* 64 */
    uint32_t sampler_dim;   /* 76 */
    float guide_albedo[3];  /* 80 */
    float guide_depth;      /* 92 */
    float guide_normal[3];  /* 96 */
    float _pad1;            /* 108, size 112 */
} GpuPathState;

/* gpu_layout.rs:624 — EXTEND's output, 64 B. */
typedef struct GpuHitRecord {
    float position[3];         /* 0 */
    float t;                   /* 12 */
    float normal[3];           /* 16 */
    uint32_t material;         /* 28 */
    float geometric_normal[3]; /* 32 */
    float light_area;          /* 44 */
    uint32_t front_face;       /* 48 */
    uint32_t valid;            /* 52 */
    uint32_t _pad0;            /* 56 */
    uint32_t _pad1;            /* 60, size 64 */
} GpuHitRecord;

/* gpu_layout.rs:644 — deferred shadow ray, 48 B. */
typedef struct GpuShadowRay {
    float origin[3];       /* 0 */
    float distance;        /* 12 */
    float direction[3];    /* 16 */
    uint32_t path;         /* 28 */
    float contribution[3]; /* 32 */
    float _pad0;           /* 44, size 48 */
} GpuShadowRay;

/* gpu_layout.rs:676 / :690 — dispatch arguments and queue counters. */
typedef struct GpuDispatchArgs {
    uint32_t trace[3];  /* 0 */
    uint32_t shadow[3]; /* 12 */
    uint32_t _pad0[2];  /* 24, size 32 */
} GpuDispatchArgs;

typedef struct GpuWavefrontCounters {
    uint32_t next_queue_len; /* 0 */
    uint32_t next_trace_x;   /* 4 */
    uint32_t shadow_len;     /* 8 */
    uint32_t shadow_count;   /* 12, size 16 */
} GpuWavefrontCounters;

/* crates/gpu/src/lbvh.rs:79-86 — the LBVH builder's box: lo, pad, hi, pad. */
typedef struct LucidaAabb {
    float lo[3];
    float _p0;
    float hi[3];
    float _p1;
} LucidaAabb;

/* Sampling modes (integrator.rs SamplingMode::index) and sampler kinds
 * (sobol.rs SamplerKind::index). */
#define LUCIDA_SAMPLING_BSDF 0u
#define LUCIDA_SAMPLING_NEE 1u
#define LUCIDA_SAMPLING_MIS 2u
#define LUCIDA_SAMPLER_INDEPENDENT 0u
#define LUCIDA_SAMPLER_SOBOL 1u

/* ------------------------------------------------------------------------ */
/* Status and context                                                       */
/* ------------------------------------------------------------------------ */

/* Mirrors GpuError {NoAdapter, NoDevice, Shader, Readback, Validation}
 * (crates/gpu/src/lib.rs:100-112) plus the two failure kinds a C caller needs. */
typedef enum lucida_status {
    LUCIDA_OK = 0,
    LUCIDA_NO_ADAPTER = 1,
    LUCIDA_NO_DEVICE = 2,
    LUCIDA_SHADER = 3,
    LUCIDA_READBACK = 4,
    LUCIDA_VALIDATION = 5,
    LUCIDA_INVALID_ARG = 6,
    LUCIDA_OOM = 7
} lucida_status;

typedef struct lucida_ctx lucida_ctx;
typedef struct lucida_scene lucida_scene;
typedef struct lucida_film lucida_film;

/* Thread-local message of the last failing call on this thread ("" if none).
 * Stands in for the String carried by GpuError. */
const char* lucida_last_error(void);

/* Gpu::new (crates/gpu/src/lib.rs:144). Binds GPU `device`, creates the
 * context's stream and loads the kernels. NO_ADAPTER when no GPU is visible. */
lucida_status lucida_ctx_create(int device, lucida_ctx** out);
/* Scenes, films and networks created on a context must be destroyed before it. */
void lucida_ctx_destroy(lucida_ctx* ctx);
/* Gpu::check_errors (crates/gpu/src/lib.rs:204): waits for the context's stream
 * and surfaces any asynchronous device error as VALIDATION. */
lucida_status lucida_ctx_synchronize(lucida_ctx* ctx);
/* Device name and compute-unit count, for reports. */
lucida_status lucida_ctx_info(lucida_ctx* ctx, char* name, size_t name_len, int* compute_units);

/* Device memory helpers for callers that keep inputs resident (the bench, the
 * batched path). Allocation is on the context's GPU; copies are ordered on the
 * context's stream and complete before return. */
lucida_status lucida_device_alloc(lucida_ctx* ctx, size_t bytes, void** dptr);
lucida_status lucida_device_free(lucida_ctx* ctx, void* dptr);
lucida_status lucida_copy_to_device(lucida_ctx* ctx, void* dst, const void* src, size_t bytes);
lucida_status lucida_copy_to_host(lucida_ctx* ctx, void* dst, const void* src, size_t bytes);
/* Device-to-device copy by a streaming kernel (16-byte loads and stores).
 * dst, src and bytes must be multiples of 16. No reference counterpart: it is
 * the known-size kernel the HBM byte counters are calibrated against. */
lucida_status lucida_copy_device(lucida_ctx* ctx, void* dst, const void* src, size_t bytes);

/* ------------------------------------------------------------------------ */
/* LBVH (crates/gpu/src/lbvh.rs)                                            */
/* ------------------------------------------------------------------------ */

/* Chooses how the LBVH fit orders its bottom-up unions: 0 (default) two
 * phases, the subtrees inside each block of 1024 sorted primitives by one
 * workgroup each (the atomic walk with its state in LDS), then the nodes that
 * span blocks by one workgroup; 1 the atomic bottom-up walk over the whole
 * tree (the second arrival at each node continues), which needs agent-scope
 * acquire/release between XCDs; 2 one launch per tree level, deepest first.
 * All produce identical boxes (crates/core/src/lbvh.rs:490-525). */
lucida_status lucida_ctx_set_lbvh_fit(lucida_ctx* ctx, int mode);

/* LbvhBuilder::sort (crates/gpu/src/lbvh.rs:669): stable ascending sort of u32
 * keys carrying u32 values. Host arrays; n==0 is not an error. */
lucida_status lucida_sort_u32(lucida_ctx* ctx, const uint32_t* keys, const uint32_t* vals, size_t n,
                              uint32_t* keys_out, uint32_t* vals_out);
/* Same sort on device arrays, in place: keys/vals are sorted, tmp_keys/tmp_vals
 * are scratch of n words each. */
lucida_status lucida_sort_u32_device(lucida_ctx* ctx, uint32_t* keys, uint32_t* vals, uint32_t* tmp_keys,
                                     uint32_t* tmp_vals, size_t n);

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): lucida_sort_u32_device with its histogram and scan
 * scratch in a fixed per-GPU array inside the library (8 MiB + 4 KiB, held
 * until the process ends) instead of the context's growing buffer. It calls
 * no allocator. n may be at most LUCIDA_SORT_NOALLOC_MAX_KEYS (2^24); larger
 * n is LUCIDA_INVALID_ARG. Queued on the context's stream like
 * lucida_sort_u32_device. Every call on one GPU shares that GPU's scratch:
 * calls that overlap on one GPU (other streams, other contexts, other host
 * threads) race and must be ordered by the caller. */
#define LUCIDA_SORT_NOALLOC_MAX_KEYS (1u << 24)
lucida_status lucida_sort_u32_device_noalloc(lucida_ctx* ctx, uint32_t* keys, uint32_t* vals, uint32_t* tmp_keys,
                                             uint32_t* tmp_vals, size_t n);

/* GpuBuildOutput (crates/gpu/src/lbvh.rs:724-733). Caller-owned host arrays of
 * the stated lengths (n = triangle count). */
typedef struct lucida_lbvh_parts {
    uint32_t* codes;         /* [n]  Morton code per triangle, input order */
    uint32_t* order;         /* [n]  triangle ids sorted by (code, id) */
    uint32_t* karras;        /* [2*(n-1)] (left, right) per internal node */
    uint32_t* parent;        /* [2n-1] parent per node, 0xFFFFFFFF at the root */
    LucidaAabb* node_bounds; /* [2n-1] fitted box per node (internal, then leaves) */
    uint32_t* subtree_size;  /* [n-1]  primitives under each internal node */
    LucidaAabb* scene_bounds;/* [1]    centroid bounds */
} lucida_lbvh_parts;

/* LbvhBuilder::build_parts (crates/gpu/src/lbvh.rs:369). n >= 1. */
lucida_status lucida_lbvh_build_parts(lucida_ctx* ctx, const GpuTriangle* tris, size_t n, const float* positions,
                                      size_t num_positions, lucida_lbvh_parts* out);

/* LbvhBuilder::build (crates/gpu/src/lbvh.rs:651) followed by Lbvh::from_parts
 * (crates/core/src/lbvh.rs:549): the relayout into GpuBvhNode runs on the GPU
 * here. nodes_out holds at least 2n-1 nodes; *num_nodes receives the count.
 * prim_indices_out holds n entries. Host arrays. */
lucida_status lucida_lbvh_build(lucida_ctx* ctx, const GpuTriangle* tris, size_t n, const float* positions,
                                size_t num_positions, GpuBvhNode* nodes_out, size_t* num_nodes,
                                uint32_t* prim_indices_out);

/* No counterpart in the reference, whose GPU path traces the LBVH as built.
 * The LBVH above, then its topology rebuilt for a lower surface-area cost
 * (treelet restructuring, Karras and Aila, HPG 2013): `rounds` bottom-up
 * passes, each of which replaces the treelet of up to LUCIDA_TREELET_LEAVES
 * leaves under every node of at least LUCIDA_TREELET_MIN_SIZE primitives with
 * the treelet topology of least cost, and subtrees of at most
 * LUCIDA_TREELET_MAX_LEAF primitives become leaves where that costs no more.
 * A node costs LUCIDA_TREELET_CI x its surface area plus its children's
 * costs, a leaf LUCIDA_TREELET_CT x area x primitives. Output layout as
 * lucida_lbvh_build's (children of an internal node adjacent, depth-first
 * left-first order, leaves index prim_indices_out); rounds 0 is taken as 1,
 * more than LUCIDA_TREELET_MAX_ROUNDS is LUCIDA_INVALID_ARG (the cost stops
 * falling after a few rounds). The closest hit along any ray is the one over
 * the LBVH, up to ties at equal distance. */
#define LUCIDA_TREELET_LEAVES 7
#define LUCIDA_TREELET_MIN_SIZE 7
#define LUCIDA_TREELET_MAX_LEAF 4
#define LUCIDA_TREELET_CI 1.2f
#define LUCIDA_TREELET_CT 1.0f
#define LUCIDA_TREELET_ROUNDS 3
#define LUCIDA_TREELET_MAX_ROUNDS 64
lucida_status lucida_bvh_build_treelet(lucida_ctx* ctx, const GpuTriangle* tris, size_t n, const float* positions,
                                       size_t num_positions, uint32_t rounds, GpuBvhNode* nodes_out,
                                       size_t* num_nodes, uint32_t* prim_indices_out);
/* The same build on device arrays, like lucida_lbvh_build_device: nodes
 * [2n-1], prim_indices [n] (optional), tris_sorted [n] (optional: the
 * triangles in leaf order), stage_us (optional, host) [3]: the LBVH up to its
 * fitted tree, the restructuring rounds, the layout, in microseconds. */
lucida_status lucida_bvh_build_treelet_device(lucida_ctx* ctx, const GpuTriangle* tris, size_t n,
                                              const float* positions, uint32_t rounds, GpuBvhNode* nodes,
                                              size_t* num_nodes, uint32_t* prim_indices, GpuTriangle* tris_sorted,
                                              float* stage_us);

/* Device-resident LBVH build: all arrays are device pointers. tris_sorted
 * receives the triangles permuted into leaf order (Bvh::compact_primitives,
 * crates/core/src/bvh.rs:517), so the leaves index triangles directly.
 * nodes holds 2n-1 entries; *num_nodes (host) receives the count. Stage
 * times in microseconds go to stage_us[0..5) when non-NULL:
 * prepare, sort, hierarchy, fit, relayout. */
lucida_status lucida_lbvh_build_device(lucida_ctx* ctx, const GpuTriangle* tris, size_t n, const float* positions,
                                       GpuBvhNode* nodes, size_t* num_nodes, uint32_t* prim_indices,
                                       GpuTriangle* tris_sorted, float* stage_us);

/* ------------------------------------------------------------------------ */
/* Scenes, films and the wavefront render (crates/gpu/src/wavefront.rs)    */
/* ------------------------------------------------------------------------ */

/* A scene as the reference uploads it (SceneBlob, gpu_layout.rs:706, plus the
 * environment map of envmap.rs). Host arrays, copied at creation. Instances
 * follow the primitives in one array exactly as SceneBlob::primitives_upload
 * (gpu_layout.rs:743) does; num_primitives counts the analytic ones only.
 * env_pixels is RGBA32F (w*h), env_conditional (w+1)*h, env_marginal h+1. */
typedef struct lucida_scene_desc {
    const GpuMaterial* materials;     size_t num_materials;
    const GpuPrimitive* primitives;   size_t num_primitives;
    const GpuInstance* instances;     size_t num_instances;
    const float* positions;           size_t num_positions;   /* vec4 each */
    const GpuVertexAttr* vertex_attrs;
    const GpuTriangle* triangles;     size_t num_triangles;
    const GpuBvhNode* bvh_nodes;      size_t num_bvh_nodes;   /* may be 0 */
    const GpuLight* lights;           size_t num_lights;
    uint32_t tlas_root;
    uint32_t env_width, env_height;
    const float* env_pixels;
    const float* env_conditional;
    const float* env_marginal;
    float env_total_weight;
    /* When nonzero and num_bvh_nodes == 0 with triangles present, the scene's
     * BVH is built on the GPU and the triangles are permuted into leaf order on
     * the device: 1 the LBVH (lucida_lbvh_build's tree), 2 the LBVH
     * restructured by treelets (lucida_bvh_build_treelet's tree,
     * LUCIDA_TREELET_ROUNDS rounds). */
    uint32_t build_lbvh;
} lucida_scene_desc;

lucida_status lucida_scene_create(lucida_ctx* ctx, const lucida_scene_desc* desc, lucida_scene** out);
void lucida_scene_destroy(lucida_scene* scene);

/* One mesh of an instanced scene: a range of desc->triangles. */
typedef struct lucida_mesh_range {
    uint32_t first_triangle;
    uint32_t triangle_count;
} lucida_mesh_range;

/* One placement (Instance::new, instance.rs:77): the forward object-to-world
 * transform as a row-major 3x4 [R | t], the mesh it places, and a material
 * override (0xFFFFFFFF keeps the triangles' own materials). */
typedef struct lucida_instance_placement {
    float object_to_world[12];
    uint32_t mesh;
    uint32_t material;
} lucida_instance_placement;

/* Scene::build_instances (crates/core/src/scene.rs:252), on the device: one
 * LBVH per mesh range (the BLASes, their triangles permuted into leaf order),
 * an LBVH over the instances' padded world bounds (the TLAS, built over two-
 * point "triangles" spanning each box), the node arrays rebased and
 * concatenated BLAS first, TLAS last (tlas_root = its first node), and the
 * instances appended to the primitives in TLAS leaf order. `desc` carries the
 * shared geometry and must have no BVH nodes and no instances of its own.
 * With desc->build_lbvh == 2 both levels are restructured by treelets
 * (lucida_bvh_build_treelet). The reference builds both levels with binned
 * SAH on the host; the tree differs, the hits do not. */
lucida_status lucida_scene_create_instanced(lucida_ctx* ctx, const lucida_scene_desc* desc,
                                            const lucida_mesh_range* meshes, size_t num_meshes,
                                            const lucida_instance_placement* placements, size_t num_placements,
                                            lucida_scene** out);
/* Counts and the device's instance records (GpuInstance, in TLAS leaf order)
 * and TLAS root of a scene, so a host tracer can walk the same structure. */
lucida_status lucida_scene_counts(lucida_scene* scene, size_t* num_triangles, size_t* num_bvh_nodes,
                                  size_t* num_instances, uint32_t* tlas_root);
lucida_status lucida_scene_download_instances(lucida_scene* scene, GpuInstance* out);
/* Number of BVH nodes the scene holds (after an on-device build). */
lucida_status lucida_scene_num_bvh_nodes(lucida_scene* scene, size_t* out);
/* Copies the scene's device BVH nodes and triangles back (host arrays sized
 * by lucida_scene_num_bvh_nodes and the triangle count). */
lucida_status lucida_scene_download_bvh(lucida_scene* scene, GpuBvhNode* nodes, GpuTriangle* triangles);

/* Batch ray queries against a resident scene without shading: the closest
 * hit of Scene::intersect (scene.rs:452) or the occlusion test of
 * Scene::occluded (scene.rs:577), through the same traversal EXTEND and
 * CONNECT use (flags below). rays: n x 8 floats (origin.xyz, unused, direction.xyz, t_max),
 * t_min is the renderer's T_MIN. hits: n x 4 (t, u, v, id bits), id = triangle
 * (leaf order), 0x80000000 | analytic primitive, or 0xFFFFFFFF; with any_hit
 * id is 0 for an occluded ray. instances (optional): the instance primitive
 * index of a triangle hit in an instanced scene. Host arrays. */
lucida_status lucida_scene_cast(lucida_ctx* ctx, const lucida_scene* scene, const float* rays, size_t n, int flags,
                                float* hits, uint32_t* instances);
/* flags for lucida_scene_cast (any other bit is LUCIDA_INVALID_ARG): */
#define LUCIDA_CAST_ANY_HIT 1 /* occlusion query */

/* A device-resident accumulation buffer (width*height GpuAccum). */
lucida_status lucida_film_create(lucida_ctx* ctx, uint32_t width, uint32_t height, lucida_film** out);
void lucida_film_destroy(lucida_film* film);
lucida_status lucida_film_clear(lucida_film* film);
lucida_status lucida_film_download(lucida_film* film, GpuAccum* host_out);
lucida_status lucida_film_upload(lucida_film* film, const GpuAccum* host_in);
/* Device address of the film's GpuAccum array. */
void* lucida_film_device_ptr(lucida_film* film);

/* Counters a render reports, all over the whole call. */
typedef struct lucida_render_stats {
    uint64_t extension_rays;        /* closest-hit rays traced (EXTEND)          */
    uint64_t shadow_rays;           /* any-hit rays traced (CONNECT)             */
    uint64_t node_visits;           /* BVH node visits, both ray kinds           */
    uint64_t triangle_tests;        /* triangle tests, both ray kinds            */
    uint64_t extend_node_visits;    /* of which closest-hit                      */
    uint64_t extend_triangle_tests; /* of which closest-hit                      */
    uint64_t stack_overflows;       /* pushes dropped for a full stack (0 expected) */
    float elapsed_ms;               /* device time of the whole render (HIP events) */
    /* Per-kernel device time and launch count when config->profile_kernels is
     * set, in the order GENERATE, EXTEND, SHADE, CONNECT, RESOLVE. */
    float kernel_ms[5];
    uint32_t kernel_launches[5];
    /* With count_traversal: 64 x the visit steps each wave executed, for
     * closest-hit and any-hit walks; visits / lane slots is the fraction of
     * lanes a walk kept busy (its SIMD efficiency). */
    uint64_t extend_lane_slots;
    uint64_t connect_lane_slots;
    /* With profile_kernels: device time and count of the extension-queue
     * sorts (see lucida_wavefront_config.ray_order). */
    float queue_sort_ms;
    uint32_t queue_sorts;
} lucida_render_stats;

/* Tunables of the wavefront. Zero fields take the defaults. */
typedef struct lucida_wavefront_config {
    uint32_t samples_per_batch; /* samples of every pixel in flight at once (0: about 2^23 paths) */
    uint32_t max_paths;         /* batched renders: paths per round at most (0: 2^24)            */
    uint32_t bvh_width;         /* 0 or 2: the binary tree, walked in the reference's order (any  *
                                 * other value is LUCIDA_INVALID_ARG)                              */
    uint32_t use_graph;         /* batched renders: capture every round's launches in one hipGraph */
    uint32_t count_traversal;   /* keep node/triangle counters (costs a little)                   */
    uint32_t profile_kernels;   /* time every kernel with HIP events (single renders)             */
    uint32_t ray_order;         /* 0: sort each bounce's extension and shadow queues by ray        *
                                 * direction and origin when the scene's BVH has at least 4096     *
                                 * nodes and a batch holds at least 2^21 paths; 1: never;          *
                                 * 2: whenever there is a BVH.                                     *
                                 * The order changes which lanes trace which rays together, not    *
                                 * what a ray hits, up to rounding: a ray meeting two triangles at *
                                 * distances equal to within rounding may keep either (the walk's  *
                                 * speculative steps depend on the rays beside it). Images were    *
                                 * bit-identical either way on every scene tested. Rounds of a     *
                                 * batched render that mix jobs sort by (job, ray), up to 256 jobs *
                                 * per round.                                                     */
    uint32_t serialize;         /* single renders, for measurement: one pipeline, and CONNECT and  *
                                 * the shadow sort on the main stream instead of beside the next  *
                                 * EXTEND, so per-kernel times (profile_kernels) are solo times.   *
                                 * Films were bit-identical either way on every scene tested.     */
} lucida_wavefront_config;

/* wavefront::render_native_on (crates/gpu/src/wavefront.rs:134): renders
 * `samples` samples per pixel, starting at uniforms->sample_offset, and ADDS
 * them to `film` (progressive, like the reference's Accum). uniforms->width /
 * height must match the film; samples_per_launch is set per batch internally.
 * stats may be NULL, config may be NULL (defaults). */
lucida_status lucida_render_wavefront(lucida_ctx* ctx, const lucida_scene* scene, const GpuUniforms* uniforms,
                                      uint32_t samples, lucida_film* film, const lucida_wavefront_config* config,
                                      lucida_render_stats* stats);

/* The megakernel render (pt_gpu::render_native_on, crates/gpu/src/lib.rs:391,
 * shaders/trace/megakernel.wgsl): the same render as lucida_render_wavefront,
 * with every path traced start to finish by one thread (one thread per
 * pixel, its samples in order). The films are equal; the comparison is the
 * architecture trade-off (README.md:675-714). With `render_with`
 * (lib.rs:69) Architecture::Megakernel maps here and ::Wavefront to
 * lucida_render_wavefront. Of config only samples_per_batch, bvh_width and
 * count_traversal apply; stats->kernel_ms[1] holds the device time. */
lucida_status lucida_render_megakernel(lucida_ctx* ctx, const lucida_scene* scene, const GpuUniforms* uniforms,
                                       uint32_t samples, lucida_film* film, const lucida_wavefront_config* config,
                                       lucida_render_stats* stats);

/* One render request of a batch: like one lucida_render_wavefront call. */
typedef struct lucida_job {
    const lucida_scene* scene;
    GpuUniforms uniforms; /* width/height must match the film; sample_offset, max_depth, seeds as usual */
    uint32_t samples;
    lucida_film* film;    /* distinct per job; the render ADDS into it */
} lucida_job;

/* Many renders of uneven resolution, samples and scene in one wavefront.
 * The reference renders one request at a time (wavefront.rs:163-212 batches
 * only the samples of one render); this entry point has no counterpart.
 * Each job is cut into the chunks its single render would use, so its film
 * is bit-identical to a lucida_render_wavefront of the same job. Jobs whose
 * chunks hold fewer than 2^21 paths go first, in mixed rounds that gather one
 * chunk from as many of them as fit in config->max_paths paths: all of a
 * round's paths share one launch of each stage, every kernel finding a path's
 * job (scene, uniforms, film) in a device table. The larger jobs follow,
 * shortest first, one chunk per round with the single render's kernels.
 * With config->use_graph the whole batch is captured once and replayed as one
 * hipGraph launch. job_done_ms (optional, n floats) receives each job's
 * completion time from the start of the batch (device clock). stats as for a
 * single render; kernel_launches[0] = rounds. */
lucida_status lucida_render_batch(lucida_ctx* ctx, const lucida_job* jobs, size_t n,
                                  const lucida_wavefront_config* config, float* job_done_ms,
                                  lucida_render_stats* stats);


/* The GGX directional-albedo table the BSDF's energy compensation reads
 * (energy::compute_table, crates/core/src/bsdf.rs:1200-1227), built on the
 * device at first use: 32 x 32 floats, row = roughness, column = sqrt(cos). */
lucida_status lucida_energy_table(lucida_ctx* ctx, float* out_1024);

/* ConvergenceMeter::record + read (crates/gpu/src/stats.rs:104, :153;
 * shaders/stats/convergence.wgsl): the mean over the film's pixels of the
 * relative standard error of each pixel's mean, from its accumulated sum and
 * sum of squares, sqrt(luminance(radiance_sq/N - mean^2) / N) /
 * luminance(mean); pixels with N < 2 or a mean luminance <= 1e-3 are left
 * out, and *out is -1 when none qualifies. Assumes independent samples, so
 * with the Sobol sampler it over-reports (a conservative figure). Waits for
 * the result. */
lucida_status lucida_film_convergence(lucida_film* film, float* out);

/* ------------------------------------------------------------------------ */
/* À-trous denoiser (crates/gpu/src/denoise.rs)                             */
/* ------------------------------------------------------------------------ */

/* DenoiseParams (crates/core/src/denoise.rs:62-94). */
typedef struct lucida_denoise_params {
    uint32_t iterations;
    float sigma_normal;
    float sigma_depth;
    float sigma_luminance;
    float sigma_albedo;
} lucida_denoise_params;

/* Denoiser::run (crates/gpu/src/denoise.rs:301): filters `in` into `out`
 * (Accum-shaped: radiance holds the filtered mean times out->samples, the
 * guide channels are copied). in and out are distinct films of one size. */
lucida_status lucida_denoise_atrous(lucida_ctx* ctx, const lucida_film* in, const lucida_denoise_params* params,
                                    lucida_film* out);

/* ---- Learned denoiser --------------------------------------------------------
 * No counterpart in the reference (its only denoiser is the a-trous filter,
 * crates/core/src/denoise.rs); it consumes the reference's Accum channels
 * (gpu_layout.rs:345). A 4-level encoder-decoder of 3x3 convolutions
 * (32/48/64/96 channels, skip concatenations, ReLU), BF16 weights and
 * activations with FP32 accumulation, trained on the device with Adam.
 * Parameters are fp32, per layer W[9][cin][cout] (tap = 3*dy + dx, dy/dx in
 * 0..2 for offsets -1..1) followed by b[cout], layers in the order
 *   16->32 32->32 | 32->48 48->48 | 48->64 64->64 | 64->96 96->96 |
 *   160->64 64->64 | 112->48 48->48 | 80->32 32->32 | 32->16 (no ReLU).
 * Crops and films must have a width that is a multiple of 128 and a height
 * that is a multiple of 8. */
typedef struct lucida_net lucida_net;

lucida_status lucida_net_create(lucida_ctx* ctx, uint32_t seed, lucida_net** out);
void lucida_net_destroy(lucida_net* net);
size_t lucida_net_param_count(void);
/* which = 0: parameters, 1: gradients of the last forward_backward. */
lucida_status lucida_net_get_params(lucida_net* net, float* out, int which);
lucida_status lucida_net_set_params(lucida_net* net, const float* in);
/* Loss = mean squared error in log1p(radiance / albedo) space against refs,
 * over `count` crops (x0[k], y0[k], w, h) of inputs[k] / refs[k] (x0, y0 may
 * be NULL for crops at the origin). Leaves the gradients in the net. */
lucida_status lucida_net_forward_backward(lucida_net* net, const lucida_film* const* inputs,
                                          const lucida_film* const* refs, uint32_t count, const uint32_t* x0,
                                          const uint32_t* y0, uint32_t w, uint32_t h, float* loss);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) with the held gradients. */
lucida_status lucida_net_adam(lucida_net* net, float lr);
lucida_status lucida_net_train_step(lucida_net* net, const lucida_film* const* inputs,
                                    const lucida_film* const* refs, uint32_t count, const uint32_t* x0,
                                    const uint32_t* y0, uint32_t w, uint32_t h, float lr, float* loss);
/* Denoises `in` into `out` (Accum-shaped, samples = 1, guides copied). */
lucida_status lucida_net_denoise(lucida_net* net, const lucida_film* in, lucida_film* out);
/* Test hook: tensor `index` (0 = input features, 1..15 = layer outputs) of
 * the last forward pass, or its gradient, as floats (count must match). */
lucida_status lucida_net_read_tensor(lucida_net* net, int index, int gradient, float* out, size_t count);

#ifdef __cplusplus
} /* extern "C" */
#endif

/* Layout contract: gpu_layout.rs:993-1166, restated. */
#ifdef __cplusplus
#define LUCIDA_STATIC_ASSERT static_assert
#else
#define LUCIDA_STATIC_ASSERT _Static_assert
#endif
LUCIDA_STATIC_ASSERT(sizeof(GpuMaterial) == 80, "GpuMaterial size");
LUCIDA_STATIC_ASSERT(offsetof(GpuMaterial, metallic) == 12, "GpuMaterial.metallic");
LUCIDA_STATIC_ASSERT(offsetof(GpuMaterial, emissive) == 16, "GpuMaterial.emissive");
LUCIDA_STATIC_ASSERT(offsetof(GpuMaterial, roughness) == 28, "GpuMaterial.roughness");
LUCIDA_STATIC_ASSERT(offsetof(GpuMaterial, ior) == 32, "GpuMaterial.ior");
LUCIDA_STATIC_ASSERT(offsetof(GpuMaterial, transmission) == 36, "GpuMaterial.transmission");
LUCIDA_STATIC_ASSERT(offsetof(GpuMaterial, eta) == 48, "GpuMaterial.eta");
LUCIDA_STATIC_ASSERT(offsetof(GpuMaterial, k) == 64, "GpuMaterial.k");
LUCIDA_STATIC_ASSERT(sizeof(GpuPrimitive) == 64, "GpuPrimitive size");
LUCIDA_STATIC_ASSERT(offsetof(GpuPrimitive, radius) == 12, "GpuPrimitive.radius");
LUCIDA_STATIC_ASSERT(offsetof(GpuPrimitive, edge_u) == 16, "GpuPrimitive.edge_u");
LUCIDA_STATIC_ASSERT(offsetof(GpuPrimitive, kind) == 28, "GpuPrimitive.kind");
LUCIDA_STATIC_ASSERT(offsetof(GpuPrimitive, edge_v) == 32, "GpuPrimitive.edge_v");
LUCIDA_STATIC_ASSERT(offsetof(GpuPrimitive, material) == 44, "GpuPrimitive.material");
LUCIDA_STATIC_ASSERT(offsetof(GpuPrimitive, normal) == 48, "GpuPrimitive.normal");
LUCIDA_STATIC_ASSERT(sizeof(GpuInstance) == sizeof(GpuPrimitive), "GpuInstance overlays GpuPrimitive");
LUCIDA_STATIC_ASSERT(offsetof(GpuInstance, blas_root) == 12, "GpuInstance.blas_root");
LUCIDA_STATIC_ASSERT(offsetof(GpuInstance, inv_row1) == 16, "GpuInstance.inv_row1");
LUCIDA_STATIC_ASSERT(offsetof(GpuInstance, kind) == offsetof(GpuPrimitive, kind), "GpuInstance.kind");
LUCIDA_STATIC_ASSERT(offsetof(GpuInstance, inv_row2) == 32, "GpuInstance.inv_row2");
LUCIDA_STATIC_ASSERT(offsetof(GpuInstance, material) == offsetof(GpuPrimitive, material), "GpuInstance.material");
LUCIDA_STATIC_ASSERT(offsetof(GpuInstance, inv_translation) == 48, "GpuInstance.inv_translation");
LUCIDA_STATIC_ASSERT(sizeof(GpuAccum) == 64, "GpuAccum size");
LUCIDA_STATIC_ASSERT(offsetof(GpuAccum, samples) == 12, "GpuAccum.samples");
LUCIDA_STATIC_ASSERT(offsetof(GpuAccum, albedo) == 16, "GpuAccum.albedo");
LUCIDA_STATIC_ASSERT(offsetof(GpuAccum, depth) == 28, "GpuAccum.depth");
LUCIDA_STATIC_ASSERT(offsetof(GpuAccum, normal) == 32, "GpuAccum.normal");
LUCIDA_STATIC_ASSERT(offsetof(GpuAccum, traversal) == 44, "GpuAccum.traversal");
LUCIDA_STATIC_ASSERT(offsetof(GpuAccum, radiance_sq) == 48, "GpuAccum.radiance_sq");
LUCIDA_STATIC_ASSERT(sizeof(GpuUniforms) == 160, "GpuUniforms size");
LUCIDA_STATIC_ASSERT(sizeof(GpuUniforms) % 16 == 0, "GpuUniforms std140 size");
LUCIDA_STATIC_ASSERT(offsetof(GpuUniforms, lens_radius) == 12, "GpuUniforms.lens_radius");
LUCIDA_STATIC_ASSERT(offsetof(GpuUniforms, cam_upper_left) == 16, "GpuUniforms.cam_upper_left");
LUCIDA_STATIC_ASSERT(offsetof(GpuUniforms, focus_distance) == 28, "GpuUniforms.focus_distance");
LUCIDA_STATIC_ASSERT(offsetof(GpuUniforms, cam_horizontal) == 32, "GpuUniforms.cam_horizontal");
LUCIDA_STATIC_ASSERT(offsetof(GpuUniforms, cam_vertical) == 48, "GpuUniforms.cam_vertical");
LUCIDA_STATIC_ASSERT(offsetof(GpuUniforms, width) == 64, "GpuUniforms.width");
LUCIDA_STATIC_ASSERT(offsetof(GpuUniforms, height) == 68, "GpuUniforms.height");
LUCIDA_STATIC_ASSERT(offsetof(GpuUniforms, sample_offset) == 72, "GpuUniforms.sample_offset");
LUCIDA_STATIC_ASSERT(offsetof(GpuUniforms, samples_per_launch) == 76, "GpuUniforms.samples_per_launch");
LUCIDA_STATIC_ASSERT(offsetof(GpuUniforms, max_depth) == 80, "GpuUniforms.max_depth");
LUCIDA_STATIC_ASSERT(offsetof(GpuUniforms, frame_seed) == 84, "GpuUniforms.frame_seed");
LUCIDA_STATIC_ASSERT(offsetof(GpuUniforms, num_primitives) == 88, "GpuUniforms.num_primitives");
LUCIDA_STATIC_ASSERT(offsetof(GpuUniforms, num_lights) == 92, "GpuUniforms.num_lights");
LUCIDA_STATIC_ASSERT(offsetof(GpuUniforms, background) == 96, "GpuUniforms.background");
LUCIDA_STATIC_ASSERT(offsetof(GpuUniforms, num_triangles) == 108, "GpuUniforms.num_triangles");
LUCIDA_STATIC_ASSERT(offsetof(GpuUniforms, num_bvh_nodes) == 112, "GpuUniforms.num_bvh_nodes");
LUCIDA_STATIC_ASSERT(offsetof(GpuUniforms, render_mode) == 116, "GpuUniforms.render_mode");
LUCIDA_STATIC_ASSERT(offsetof(GpuUniforms, sampling_mode) == 120, "GpuUniforms.sampling_mode");
LUCIDA_STATIC_ASSERT(offsetof(GpuUniforms, env_width) == 124, "GpuUniforms.env_width");
LUCIDA_STATIC_ASSERT(offsetof(GpuUniforms, env_height) == 128, "GpuUniforms.env_height");
LUCIDA_STATIC_ASSERT(offsetof(GpuUniforms, env_total_weight) == 132, "GpuUniforms.env_total_weight");
LUCIDA_STATIC_ASSERT(offsetof(GpuUniforms, sampler_kind) == 136, "GpuUniforms.sampler_kind");
LUCIDA_STATIC_ASSERT(offsetof(GpuUniforms, num_instances) == 140, "GpuUniforms.num_instances");
LUCIDA_STATIC_ASSERT(offsetof(GpuUniforms, tlas_root) == 144, "GpuUniforms.tlas_root");
LUCIDA_STATIC_ASSERT(sizeof(GpuLight) == 64, "GpuLight size");
LUCIDA_STATIC_ASSERT(offsetof(GpuLight, area) == 12, "GpuLight.area");
LUCIDA_STATIC_ASSERT(offsetof(GpuLight, edge_u) == 16, "GpuLight.edge_u");
LUCIDA_STATIC_ASSERT(offsetof(GpuLight, kind) == 28, "GpuLight.kind");
LUCIDA_STATIC_ASSERT(offsetof(GpuLight, edge_v) == 32, "GpuLight.edge_v");
LUCIDA_STATIC_ASSERT(offsetof(GpuLight, material) == 44, "GpuLight.material");
LUCIDA_STATIC_ASSERT(offsetof(GpuLight, normal) == 48, "GpuLight.normal");
LUCIDA_STATIC_ASSERT(sizeof(GpuPathState) == 112, "GpuPathState size");
LUCIDA_STATIC_ASSERT(offsetof(GpuPathState, pixel) == 12, "GpuPathState.pixel");
LUCIDA_STATIC_ASSERT(offsetof(GpuPathState, direction) == 16, "GpuPathState.direction");
LUCIDA_STATIC_ASSERT(offsetof(GpuPathState, rng_state) == 28, "GpuPathState.rng_state");
LUCIDA_STATIC_ASSERT(offsetof(GpuPathState, throughput) == 32, "GpuPathState.throughput");
LUCIDA_STATIC_ASSERT(offsetof(GpuPathState, prev_bsdf_pdf) == 44, "GpuPathState.prev_bsdf_pdf");
LUCIDA_STATIC_ASSERT(offsetof(GpuPathState, radiance) == 48, "GpuPathState.radiance");
LUCIDA_STATIC_ASSERT(offsetof(GpuPathState, depth) == 60, "GpuPathState.depth");
LUCIDA_STATIC_ASSERT(offsetof(GpuPathState, prev_position) == 64, "GpuPathState.prev_position");
LUCIDA_STATIC_ASSERT(offsetof(GpuPathState, guide_albedo) == 80, "GpuPathState.guide_albedo");
LUCIDA_STATIC_ASSERT(offsetof(GpuPathState, guide_depth) == 92, "GpuPathState.guide_depth");
LUCIDA_STATIC_ASSERT(offsetof(GpuPathState, guide_normal) == 96, "GpuPathState.guide_normal");
LUCIDA_STATIC_ASSERT(sizeof(GpuHitRecord) == 64, "GpuHitRecord size");
LUCIDA_STATIC_ASSERT(offsetof(GpuHitRecord, geometric_normal) == 32, "GpuHitRecord.geometric_normal");
LUCIDA_STATIC_ASSERT(offsetof(GpuHitRecord, light_area) == 44, "GpuHitRecord.light_area");
LUCIDA_STATIC_ASSERT(offsetof(GpuHitRecord, valid) == 52, "GpuHitRecord.valid");
LUCIDA_STATIC_ASSERT(sizeof(GpuShadowRay) == 48, "GpuShadowRay size");
LUCIDA_STATIC_ASSERT(offsetof(GpuShadowRay, path) == 28, "GpuShadowRay.path");
LUCIDA_STATIC_ASSERT(offsetof(GpuShadowRay, contribution) == 32, "GpuShadowRay.contribution");
LUCIDA_STATIC_ASSERT(offsetof(GpuDispatchArgs, shadow) == 12, "GpuDispatchArgs.shadow");
LUCIDA_STATIC_ASSERT(sizeof(GpuWavefrontCounters) == 16, "GpuWavefrontCounters size");
LUCIDA_STATIC_ASSERT(offsetof(GpuWavefrontCounters, shadow_count) == 12, "GpuWavefrontCounters.shadow_count");
LUCIDA_STATIC_ASSERT(sizeof(GpuVertexAttr) == 32, "GpuVertexAttr size");
LUCIDA_STATIC_ASSERT(offsetof(GpuVertexAttr, uv) == 16, "GpuVertexAttr.uv");
LUCIDA_STATIC_ASSERT(sizeof(GpuTriangle) == 16, "GpuTriangle size");
LUCIDA_STATIC_ASSERT(offsetof(GpuTriangle, material) == 12, "GpuTriangle.material");
LUCIDA_STATIC_ASSERT(sizeof(GpuBvhNode) == 32, "GpuBvhNode size");
LUCIDA_STATIC_ASSERT(offsetof(GpuBvhNode, left_first) == 12, "GpuBvhNode.left_first");
LUCIDA_STATIC_ASSERT(offsetof(GpuBvhNode, bounds_max) == 16, "GpuBvhNode.bounds_max");
LUCIDA_STATIC_ASSERT(offsetof(GpuBvhNode, count) == 28, "GpuBvhNode.count");
LUCIDA_STATIC_ASSERT(sizeof(LucidaAabb) == 32, "LucidaAabb size");

#endif /* LUCIDA_H */
