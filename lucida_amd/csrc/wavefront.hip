// Wavefront path tracer for gfx950.
//
// Reference: shaders/wavefront/{generate,extend,shade,connect,reset,resolve}.wgsl
// and the host loop of crates/gpu/src/wavefront.rs:152-676; the CPU twin is
// integrator.rs:209-699.
//
// Differences from the reference's WebGPU design, all MI355X-driven:
//  * Path state is kept as 16-byte planes (struct-of-arrays) instead of one
//    112-byte record, so each stage streams only the planes it needs: EXTEND
//    reads 32 bytes of a path, not 112.
//  * EXTEND is pure traversal and writes a 16-byte hit (t, u, v, primitive);
//    SHADE reconstructs the surface, so the traversal kernel stays small.
//  * Traversal carries a visited child's (first, count) with it, so each
//    node visit reads only the two 32-byte children (64 bytes), and the
//    stack lives in LDS, one column per lane.
//  * Queues are compacted with wave-level prefix sums: one atomic per wave
//    (ballot + popcount), not one per path.
//  * There is no indirect dispatch on HIP; kernels read their item counts
//    from device memory and the grid covers the batch's upper bound. Queue
//    and shadow counters have one slot per bounce, so no RESET pass is needed.
//  * The guide channels and radiance stay per path and RESOLVE sums a pixel's
//    samples in sample order, as the reference does, so the image is the
//    CPU twin's up to float reassociation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "context.hpp"
#include "shading.hpp"
#include "raystack.hpp"

namespace lucida {
namespace {

using namespace dev;

constexpr uint32_t MISS = 0xFFFFFFFFu;
constexpr uint32_t ANALYTIC = 0x80000000u;
constexpr int SHADE_WG = 256;   // SHADE: four waves, one queue atomic per block
constexpr int TRACE_WG = 128;   // traversal workgroups (two waves)
constexpr int STACK = 32;       // traversal stack entries per lane (shaders/common/bvh.wgsl:133)
constexpr int LDS_STACK = 16;   // the newest entries, in LDS; older ones spill to memory
constexpr int INST_STACK = 48;  // the two-level walk adds a RETURN marker and up to three waiting instances
// Traversal variants a kernel is compiled for.
constexpr int TRAV_BINARY = 0, TRAV_INSTANCED = 1;
constexpr int MAX_BOUNCES = 64;
// Counter words: queue length per bounce at [0, 72) (bounce b+1's count is
// written by bounce b's SHADE), shadow-ray count per bounce at [SHADOW_AT,
// +64), this batch's traversal counters (u64) at STATS_AT in STATS_STRIPES
// copies of 8 (a block adds to copy blockIdx % STATS_STRIPES, so the waves
// of a launch do not queue on a few addresses), the render's running totals
// (10 u64) at TOTALS_AT.
constexpr uint32_t SHADOW_AT = 80;
constexpr uint32_t STATS_AT = 160;
constexpr uint32_t STATS_STRIPES = 32;
constexpr uint32_t TOTALS_AT = STATS_AT + STATS_STRIPES * 16;
constexpr size_t COUNTER_BYTES = 4096;
static_assert((TOTALS_AT + 24) * 4 <= COUNTER_BYTES, "counter words exceed their buffer");

// Path and shadow-ray state has two layouts. Read in queue order (the queue
// a compacted, mostly ascending list of paths), one array per field keeps a
// wave's loads of a field contiguous. Read through sorted queues, in no
// particular order, records do better: a path's seven float4 fields and a
// pad in one 128-byte record, a shadow ray's three fields and a pad in 64
// bytes, so a lane's reads share one cache line. A Field addresses one field
// of every path (or shadow ray) in either layout: field[i] = base[i * stride].
constexpr uint32_t PATH_RECORD = 8;    // float4s per path record
constexpr uint32_t SHADOW_RECORD = 4;  // float4s per shadow record
struct Field {
    float4* base;
    uint32_t stride;  // 1: arrays of fields; else the record size in float4s
    __device__ float4& operator[](size_t i) const { return base[i * stride]; }
};
using PathField = Field;
using ShadowField = Field;

// Per-render device state: path records and queue/shadow planes of P entries.
struct Planes {
    PathField org;  // origin.xyz, depth (uint bits)
    PathField dir;  // direction.xyz, rng_state (uint bits)
    PathField thr;  // throughput.xyz, prev_bsdf_pdf
    PathField prv;  // prev_position.xyz, sampler_dim (uint bits)
    PathField rad;  // radiance.xyz, unused
    PathField gal;  // guide albedo.xyz, guide depth
    PathField gnr;  // guide normal.xyz, unused
    float4* hit;   // t, u, v, primitive id (uint bits)
    uint32_t* hin; // instance of a triangle hit (instanced scenes only)
    ShadowField sro;  // shadow: origin.xyz, t_max
    ShadowField srd;  // shadow: direction.xyz, path (uint bits)
    ShadowField src;  // shadow: contribution.xyz, unused
    uint32_t* q0;
    uint32_t* q1;
    uint32_t* counters;  // see SHADOW_AT / STATS_AT / TOTALS_AT
    uint32_t* job;       // job of each path (batched renders only)
    uint32_t* spill;     // traversal stack spill columns of EXTEND (INST_STACK - LDS_STACK per lane)
    uint32_t* spill2;    // the same for CONNECT, which may run beside the next EXTEND
    uint32_t* qkey;      // sort key of each extension-queue entry SHADE appends (ray_order_key)
    uint32_t* skey;      // sort key of each shadow ray SHADE appends, and
    uint32_t* sidx;      // its slot: after the sort, CONNECT's order over the shadow records
    uint32_t* qtmp;      // the extension sort's temporaries: entries and keys (after an
    uint32_t* ktmp;      // odd number of 8-bit passes, qtmp holds the sorted queue)
    uint32_t* qtmp2;     // the same for the shadow sort
    uint32_t* ktmp2;
    uint32_t* sort_scratch;  // the extension sort's digit histograms and scan partials
    uint32_t* sort_scratch2; // the shadow sort's (the two sorts run at once)
};

struct SceneView {
    const GpuMaterial* __restrict__ materials;
    const GpuPrimitive* __restrict__ prims;
    const float4* __restrict__ prim_w;  // a quad's w (quad_w), per analytic primitive
    const float4* __restrict__ pos;
    const GpuVertexAttr* __restrict__ attrs;
    const GpuTriangle* __restrict__ tris;
    const float4* __restrict__ edges;  // per triangle: p0, p1 - p0, p2 - p0
    const GpuBvhNode* __restrict__ nodes;
    const GpuInstance* __restrict__ insts;  // the primitive array read as instances (absolute indices)
    const GpuLight* __restrict__ lights;
    const float* __restrict__ ggx;
    EnvView env;
    uint32_t num_prims, num_tris, num_nodes, num_lights;
    uint32_t num_instances, tlas_root;
};

// One job of a batched render (lucida_render_batch): its scene, this round's
// uniforms (sample_offset and samples_per_launch of the round's chunk of its
// samples), its film, and where its paths and pixels start in the round.
struct BatchJob {
    SceneView sc;
    GpuUniforms u;
    GpuAccum* film;
    uint32_t path_base, pixel_base, pixels, _pad;
};

// The wave-uniform job of a lane: kernels of a batched render process the
// distinct jobs of a wave one at a time (a waterfall over readfirstlane), so
// each job's scene and uniforms are read once per wave through the scalar
// unit. Paths of one job are contiguous, so a wave almost always holds one.
__device__ inline uint32_t first_job(uint32_t jid) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)jid); }

__device__ inline uint64_t lanemask_lt() {
    const int lane = threadIdx.x & 63;
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}


// Queue append for a block of NW waves: each thread that wants a slot gets
// one, with one atomic per block on the queue's counter (a single address
// that every block of a launch hits; one atomic per wave made SHADE twice as
// slow). Slots are handed out in wave order, then lane order. Every thread
// of the block must call it; `lds` holds NW + 1 words.
template <int NW>
__device__ inline uint32_t block_append(bool want, uint32_t* counter, uint32_t* lds) {
    const uint64_t m = __ballot(want);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) lds[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t total = 0;
        for (int w = 0; w < NW; ++w) {
            const uint32_t c = lds[w];
            lds[w] = total;
            total += c;
        }
        lds[NW] = total ? atomicAdd(counter, total) : 0u;
    }
    __syncthreads();
    const uint32_t slot = lds[NW] + lds[wave] + (uint32_t)__popcll(m & lanemask_lt());
    __syncthreads();  // lds is reused by the next call
    return slot;
}

__device__ inline uint64_t wave_sum_u32(uint32_t v) {
    uint64_t s = v;
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    return s;
}

__device__ inline uint32_t wave_max_u32(uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o, 64));
    return v;
}

// Per-wave totals of the traversal counters, one atomic each. Overflows are
// always counted; visits and tests only when asked for.
// Per-batch counters live at word STATS_AT as u64: closest-hit visits and
// tests, any-hit visits and tests, stack overflows, then the lane slots of
// closest-hit and any-hit walks (64 x the wave's longest walk, the visit
// steps it executed, so visits / slots is the walk's SIMD efficiency).
// `slot` is 0 for EXTEND and 2 for CONNECT.
__device__ inline void flush_traversal_stats(uint32_t* counters, uint32_t visits, uint32_t tests, uint32_t overflow,
                                             uint32_t stats_on, int slot) {
    uint64_t* s64 = reinterpret_cast<uint64_t*>(counters + STATS_AT + (blockIdx.x % STATS_STRIPES) * 16);
    const uint64_t so = wave_sum_u32(overflow);
    if (stats_on) {
        const uint64_t sv = wave_sum_u32(visits), st = wave_sum_u32(tests);
        const uint64_t steps = wave_max_u32(visits);
        if ((threadIdx.x & 63) == 0) {
            if (sv) atomicAdd((unsigned long long*)&s64[slot], (unsigned long long)sv);
            if (st) atomicAdd((unsigned long long*)&s64[slot + 1], (unsigned long long)st);
            if (steps) atomicAdd((unsigned long long*)&s64[5 + slot / 2], (unsigned long long)(64 * steps));
        }
    }
    if ((threadIdx.x & 63) == 0 && so) atomicAdd((unsigned long long*)&s64[4], (unsigned long long)so);
}

// ---------------------------------------------------------------------------
// GENERATE (generate.wgsl:36-88): camera ray per (pixel, sample); identity
// queue; bounce-0 count.
// ---------------------------------------------------------------------------
// The pixels a batch covers: `count` consecutive pixels from `first` (rows of
// the image; a whole image is {0, width * height}). Path `local` of the batch
// is sample local / count of pixel first + local % count.
struct Band {
    uint32_t first, count;
};

// The primary ray of (pixel, sample) (camera.rs:84 generate_ray): the
// sampler is created and its pixel and lens draws taken.
__device__ inline void camera_ray(const GpuUniforms& u, uint32_t pixel, uint32_t sample, Sampler* s, V3* origin_out,
                                  V3* dir_out) {
    const uint32_t x = pixel % u.width, y = pixel / u.width;
    *s = make_sampler(u, pixel, sample);
    float pu, pv, lu, lv;
    s->next_vec2(&pu, &pv);
    s->next_vec2(&lu, &lv);
    float sx = ((float)x + pu) / (float)u.width;
    float ty = ((float)y + pv) / (float)u.height;
    V3 origin = ld3(u.cam_origin);
    V3 target = ld3(u.cam_upper_left) + sx * ld3(u.cam_horizontal) + ty * ld3(u.cam_vertical);
    if (u.lens_radius > 0.0f) {
        float dx, dy;
        concentric_disk(lu, lv, &dx, &dy);
        dx *= u.lens_radius;
        dy *= u.lens_radius;
        V3 right = normalize(ld3(u.cam_horizontal)), down = normalize(ld3(u.cam_vertical));
        origin = origin + (right * dx + down * dy);
    }
    *origin_out = origin;
    *dir_out = normalize(target - origin);
}

// Path `index` of the batch; `local` is its index within its own render
// (sample-major over the band: local = sample * band.count + pixel - band.first).
__device__ inline void generate_path(const GpuUniforms& u, uint32_t index, uint32_t local, Planes& p, Band band) {
    const uint32_t pixel = band.first + local % band.count;
    const uint32_t sample = u.sample_offset + local / band.count;
    Sampler s;
    V3 origin, d;
    camera_ray(u, pixel, sample, &s, &origin, &d);
    p.org[index] = make_float4(origin.x, origin.y, origin.z, __uint_as_float(0u));
    p.dir[index] = make_float4(d.x, d.y, d.z, __uint_as_float(s.state));
    p.thr[index] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
    p.prv[index] = make_float4(origin.x, origin.y, origin.z, __uint_as_float(s.dim));
    p.rad[index] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    p.gal[index] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    p.gnr[index] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    p.q0[index] = index;
}

__global__ __launch_bounds__(256) void k_generate(const GpuUniforms* __restrict__ U, Planes p, Band band) {
    const GpuUniforms u = *U;
    const uint32_t total = band.count * u.samples_per_launch;
    const uint32_t index = blockIdx.x * 256 + threadIdx.x;
    if (index == 0) p.counters[0] = total;
    if (index >= total) return;
    generate_path(u, index, index, p, band);
}

// Largest j < n with base(j) <= x (bases ascending, base(0) = 0).
template <class F>
__device__ inline uint32_t find_job(uint32_t n, uint32_t x, F base) {
    uint32_t lo = 0, hi = n - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) / 2;
        if (base(mid) <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// GENERATE for one round of a batched render: the round's paths are the
// jobs' chunks laid end to end, each in the single render's order.
__global__ __launch_bounds__(256) void k_generate_batch(const BatchJob* __restrict__ jobs, uint32_t njobs,
                                                        uint32_t total, Planes p) {
    const uint32_t index = blockIdx.x * 256 + threadIdx.x;
    if (index == 0) p.counters[0] = total;
    if (index >= total) return;
    const uint32_t jid = find_job(njobs, index, [&](uint32_t k) { return jobs[k].path_base; });
    for (;;) {
        const uint32_t j = first_job(jid);
        if (jid != j) continue;
        generate_path(jobs[j].u, index, index - jobs[j].path_base, p, Band{0u, jobs[j].pixels});
        p.job[index] = j;
        break;
    }
}

// ---------------------------------------------------------------------------
// Traversal (bvh.rs:1045 / bvh.wgsl:137): near-first, children tested as a
// pair, the far child pushed. A stack entry packs a child's (first, count) so
// a popped node is never re-read: (first << 5) | count, count <= 31 (leaf) or
// 0 (internal). An overflowing push is dropped, as the reference GPU does,
// and counted.
// ---------------------------------------------------------------------------

__device__ inline bool slab(const float4 lo, const float4 hi, V3 o, V3 inv, float t_min, float t_max, float* entry) {
    float t0x = (lo.x - o.x) * inv.x, t1x = (hi.x - o.x) * inv.x;
    float t0y = (lo.y - o.y) * inv.y, t1y = (hi.y - o.y) * inv.y;
    float t0z = (lo.z - o.z) * inv.z, t1z = (hi.z - o.z) * inv.z;
    float nx = fminf(t0x, t1x), ny = fminf(t0y, t1y), nz = fminf(t0z, t1z);
    float fx = fmaxf(t0x, t1x), fy = fmaxf(t0y, t1y), fz = fmaxf(t0z, t1z);
    float en = fmaxf(fmaxf(fmaxf(nx, ny), nz), t_min);
    float ex = fminf(fminf(fminf(fx, fy), fz), t_max);
    *entry = en;
    return en <= ex;
}

__device__ inline V3 safe_inv_dir(V3 d) {
    const float BIG = 1.0e16f, SMALL = 1.0e-16f;
    return V3{fabsf(d.x) > SMALL ? 1.0f / d.x : copysignf(BIG, d.x),
              fabsf(d.y) > SMALL ? 1.0f / d.y : copysignf(BIG, d.y),
              fabsf(d.z) > SMALL ? 1.0f / d.z : copysignf(BIG, d.z)};
}

// When every active lane tests the same triangle its 48 bytes are read once,
// through the scalar cache (the uniform fast path of load4 below).
__device__ inline void load_tri(const float4* edges, uint32_t ti, V3* p0, V3* e1, V3* e2) {
#if __HIP_DEVICE_COMPILE__
    const uint32_t t0 = __builtin_amdgcn_readfirstlane(ti);
    if (__builtin_amdgcn_ballot_w64(ti != t0) == 0) {
        typedef const __attribute__((address_space(4))) float4 ConstF4;
        const ConstF4* e = reinterpret_cast<const ConstF4*>(reinterpret_cast<uintptr_t>(edges + 3 * (size_t)t0));
        *p0 = f4xyz(e[0]);
        *e1 = f4xyz(e[1]);
        *e2 = f4xyz(e[2]);
        return;
    }
#endif
    const float4* e = edges + 3 * (size_t)ti;
    *p0 = f4xyz(e[0]);
    *e1 = f4xyz(e[1]);
    *e2 = f4xyz(e[2]);
}

// bvh.rs:930 intersect_triangle, in the twin's operation order. The edges
// e1 = p1 - p0 and e2 = p2 - p0 come precomputed (the same f32 subtractions,
// done once per triangle by k_tri_edges), so a test is three independent
// 16-byte loads instead of an index load followed by three gathers.
__device__ inline bool tri_test(const float4* __restrict__ edges, uint32_t ti, V3 o, V3 d, float t_min,
                                float t_max, float* t_out, float* u_out, float* v_out) {
    V3 p0, e1, e2;
    load_tri(edges, ti, &p0, &e1, &e2);
    V3 pv = cross(d, e2);
    float det = dot(e1, pv);
    if (fabsf(det) < 1.0e-12f) return false;
    float inv_det = 1.0f / det;
    V3 tv = o - p0;
    float u = dot(tv, pv) * inv_det;
    if (!(u >= 0.0f && u <= 1.0f)) return false;
    V3 qv = cross(tv, e1);
    float v = dot(d, qv) * inv_det;
    if (v < 0.0f || u + v > 1.0f) return false;
    float t = dot(e2, qv) * inv_det;
    if (t < t_min || t > t_max) return false;
    *t_out = t;
    *u_out = u;
    *v_out = v;
    return true;
}

// Four consecutive float4s q[k .. k+3]. When every active lane of the wave
// reads the same ones (common once the queues are sorted) they are read once
// through the scalar cache into SGPRs (the offset made uniform with
// readfirstlane, the loads through a constant-address-space pointer) instead
// of by every lane through the vector memory path. Nodes, triangle edges and
// instances are written only by earlier launches, and every launch starts
// with the scalar cache invalidated, so these reads never see stale data.
__device__ inline void load4(const float4* q, uint32_t k, float4* a, float4* b, float4* c, float4* e) {
#if __HIP_DEVICE_COMPILE__
    const uint32_t k0 = __builtin_amdgcn_readfirstlane(k);
    if (__builtin_amdgcn_ballot_w64(k != k0) == 0) {
        typedef const __attribute__((address_space(4))) float4 ConstF4;
        const ConstF4* r = reinterpret_cast<const ConstF4*>(reinterpret_cast<uintptr_t>(q + k0));
        *a = r[0];
        *b = r[1];
        *c = r[2];
        *e = r[3];
        return;
    }
#endif
    const float4* r = q + k;
    *a = r[0];
    *b = r[1];
    *c = r[2];
    *e = r[3];
}

// The child pair of a node: nodes[first] and nodes[first + 1], 32 bytes each.
__device__ inline void load_pair(const GpuBvhNode* nodes, uint32_t first, float4* a0, float4* a1, float4* b0,
                                 float4* b1) {
    load4(reinterpret_cast<const float4*>(nodes), 2 * first, a0, a1, b0, b1);
}

// Slab tests of a node's child pair, the pair read as load_pair reads it. In
// the uniform case the tests take the planes straight from SGPRs; the empty
// asm markers (distinct per branch) keep the compiler from merging the two
// branches' tests into one on copies of the planes.
__device__ inline void test_pair(const GpuBvhNode* nodes, uint32_t first, V3 o, V3 inv, float t_min, float t_max,
                                 bool* ha, bool* hb, float* ea, float* eb, uint32_t* ca, uint32_t* cb) {
#if __HIP_DEVICE_COMPILE__
    const uint32_t k0 = __builtin_amdgcn_readfirstlane(first);
    if (__builtin_amdgcn_ballot_w64(first != k0) == 0) {
        typedef const __attribute__((address_space(4))) float4 ConstF4;
        const ConstF4* r = reinterpret_cast<const ConstF4*>(reinterpret_cast<uintptr_t>(nodes + k0));
        const float4 a0 = r[0], a1 = r[1], b0 = r[2], b1 = r[3];
        *ha = slab(a0, a1, o, inv, t_min, t_max, ea);
        *hb = slab(b0, b1, o, inv, t_min, t_max, eb);
        *ca = (__float_as_uint(a0.w) << 5) | __float_as_uint(a1.w);
        *cb = (__float_as_uint(b0.w) << 5) | __float_as_uint(b1.w);
        asm volatile("; uniform pair");
        return;
    }
#endif
    const float4* r = reinterpret_cast<const float4*>(nodes + first);
    const float4 a0 = r[0], a1 = r[1], b0 = r[2], b1 = r[3];
    *ha = slab(a0, a1, o, inv, t_min, t_max, ea);
    *hb = slab(b0, b1, o, inv, t_min, t_max, eb);
    *ca = (__float_as_uint(a0.w) << 5) | __float_as_uint(a1.w);
    *cb = (__float_as_uint(b0.w) << 5) | __float_as_uint(b1.w);
#if __HIP_DEVICE_COMPILE__
    asm volatile("; per-lane pair");
#endif
}

// A child carried as (first << 5) | count: count <= 31 (0: internal), first < 2^27.
constexpr uint32_t COUNT_BITS = 5, COUNT_MASK = 31u;
__device__ inline uint32_t pack_child(uint32_t first, uint32_t count) { return (first << COUNT_BITS) | count; }

using RayStack = RayStackT<LDS_STACK, TRACE_WG>;

__device__ inline RayStack make_stack(uint32_t* lds_base, uint32_t* spill_base, int cap) {
    const uint32_t lanes = gridDim.x * TRACE_WG;
    return RayStack{lds_base + threadIdx.x, spill_base + blockIdx.x * TRACE_WG + threadIdx.x, lanes, 0, 0, cap, 0u,
                    false};
}

// Closest hit (any_hit = false) or occlusion (any_hit = true) against the
// triangles of the tree rooted at `root` (intersect_bvh_from, bvh.wgsl:137:
// the root's own box is not tested). Returns the hit flag; closest-hit fills
// t/u/v/tri. Entries the walk pushes sit above the stack's current top.
//
// Children are taken nearest first, as the reference takes them, and the
// walk is bound by how many steps a wave executes over its diverged lanes
// (vector-instruction issue), not by memory. Closest hit runs as two loops,
// inner nodes then leaves, and a lane that reaches a leaf postpones it and
// keeps walking inner nodes while some other lane of its wave is still
// looking for a leaf (speculative traversal, Aila and Laine, "Understanding
// the Efficiency of Ray Traversal on GPUs", HPG 2009): the wave's leaf steps
// then serve most of its lanes at once. The leaf step does not wait for the
// last four searching lanes (the lanes holding leaves would otherwise keep
// stepping speculatively through those few lanes' long searches: EXTEND
// 24.8 -> 22.6 ms per 16-spp step at 1024^2, that film unchanged). Leaves
// are still intersected in the walk's order, so the hit is the one the
// reference's walk returns, up to rounding: the inner nodes visited past a
// postponed leaf are tested against the closest hit before that leaf's
// triangles, so the walk may visit nodes the reference's walk culls (never
// fewer), and a triangle there whose hit ties the closest one to within
// rounding may then be kept. Occlusion rays mostly end at their
// first hit leaf, where one loop that takes either kind of visit per step
// measured faster. A tree visits each node at most once per ray; the cap
// only guarantees termination on a malformed (cyclic) node array.
template <bool ANY_HIT>
__device__ inline bool traverse(const SceneView& sc, RayStack& st, uint32_t root_index, V3 o, V3 d, float t_min,
                                float t_max, float* t_out, float* u_out, float* v_out, uint32_t* tri_out,
                                uint32_t* visits, uint32_t* tests, uint32_t* overflow) {
    bool found = false;
    float closest = t_max;
    if (sc.num_nodes == 0) {
        for (uint32_t i = 0; i < sc.num_tris; ++i) {
            float t, u, v;
            if (tri_test(sc.edges, i, o, d, t_min, closest, &t, &u, &v)) {
                if (ANY_HIT) return true;
                closest = t;
                *t_out = t; *u_out = u; *v_out = v; *tri_out = i;
                found = true;
            }
        }
        return found;
    }
    const V3 inv = safe_inv_dir(d);
    const GpuBvhNode root = sc.nodes[root_index];
    uint32_t cur = pack_child(root.left_first, root.count);
    const int base = st.sp;
    const uint32_t cap = 2u * sc.num_nodes + 8u;
    uint32_t iter = 0;
    if (ANY_HIT) {
        // one visit, of either kind, per step
        for (;;) {
            *visits += 1;
            const uint32_t first = cur >> COUNT_BITS, count = cur & COUNT_MASK;
            if (count > 0) {
                for (uint32_t k = 0; k < count; ++k) {
                    *tests += 1;
                    float t, u, v;
                    if (tri_test(sc.edges, first + k, o, d, t_min, closest, &t, &u, &v)) {
                        st.truncate(base);
                        return true;
                    }
                }
            } else {
                float4 a0, a1, b0, b1;
                load_pair(sc.nodes, first, &a0, &a1, &b0, &b1);
                float ea, eb;
                const bool ha = slab(a0, a1, o, inv, t_min, closest, &ea);
                const bool hb = slab(b0, b1, o, inv, t_min, closest, &eb);
                const uint32_t ca = pack_child(__float_as_uint(a0.w), __float_as_uint(a1.w));
                const uint32_t cb = pack_child(__float_as_uint(b0.w), __float_as_uint(b1.w));
                if (ha && hb) {
                    const bool a_near = ea <= eb;
                    if (!st.push(a_near ? cb : ca)) *overflow += 1;
                    cur = a_near ? ca : cb;
                    if (++iter >= cap) break;
                    continue;
                } else if (ha || hb) {
                    cur = ha ? ca : cb;
                    if (++iter >= cap) break;
                    continue;
                }
            }
            if (st.sp == base) break;
            cur = st.pop();
            if (++iter >= cap) break;
        }
        st.truncate(base);
        return false;
    }
    constexpr uint32_t NONE = 0xFFFFFFFFu;  // never a packed child: node indices are below 2^26
    constexpr int LEAF_WAIT = 4;            // searching lanes the leaf step does not wait for
    uint32_t leaf = NONE;                   // the postponed leaf
    for (;;) {
        // inner nodes, until every lane that still walks holds a leaf
        for (;;) {
            if (cur != NONE && (cur & COUNT_MASK) != 0 && leaf == NONE) {
                leaf = cur;
                cur = st.sp == base ? NONE : st.pop();
            }
            const bool inner = cur != NONE && (cur & COUNT_MASK) == 0;
            // To the leaf step once no lane steps an inner node, no lane still
            // looks for a leaf, or at most LEAF_WAIT lanes do while some lane
            // holds one. Each pass of the outer loop therefore intersects a
            // leaf or steps an inner node, so the walk always ends; with no
            // leaf held, the searching lanes step on however few they are.
            const uint64_t searching = __ballot(inner && leaf == NONE);
            if (__ballot(inner) == 0 || searching == 0 ||
                (__popcll(searching) <= LEAF_WAIT && __ballot(leaf != NONE) != 0))
                break;
            if (inner) {
                *visits += 1;
                bool ha, hb;
                float ea, eb;
                uint32_t ca, cb;
                test_pair(sc.nodes, cur >> COUNT_BITS, o, inv, t_min, closest, &ha, &hb, &ea, &eb, &ca, &cb);
                if (ha && hb) {
                    const bool a_near = ea <= eb;
                    if (!st.push(a_near ? cb : ca)) *overflow += 1;
                    cur = a_near ? ca : cb;
                } else if (ha || hb) {
                    cur = ha ? ca : cb;
                } else {
                    cur = st.sp == base ? NONE : st.pop();
                }
                if (++iter >= cap) {
                    cur = NONE;
                    leaf = NONE;
                }
            }
        }
        // the postponed leaves
        if (leaf != NONE) {
            *visits += 1;
            const uint32_t first = leaf >> COUNT_BITS, count = leaf & COUNT_MASK;
            for (uint32_t k = 0; k < count; ++k) {
                const uint32_t ti = first + k;
                *tests += 1;
                float t, u, v;
                if (tri_test(sc.edges, ti, o, d, t_min, closest, &t, &u, &v)) {
                    found = true;
                    closest = t;
                    *t_out = t; *u_out = u; *v_out = v; *tri_out = ti;
                }
            }
            leaf = NONE;
            if (++iter >= cap) cur = NONE;
        }
        if (__ballot(cur != NONE) == 0) break;
    }
    st.truncate(base);
    return found;
}

// An instance's world-to-object transform (GpuInstance, gpu_layout.rs:496).
struct InstanceXf {
    V3 r0, r1, r2, t;
    uint32_t blas_root, material;
};

__device__ inline InstanceXf load_instance(const GpuInstance* __restrict__ insts, uint32_t i) {
    const float4* q = reinterpret_cast<const float4*>(insts + i);
    const float4 a = q[0], b = q[1], c = q[2], e = q[3];
    return InstanceXf{V3{a.x, a.y, a.z}, V3{b.x, b.y, b.z}, V3{c.x, c.y, c.z}, V3{e.x, e.y, e.z},
                      __float_as_uint(a.w), __float_as_uint(c.w)};
}

// Instance::transform_ray (instance.rs:120, bvh.wgsl instance_transform_*):
// the direction is not normalised, so t means the same in both spaces.
__device__ inline V3 xf_point(const InstanceXf& I, V3 p) {
    return V3{dot(I.r0, p) + I.t.x, dot(I.r1, p) + I.t.y, dot(I.r2, p) + I.t.z};
}
__device__ inline V3 xf_vector(const InstanceXf& I, V3 v) { return V3{dot(I.r0, v), dot(I.r1, v), dot(I.r2, v)}; }
// Instance::transform_normal (instance.rs:138): by the transpose of the stored inverse.
__device__ inline V3 xf_normal(const InstanceXf& I, V3 n) {
    return normalize_or(V3{I.r0.x * n.x + I.r1.x * n.y + I.r2.x * n.z, I.r0.y * n.x + I.r1.y * n.y + I.r2.y * n.z,
                           I.r0.z * n.x + I.r1.z * n.y + I.r2.z * n.z},
                        v3(0, 0, 1));
}

// Two-level traversal (intersect_instances, bvh.wgsl:305): the TLAS rooted at
// sc.tlas_root in world space; at each instance leaf the ray goes to object
// space and walks that instance's BLAS with `closest` carried across levels.
// With ANY_HIT it is the two-level occlusion query (the reference's shadow
// rays walk only the tree at node 0, see connect.wgsl:192-212; here they see
// every instance).
//
// One loop for both levels: a lane is either in the TLAS (world ray) or in a
// BLAS (object-space ray). Entering an instance pushes a RETURN marker below
// the BLAS's entries and the leaf's remaining instances as tagged entries
// below that, so popping the marker restores the world ray and popping a
// tagged entry enters the next instance. Nodes, instances and triangles are
// visited in the order of a BLAS walk nested in the TLAS walk (so the hits
// are the same), but a lane never waits in the TLAS loop for another lane's
// whole BLAS walk. Each level's visits are capped as a nested walk's would be.
template <bool ANY_HIT>
__device__ inline bool traverse_instances(const SceneView& sc, RayStack& st, V3 o_w, V3 d_w, float t_min,
                                          float t_max, float* t_out, float* u_out, float* v_out, uint32_t* tri_out,
                                          uint32_t* inst_out, uint32_t* visits, uint32_t* tests,
                                          uint32_t* overflow) {
    constexpr uint32_t RETURN = 0xFFFFFFFFu;  // never a packed child: node indices are below 2^26
    constexpr uint32_t INST = 0x80000000u;    // tag of a waiting instance (primitive index below 2^26)
    bool found = false;
    float closest = t_max;
    const V3 inv_w = safe_inv_dir(d_w);
    V3 o = o_w, d = d_w, inv = inv_w;
    const GpuBvhNode root = sc.nodes[sc.tlas_root];
    uint32_t cur = pack_child(root.left_first, root.count);
    bool in_blas = false;
    uint32_t inst = 0, tlas_iter = 0, blas_iter = 0;
    const uint32_t cap = 2u * sc.num_nodes + 8u;
    const int base = st.sp;
    auto enter = [&](uint32_t ii) {
        if (!st.push(RETURN)) *overflow += 1;
        const InstanceXf I = load_instance(sc.insts, ii);
        o = xf_point(I, o_w);
        d = xf_vector(I, d_w);
        inv = safe_inv_dir(d);
        const GpuBvhNode br = sc.nodes[I.blas_root];
        cur = pack_child(br.left_first, br.count);
        in_blas = true;
        inst = ii;
        blas_iter = 0;
    };
    // Pops the next node into cur, crossing RETURN markers and entering
    // waiting instances; false when the walk is over.
    auto pop_next = [&]() -> bool {
        for (;;) {
            if (st.sp == base) return false;
            const uint32_t e = st.pop();
            if (e == RETURN) {
                in_blas = false;
                o = o_w;
                d = d_w;
                inv = inv_w;
                continue;
            }
            if (e & INST) enter(e & ~INST);
            else cur = e;
            return ++(in_blas ? blas_iter : tlas_iter) < cap;
        }
    };
    // Visits the inner node cur; false when neither child is hit.
    auto inner = [&]() -> bool {
        float4 a0, a1, b0, b1;
        load_pair(sc.nodes, cur >> COUNT_BITS, &a0, &a1, &b0, &b1);
        float ea, eb;
        const bool ha = slab(a0, a1, o, inv, t_min, closest, &ea);
        const bool hb = slab(b0, b1, o, inv, t_min, closest, &eb);
        const uint32_t ca = pack_child(__float_as_uint(a0.w), __float_as_uint(a1.w));
        const uint32_t cb = pack_child(__float_as_uint(b0.w), __float_as_uint(b1.w));
        if (ha && hb) {
            const bool a_near = ea <= eb;
            if (!st.push(a_near ? cb : ca)) *overflow += 1;
            cur = a_near ? ca : cb;
            return true;
        }
        if (ha || hb) {
            cur = ha ? ca : cb;
            return true;
        }
        return false;
    };
    // The leaf cur: a BLAS leaf's triangles, or a TLAS leaf's instances
    // (the first entered, the later ones waiting). True for an any-hit.
    auto leaf = [&]() -> bool {
        const uint32_t first = cur >> COUNT_BITS, count = cur & COUNT_MASK;
        if (in_blas) {
            for (uint32_t k = 0; k < count; ++k) {
                const uint32_t ti = first + k;
                *tests += 1;
                float t, u, v;
                if (tri_test(sc.edges, ti, o, d, t_min, closest, &t, &u, &v)) {
                    if (ANY_HIT) return true;
                    closest = t;
                    *t_out = t; *u_out = u; *v_out = v; *tri_out = ti; *inst_out = inst;
                    found = true;
                }
            }
            return false;
        }
        for (uint32_t k = count - 1; k >= 1; --k)
            if (!st.push(INST | (first + k))) *overflow += 1;
        enter(first);
        return false;
    };
    if (ANY_HIT) {
        // one visit, of either kind, per step
        for (;;) {
            *visits += 1;
            bool more;
            if ((cur & COUNT_MASK) == 0) {
                more = inner() ? ++(in_blas ? blas_iter : tlas_iter) < cap : pop_next();
            } else {
                const bool was_blas = in_blas;
                if (leaf()) {
                    st.truncate(base);
                    return true;
                }
                more = was_blas ? pop_next() : ++blas_iter < cap;
            }
            if (!more) break;
        }
    } else {
        // closest hit: inner nodes of either level until a leaf, then the leaf
        bool more = true;
        while (more) {
            while (more && (cur & COUNT_MASK) == 0) {
                *visits += 1;
                more = inner() ? ++(in_blas ? blas_iter : tlas_iter) < cap : pop_next();
            }
            if (!more) break;
            *visits += 1;
            const bool was_blas = in_blas;
            leaf();
            more = was_blas ? pop_next() : ++blas_iter < cap;
        }
    }
    st.truncate(base);
    return found;
}

// The closest of the analytic primitives (scene.wgsl: each tested in turn
// before the tree; on equal t the later one wins), into best_t / best. The
// primitives' padded boxes (k_prim_box, after their quad_w planes) are
// tested first: the box of them all, then that of each run of PRIM_CHUNK
// consecutive primitives. A ray that misses a box misses every primitive in
// it, so skipping them leaves the result as the full loop's.
constexpr uint32_t PRIM_CHUNK = 4;
__device__ inline void analytic_closest(const SceneView& sc, V3 o, V3 d, float* best_t, uint32_t* best) {
    const uint32_t n = sc.num_prims;
    if (n == 0) return;
    const V3 inv = safe_inv_dir(d);
    float entry;
    if (!slab(sc.prim_w[n], sc.prim_w[n + 1], o, inv, T_MIN, *best_t, &entry)) return;
    for (uint32_t g = 0; g * PRIM_CHUNK < n; ++g) {
        if (!slab(sc.prim_w[n + 2 + 2 * g], sc.prim_w[n + 3 + 2 * g], o, inv, T_MIN, *best_t, &entry)) continue;
        const uint32_t end = min(g * PRIM_CHUNK + PRIM_CHUNK, n);
        for (uint32_t k = g * PRIM_CHUNK; k < end; ++k) {
            const GpuPrimitive& pr = sc.prims[k];
            float t;
            const bool h = pr.kind == LUCIDA_PRIM_KIND_QUAD
                               ? intersect_quad(pr, f4xyz(sc.prim_w[k]), o, d, T_MIN, *best_t, &t)
                               : intersect_sphere(pr, o, d, T_MIN, *best_t, &t);
            if (h) {
                *best_t = t;
                *best = k;
            }
        }
    }
}

// Whether any analytic primitive blocks the segment (T_MIN, t_max)
// (light.rs:327 unoccluded; scene.rs:577), behind the same boxes.
__device__ inline bool analytic_any(const SceneView& sc, V3 o, V3 d, float t_max) {
    const uint32_t n = sc.num_prims;
    if (n == 0) return false;
    const V3 inv = safe_inv_dir(d);
    float entry;
    if (!slab(sc.prim_w[n], sc.prim_w[n + 1], o, inv, T_MIN, t_max, &entry)) return false;
    for (uint32_t g = 0; g * PRIM_CHUNK < n; ++g) {
        if (!slab(sc.prim_w[n + 2 + 2 * g], sc.prim_w[n + 3 + 2 * g], o, inv, T_MIN, t_max, &entry)) continue;
        const uint32_t end = min(g * PRIM_CHUNK + PRIM_CHUNK, n);
        for (uint32_t k = g * PRIM_CHUNK; k < end; ++k) {
            const GpuPrimitive& pr = sc.prims[k];
            float t;
            if (pr.kind == LUCIDA_PRIM_KIND_QUAD ? intersect_quad(pr, f4xyz(sc.prim_w[k]), o, d, T_MIN, t_max, &t)
                                                 : intersect_sphere(pr, o, d, T_MIN, t_max, &t))
                return true;
        }
    }
    return false;
}

// ---------------------------------------------------------------------------
// EXTEND (extend.wgsl:124, scene.rs:452): analytic primitives by brute force,
// then the BVH with t_max = the best analytic hit. Writes (t, u, v, id):
// id = triangle, ANALYTIC | primitive, or MISS.
// ---------------------------------------------------------------------------
// MODE (TRAV_*) selects the walk at compile time, so each kernel carries
// only its own traversal's registers.
template <bool BATCH, int MODE>
__global__ __launch_bounds__(TRACE_WG) void k_extend(SceneView sc0, const BatchJob* __restrict__ jobs, Planes p,
                                                     const uint32_t* __restrict__ queue,
                                                     const uint32_t* __restrict__ count_ptr, uint32_t stats_on) {
    __shared__ uint32_t stack_mem[LDS_STACK * TRACE_WG];
    const uint32_t i = blockIdx.x * TRACE_WG + threadIdx.x;
    const uint32_t count = *count_ptr;
    uint32_t visits = 0, tests = 0, overflow = 0;
    if (i < count) {
      RayStack st = make_stack(stack_mem, p.spill, MODE == TRAV_BINARY ? STACK : INST_STACK);
      const uint32_t pid = queue[i];
      const uint32_t jid = BATCH ? p.job[pid] : 0u;
      for (;;) {
        const uint32_t j = BATCH ? first_job(jid) : 0u;
        if (jid != j) continue;
        const SceneView& sc = BATCH ? jobs[j].sc : sc0;
        const float4 of = p.org[pid], df = p.dir[pid];
        const V3 o = f4xyz(of), d = f4xyz(df);
        float best_t = T_MAX;
        uint32_t best = MISS;
        analytic_closest(sc, o, d, &best_t, &best);
        float t = 0.0f, u = 0.0f, v = 0.0f;
        uint32_t tri = 0, inst = 0;
        // Instanced scenes take the two-level path instead of the
        // single-level one (scene.wgsl:31-75): their triangles live in object space.
        const bool two_level = MODE == TRAV_INSTANCED && sc.num_instances;
        bool th;
        if (two_level)
            th = traverse_instances<false>(sc, st, o, d, T_MIN, best_t, &t, &u, &v, &tri, &inst, &visits, &tests,
                                           &overflow);
        else
            th = traverse<false>(sc, st, 0, o, d, T_MIN, best_t, &t, &u, &v, &tri, &visits, &tests, &overflow);
        if (two_level) p.hin[i] = inst;
        float4 out;
        if (th) out = make_float4(t, u, v, __uint_as_float(tri));
        else if (best != MISS) out = make_float4(best_t, 0.0f, 0.0f, __uint_as_float(ANALYTIC | best));
        else out = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(MISS));
        p.hit[i] = out;
        break;
      }
    }
    flush_traversal_stats(p.counters, visits, tests, overflow, stats_on, 0);
}

// Sort key of an extension or shadow ray. Between bounces the queues are
// sorted by it, so a wave walks rays that leave nearby points in similar
// directions: they visit the same nodes in the same order, which keeps more
// of the wave's lanes busy (the walks are bound by the steps a wave executes
// over diverged lanes) and shares node fetches. Three bits per direction
// component over [-1, 1] and five per origin coordinate over the scene box,
// interleaved from the most significant bit (origin, then direction, at each
// level; the origin's two lowest levels last): 24 bits, three 8-bit radix
// passes. Measured on the 1.05M-triangle scene at 1024^2 x 16 spp: this key
// 259 Mpaths/s; direction before origin 255, 4 + 4 bits 253, 2 + 6 bits 245,
// 5 + 5 bits (four passes) 250, 16 bits (two passes) 236.
constexpr uint32_t RAY_KEY_BITS = 24;
constexpr int KEY_DIR_BITS = 3, KEY_ORG_BITS = 5;
__device__ inline uint32_t ray_order_key(V3 o, V3 d, const GpuBvhNode& box) {
    auto cell = [](float x, float n) { return (uint32_t)fminf(fmaxf(x, 0.0f), n - 1.0f); };
    constexpr float ND = (float)(1 << KEY_DIR_BITS), NO = (float)(1 << KEY_ORG_BITS);
    auto ocell = [&](float x, float lo, float hi) { return cell((x - lo) * (NO / fmaxf(hi - lo, 1.0e-30f)), NO); };
    const uint32_t dc[3] = {cell((d.x + 1.0f) * (0.5f * ND), ND), cell((d.y + 1.0f) * (0.5f * ND), ND),
                            cell((d.z + 1.0f) * (0.5f * ND), ND)};
    const uint32_t oc[3] = {ocell(o.x, box.bounds_min[0], box.bounds_max[0]),
                            ocell(o.y, box.bounds_min[1], box.bounds_max[1]),
                            ocell(o.z, box.bounds_min[2], box.bounds_max[2])};
    constexpr int L = KEY_DIR_BITS > KEY_ORG_BITS ? KEY_DIR_BITS : KEY_ORG_BITS;
    uint32_t k = 0;
#pragma unroll
    for (int b = L - 1; b >= 0; --b) {
        const int bd = b - (L - KEY_DIR_BITS), bo = b - (L - KEY_ORG_BITS);
#pragma unroll
        for (int a = 0; a < 3; ++a)
            if (bo >= 0) k = (k << 1) | ((oc[a] >> bo) & 1u);
#pragma unroll
        for (int a = 0; a < 3; ++a)
            if (bd >= 0) k = (k << 1) | ((dc[a] >> bd) & 1u);
    }
    return k;
}

// Queue sort key of a ray: in a single-scene render the ray's key; in a round
// of a batched render (job >= 0) the job above the key's top 16 bits, so a
// sorted round keeps each job's rays together (a wave of one job walks one
// tree) and still sorts in three passes.
__device__ inline uint32_t queue_key(const SceneView& sc, int job, V3 o, V3 d) {
    const uint32_t k = sc.num_nodes ? ray_order_key(o, d, sc.nodes[sc.num_instances ? sc.tlas_root : 0]) : 0u;
    return job < 0 ? k : ((uint32_t)job << 16) | (k >> (RAY_KEY_BITS - 16));
}

// ---------------------------------------------------------------------------
// SHADE (shade.wgsl:120-348 = integrator.rs radiance_with_guide, one bounce).
// ---------------------------------------------------------------------------
// A path between bounces, in registers: its ray, throughput, radiance so
// far, the previous vertex and BSDF pdf (for MIS on emission), its bounce
// depth and its sampler.
struct PathRegs {
    V3 origin, dir, throughput, radiance, prev_position;
    float prev_bsdf_pdf;
    uint32_t depth;
    Sampler s;
};
// A deferred shadow ray: origin, direction, t_max and the contribution it
// adds to its path's radiance when unoccluded.
struct ShadowRay {
    V3 o, d, c;
    float t_max;
};

// One bounce of radiance_with_guide for the ray's hit (t, u, v, id; `inst`
// is the hit's instance in an instanced scene): adds emission to
// P.radiance, emits at most one shadow ray (`shadow`, `sr`), writes the
// depth-0 guides (`guide`, `gal` = albedo and hit distance, `gnr` = normal)
// and samples the BSDF. Returns whether the path continues; P then holds the
// next ray.
__device__ __forceinline__ bool shade_bounce(const SceneView& sc, const GpuUniforms& u, const float4 hit, uint32_t inst,
                                    PathRegs& P, bool& shadow, ShadowRay& sr, bool& guide, float4& gal,
                                    float4& gnr) {
    shadow = false;
    guide = false;
    const V3 origin = P.origin, dir = P.dir;
    const uint32_t depth = P.depth;
    const uint32_t mode = u.sampling_mode > 2 ? 2 : u.sampling_mode;
    const uint32_t total = (uint32_t)sc.num_lights + (sc.env.present() ? 1u : 0u);
    const uint32_t id = __float_as_uint(hit.w);

    if (id == MISS) {
        if (sc.env.present()) {
            float w = 1.0f;
            if (depth > 0) {
                if (mode == LUCIDA_SAMPLING_NEE) w = 0.0f;
                else if (mode == LUCIDA_SAMPLING_MIS)
                    w = power_heuristic(P.prev_bsdf_pdf, env_pdf(sc.env, dir) / (float)max(total, 1u));
            }
            if (w > 0.0f) P.radiance = P.radiance + P.throughput * env_radiance(sc.env, dir) * w;
        } else {
            P.radiance = P.radiance + P.throughput * ld3(u.background);
        }
        return false;
    }
    // Surface at the hit: shade_triangle / shade_analytic (scene.rs:405-560).
    const float t = hit.x;
    V3 position = origin + t * dir;
    V3 normal, gnormal;
    bool front;
    uint32_t material;
    float light_area;
    if (id & ANALYTIC) {
        const GpuPrimitive& pr = sc.prims[id & ~ANALYTIC];
        const bool quad = pr.kind == LUCIDA_PRIM_KIND_QUAD;
        light_area = quad ? length(cross(ld3(pr.edge_u), ld3(pr.edge_v))) : 0.0f;
        V3 gn = quad ? ld3(pr.normal) : (position - ld3(pr.position)) / pr.radius;
        front = dot(gn, dir) < 0.0f;
        gnormal = front ? gn : -gn;
        normal = gnormal;
        material = pr.material;
    } else {
        const GpuTriangle tri = sc.tris[id];
        const V3 p0 = f4xyz(sc.pos[tri.i0]), p1 = f4xyz(sc.pos[tri.i1]), p2 = f4xyz(sc.pos[tri.i2]);
        const float w = 1.0f - hit.y - hit.z;
        V3 sh = normalize_or(w * ld3(sc.attrs[tri.i0].normal) + hit.y * ld3(sc.attrs[tri.i1].normal) +
                                 hit.z * ld3(sc.attrs[tri.i2].normal),
                             v3(0, 0, 1));
        V3 cr = cross(p1 - p0, p2 - p0);
        V3 g = normalize_or(cr, v3(0, 0, 1));
        material = tri.material;
        if (sc.num_instances) {
            // shade_instanced_triangle (scene.rs:375): orient in object
            // space, bring both normals to world space, re-decide the face
            // there, apply the material override; instanced triangles are
            // not lights.
            const InstanceXf I = load_instance(sc.insts, inst);
            const float oflip = dot(g, xf_vector(I, dir)) < 0.0f ? 1.0f : -1.0f;
            const V3 n_w = xf_normal(I, sh * oflip), g_w = xf_normal(I, g * oflip);
            front = dot(g_w, dir) < 0.0f;
            const float flip = front ? 1.0f : -1.0f;
            normal = n_w * flip;
            gnormal = g_w * flip;
            if (I.material != 0xFFFFFFFFu) material = I.material;
            light_area = 0.0f;
        } else {
            front = dot(g, dir) < 0.0f;
            const float flip = front ? 1.0f : -1.0f;
            normal = sh * flip;
            gnormal = g * flip;
            light_area = 0.5f * length(cr);
        }
    }
    const GpuMaterial m = sc.materials[material];
    const V3 emissive = ld3(m.emissive);
    if (front && max3(emissive) > 0.0f) {
        float w = 1.0f;
        if (depth > 0) {
            if (mode == LUCIDA_SAMPLING_NEE) w = 0.0f;
            else if (mode == LUCIDA_SAMPLING_MIS) {
                // light_pdf_from_geometry (light.rs:241) * n / total
                float pdf_light = 0.0f;
                const uint32_t n = sc.num_lights;
                if (n != 0 && light_area > 0.0f) {
                    V3 tl = position - P.prev_position;
                    float dist_sq = dot(tl, tl);
                    if (dist_sq > 0.0f) {
                        V3 wl = tl / sqrtf(dist_sq);
                        float cos_l = dot(gnormal, -wl);
                        if (cos_l > 1.0e-6f) pdf_light = (1.0f / ((float)n * light_area)) * dist_sq / cos_l;
                    }
                }
                pdf_light = pdf_light * ((float)n / (float)max(total, 1u));
                w = power_heuristic(P.prev_bsdf_pdf, pdf_light);
            }
        }
        if (w > 0.0f) P.radiance = P.radiance + P.throughput * emissive * w;
    }
    Surface surf = surface_from_material(m, front, sc.ggx);
    if (depth == 0) {
        V3 ga = vmax(surf.diffuse_albedo + surf.f0 + emissive, v3s(1.0e-3f));
        guide = true;
        gal = make_float4(ga.x, ga.y, ga.z, t);
        gnr = make_float4(normal.x, normal.y, normal.z, 0.0f);
    }
    const V3 wo = to_local(-dir, normal);
    if (!(wo.z > 0.0f)) return false;
    surf.bind(wo);
    Sampler& s = P.s;
    if (mode != LUCIDA_SAMPLING_BSDF && depth + 1 < u.max_depth && total > 0) {
        // direct_light (integrator.rs:418): the contribution is resolved
        // here; the shadow ray only answers visibility.
        const float u_select = s.next_f32();
        float ua_x, ua_y;
        s.next_vec2(&ua_x, &ua_y);
        const uint32_t n_lights = sc.num_lights;
        const uint32_t pick = min((uint32_t)(u_select * (float)total), total - 1u);
        V3 wi_world{0, 1, 0}, rin{0, 0, 0};
        float distance = 0.0f, pdf = 0.0f;
        bool valid = false;
        if (pick < n_lights) {
            const float u_light = ((float)pick + 0.5f) / (float)n_lights;
            const uint32_t li = min((uint32_t)(u_light * (float)n_lights), n_lights - 1u);
            const GpuLight L = sc.lights[li];
            V3 lo = ld3(L.origin), eu = ld3(L.edge_u), ev = ld3(L.edge_v);
            V3 pt;
            if (L.kind == LUCIDA_LIGHT_KIND_TRIANGLE) {
                float su = sqrtf(ua_x);
                pt = lo + (1.0f - su) * eu + (ua_y * su) * ev;
            } else {
                pt = lo + ua_x * eu + ua_y * ev;
            }
            V3 tl = pt - position;
            float dist_sq = dot(tl, tl);
            if (dist_sq > 0.0f) {
                float dist = sqrtf(dist_sq);
                V3 wl = tl / dist;
                float cos_l = dot(ld3(L.normal), -wl);
                if (cos_l > 1.0e-6f) {
                    float lp = (1.0f / ((float)n_lights * L.area)) * dist_sq / cos_l;
                    if (isfinite(lp) && lp > 0.0f) {
                        wi_world = wl;
                        rin = ld3(sc.materials[L.material].emissive);
                        distance = dist;
                        pdf = lp * ((float)n_lights / (float)total);
                        valid = true;
                    }
                }
            }
        } else {
            V3 ed, er;
            float ep = env_sample(sc.env, ua_x, ua_y, &ed, &er);
            if (ep > 0.0f) {
                wi_world = ed;
                rin = er;
                distance = INFINITY;
                pdf = ep / (float)total;
                valid = true;
            }
        }
        if (valid && pdf > 0.0f) {
            V3 wi = to_local(wi_world, normal);
            if (wi.z != 0.0f && (wi.z > 0.0f || surf.transmission > 0.0f)) {
                V3 f;
                float bsdf_pdf;
                surface_eval_pdf(surf, wo, wi, &f, &bsdf_pdf);
                if (max3(f) > 0.0f) {
                    float mis = mode == LUCIDA_SAMPLING_MIS ? power_heuristic(pdf, bsdf_pdf) : 1.0f;
                    if (mis > 0.0f) {
                        sr.o = offset_ray_origin(position, wi.z < 0.0f ? -gnormal : gnormal);
                        sr.d = wi_world;
                        sr.t_max = distance * (1.0f - 1.0e-3f);
                        sr.c = P.throughput * (f * fabsf(wi.z) * rin * (mis / pdf));
                        shadow = true;
                    }
                }
            }
        }
    }
    const float u_lobe = s.next_f32();
    float bx, by;
    s.next_vec2(&bx, &by);
    V3 wi, weight;
    const float bpdf = surface_sample(surf, wo, u_lobe, bx, by, &wi, &weight);
    if (!(bpdf > 0.0f)) return false;
    V3 nt = P.throughput * weight;
    if (!(max3(nt) > 0.0f && depth + 1 < u.max_depth)) return false;
    P.origin = offset_ray_origin(position, wi.z < 0.0f ? -gnormal : gnormal);
    P.dir = to_world(wi, normal);
    P.throughput = nt;
    P.prev_bsdf_pdf = bpdf;
    P.prev_position = position;
    P.depth = depth + 1;
    return true;
}

// With `key_out`, each appended extension-queue entry gets its ray's sort key
// (queue_key) at the same slot, and each shadow ray its key and slot in
// skey / sidx.
template <bool BATCH>
__global__ __launch_bounds__(SHADE_WG) void k_shade(SceneView sc0, const GpuUniforms* __restrict__ U,
                                              const BatchJob* __restrict__ jobs, Planes p,
                                              const uint32_t* __restrict__ queue_in,
                                              const uint32_t* __restrict__ count_ptr, uint32_t* __restrict__ queue_out,
                                              uint32_t* __restrict__ next_count, uint32_t* __restrict__ shadow_count,
                                              uint32_t* __restrict__ key_out, Band band0) {
    const uint32_t i = blockIdx.x * SHADE_WG + threadIdx.x;
    const uint32_t count = *count_ptr;
    bool alive = false, shadow = false;
    uint32_t pid = 0, key = 0, skey = 0;
    ShadowRay sr{};
    if (i < count) {
      pid = queue_in[i];
      const uint32_t jid = BATCH ? p.job[pid] : 0u;
      for (;;) {
        const uint32_t j = BATCH ? first_job(jid) : 0u;
        if (jid != j) continue;
        const SceneView& sc = BATCH ? jobs[j].sc : sc0;
        const GpuUniforms& u = BATCH ? jobs[j].u : *U;
        const uint32_t local = pid - (BATCH ? jobs[j].path_base : 0u);  // index within its own job
        const float4 hit = p.hit[i];
        const float4 of = p.org[pid], df = p.dir[pid], tf = p.thr[pid], pf = p.prv[pid], rf = p.rad[pid];
        const Band band = BATCH ? Band{0u, jobs[j].pixels} : band0;
        PathRegs P;
        P.origin = f4xyz(of);
        P.dir = f4xyz(df);
        P.throughput = f4xyz(tf);
        P.radiance = f4xyz(rf);
        P.prev_position = f4xyz(pf);
        P.prev_bsdf_pdf = tf.w;
        P.depth = __float_as_uint(of.w);
        P.s.sobol = u.sampler_kind == LUCIDA_SAMPLER_SOBOL;
        P.s.state = __float_as_uint(df.w);
        P.s.index = u.sample_offset + local / band.count;
        P.s.seed = hash_u32(band.first + local % band.count + u.frame_seed * 0x9e3779b9u);
        P.s.dim = __float_as_uint(pf.w);
        bool guide;
        float4 gal, gnr;
        alive = shade_bounce(sc, u, hit, (sc.num_instances && __float_as_uint(hit.w) != MISS) ? p.hin[i] : 0u, P,
                             shadow, sr, guide, gal, gnr);
        if (guide) {
            p.gal[pid] = gal;
            p.gnr[pid] = gnr;
        }
        p.rad[pid] = make_float4(P.radiance.x, P.radiance.y, P.radiance.z, 0.0f);
        if (shadow && key_out) skey = queue_key(sc, BATCH ? (int)j : -1, sr.o, sr.d);
        if (alive) {
            p.org[pid] = make_float4(P.origin.x, P.origin.y, P.origin.z, __uint_as_float(P.depth));
            p.dir[pid] = make_float4(P.dir.x, P.dir.y, P.dir.z, __uint_as_float(P.s.state));
            p.thr[pid] = make_float4(P.throughput.x, P.throughput.y, P.throughput.z, P.prev_bsdf_pdf);
            p.prv[pid] = make_float4(P.prev_position.x, P.prev_position.y, P.prev_position.z, __uint_as_float(P.s.dim));
            if (key_out) key = queue_key(sc, BATCH ? (int)j : -1, P.origin, P.dir);
        }
        break;
      }
    }
    // Compaction: one atomic per block for each queue.
    __shared__ uint32_t append_lds[SHADE_WG / 64 + 1];
    const uint32_t sslot = block_append<SHADE_WG / 64>(shadow, shadow_count, append_lds);
    if (shadow) {
        p.sro[sslot] = make_float4(sr.o.x, sr.o.y, sr.o.z, sr.t_max);
        p.srd[sslot] = make_float4(sr.d.x, sr.d.y, sr.d.z, __uint_as_float(pid));
        p.src[sslot] = make_float4(sr.c.x, sr.c.y, sr.c.z, 0.0f);
        if (key_out) {
            p.skey[sslot] = skey;
            p.sidx[sslot] = sslot;
        }
    }
    const uint32_t qslot = block_append<SHADE_WG / 64>(alive, next_count, append_lds);
    if (alive) {
        queue_out[qslot] = pid;
        if (key_out) key_out[qslot] = key;
    }
}

// ---------------------------------------------------------------------------
// CONNECT (connect.wgsl:192-240, light.rs:327 unoccluded): any-hit over the
// analytic primitives and the BVH; an unoccluded ray adds its contribution
// to its path's radiance (one shadow ray per path per bounce: no races).
// ---------------------------------------------------------------------------
template <bool BATCH, int MODE>
__global__ __launch_bounds__(TRACE_WG) void k_connect(SceneView sc0, const BatchJob* __restrict__ jobs, Planes p,
                                                      const uint32_t* __restrict__ count_ptr,
                                                      const uint32_t* __restrict__ order, uint32_t stats_on) {
    __shared__ uint32_t stack_mem[LDS_STACK * TRACE_WG];
    const uint32_t q = blockIdx.x * TRACE_WG + threadIdx.x;
    const uint32_t count = *count_ptr;
    uint32_t visits = 0, tests = 0, overflow = 0;
    if (q < count) {
      const uint32_t i = order ? order[q] : q;  // the shadow record this lane traces
      RayStack st = make_stack(stack_mem, p.spill2, MODE == TRAV_BINARY ? STACK : INST_STACK);
      const float4 a = p.sro[i], b = p.srd[i];
      const uint32_t jid = BATCH ? p.job[__float_as_uint(b.w)] : 0u;
      for (;;) {
        const uint32_t j = BATCH ? first_job(jid) : 0u;
        if (jid != j) continue;
        const SceneView& sc = BATCH ? jobs[j].sc : sc0;
        const V3 o = f4xyz(a), d = f4xyz(b);
        const float t_max = a.w;
        bool blocked = analytic_any(sc, o, d, t_max);
        if (!blocked) {
            float t, u, v;
            uint32_t tri, inst;
            if (MODE == TRAV_INSTANCED && sc.num_instances)
                blocked = traverse_instances<true>(sc, st, o, d, T_MIN, t_max, &t, &u, &v, &tri, &inst, &visits,
                                                   &tests, &overflow);
            else
                blocked = traverse<true>(sc, st, 0, o, d, T_MIN, t_max, &t, &u, &v, &tri, &visits, &tests, &overflow);
        }
        if (!blocked) {
            const uint32_t pid = __float_as_uint(b.w);
            const float4 c = p.src[i];
            float4 r = p.rad[pid];
            r.x += c.x;
            r.y += c.y;
            r.z += c.z;
            p.rad[pid] = r;
        }
        break;
      }
    }
    flush_traversal_stats(p.counters, visits, tests, overflow, stats_on, 2);
}

// ---------------------------------------------------------------------------
// Ray casting without shading (the scene_intersect / scene_occluded queries
// EXTEND and CONNECT make): rays are (origin, unused, direction, t_max);
// hits are (t, u, v, id) with id = triangle, ANALYTIC | primitive or MISS
// (any-hit: id 0 when occluded), plus the instance of an instanced hit.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(TRACE_WG) void k_cast(SceneView sc, const float4* __restrict__ rays, uint32_t n,
                                                   uint32_t any_hit, float4* __restrict__ hits,
                                                   uint32_t* __restrict__ insts, uint32_t* __restrict__ spill) {
    __shared__ uint32_t stack_mem[LDS_STACK * TRACE_WG];
    const uint32_t i = blockIdx.x * TRACE_WG + threadIdx.x;
    if (i >= n) return;
    RayStack st = make_stack(stack_mem, spill, sc.num_instances ? INST_STACK : STACK);
    const float4 a = rays[2 * i], b = rays[2 * i + 1];
    const V3 o = f4xyz(a), d = f4xyz(b);
    float best_t = b.w;
    uint32_t best = MISS;
    for (uint32_t k = 0; k < sc.num_prims; ++k) {
        const GpuPrimitive& pr = sc.prims[k];
        float t;
        const bool h = pr.kind == LUCIDA_PRIM_KIND_QUAD ? intersect_quad(pr, f4xyz(sc.prim_w[k]), o, d, T_MIN, best_t, &t)
                                                        : intersect_sphere(pr, o, d, T_MIN, best_t, &t);
        if (h) {
            best_t = t;
            best = k;
            if (any_hit) break;
        }
    }
    uint32_t visits = 0, tests = 0, overflow = 0, tri = 0, inst = 0;
    float t = 0.0f, u = 0.0f, v = 0.0f;
    float4 out = best != MISS ? make_float4(best_t, 0.0f, 0.0f, __uint_as_float(ANALYTIC | best))
                              : make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(MISS));
    if (!(any_hit && best != MISS)) {
        bool got;
        if (any_hit)
            got = sc.num_instances ? traverse_instances<true>(sc, st, o, d, T_MIN, best_t, &t, &u, &v, &tri, &inst,
                                                              &visits, &tests, &overflow)
                                   : traverse<true>(sc, st, 0, o, d, T_MIN, best_t, &t, &u, &v, &tri, &visits, &tests,
                                                    &overflow);
        else
            got = sc.num_instances ? traverse_instances<false>(sc, st, o, d, T_MIN, best_t, &t, &u, &v, &tri, &inst,
                                                               &visits, &tests, &overflow)
                                   : traverse<false>(sc, st, 0, o, d, T_MIN, best_t, &t, &u, &v, &tri, &visits, &tests,
                                                     &overflow);
        if (got) out = any_hit ? make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(0u)) : make_float4(t, u, v, __uint_as_float(tri));
    }
    hits[i] = out;
    insts[i] = inst;
}

// ---------------------------------------------------------------------------
// MEGAKERNEL (shaders/trace/megakernel.wgsl:61-216): the whole path in one
// thread. A thread owns one pixel of the band and traces its samples one
// after another, each through every bounce: the closest-hit walk, the
// bounce's shading (shade_bounce, the wavefront SHADE's own code), the shadow
// ray's occlusion walk, then the next ray. The samples' sums are added to the
// film in sample order, as RESOLVE adds a batch, so the film equals the
// wavefront render's. The comparison architecture: every lane carries its
// path to the end, so a wave runs as long as its longest path and its lanes
// diverge through every stage, where the wavefront regroups live paths
// between stages.
// ---------------------------------------------------------------------------
template <int MODE>
__global__ __launch_bounds__(TRACE_WG) void k_megakernel(SceneView sc, const GpuUniforms* __restrict__ U,
                                                         GpuAccum* __restrict__ accum, Band band,
                                                         uint32_t* __restrict__ spill, uint32_t* __restrict__ counters,
                                                         uint32_t stats_on) {
    __shared__ uint32_t stack_mem[LDS_STACK * TRACE_WG];
    const uint32_t px = blockIdx.x * TRACE_WG + threadIdx.x;
    uint32_t visits = 0, tests = 0, overflow = 0, rays = 0, shadow_rays = 0;
    if (px < band.count) {
        const GpuUniforms& u = *U;
        RayStack st = make_stack(stack_mem, spill, MODE == TRAV_BINARY ? STACK : INST_STACK);
        const uint32_t pixel = band.first + px;
        V3 sum{0, 0, 0}, sq{0, 0, 0}, al{0, 0, 0}, nr{0, 0, 0};
        float dp = 0.0f;
        for (uint32_t j = 0; j < u.samples_per_launch; ++j) {
            // GENERATE
            PathRegs P;
            V3 origin;
            camera_ray(u, pixel, u.sample_offset + j, &P.s, &origin, &P.dir);
            P.origin = origin;
            P.throughput = v3(1.0f, 1.0f, 1.0f);
            P.radiance = v3(0.0f, 0.0f, 0.0f);
            P.prev_position = origin;
            P.prev_bsdf_pdf = 0.0f;
            P.depth = 0;
            float4 g_al = make_float4(0.0f, 0.0f, 0.0f, 0.0f), g_nr = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            for (;;) {
                // EXTEND: analytic primitives by brute force, then the tree
                ++rays;
                const V3 o = P.origin, d = P.dir;
                float best_t = T_MAX;
                uint32_t best = MISS;
                analytic_closest(sc, o, d, &best_t, &best);
                float t = 0.0f, hu = 0.0f, hv = 0.0f;
                uint32_t tri = 0, inst = 0;
                const bool two_level = MODE == TRAV_INSTANCED && sc.num_instances;
                bool th;
                if (two_level)
                    th = traverse_instances<false>(sc, st, o, d, T_MIN, best_t, &t, &hu, &hv, &tri, &inst, &visits,
                                                   &tests, &overflow);
                else
                    th = traverse<false>(sc, st, 0, o, d, T_MIN, best_t, &t, &hu, &hv, &tri, &visits, &tests,
                                         &overflow);
                float4 hit;
                if (th) hit = make_float4(t, hu, hv, __uint_as_float(tri));
                else if (best != MISS) hit = make_float4(best_t, 0.0f, 0.0f, __uint_as_float(ANALYTIC | best));
                else hit = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(MISS));
                // SHADE
                bool shadow, guide;
                ShadowRay sr;
                float4 gal, gnr;
                const bool alive = shade_bounce(sc, u, hit, two_level ? inst : 0u, P, shadow, sr, guide, gal, gnr);
                if (guide) {
                    g_al = gal;
                    g_nr = gnr;
                }
                // CONNECT
                if (shadow) {
                    ++shadow_rays;
                    bool blocked = analytic_any(sc, sr.o, sr.d, sr.t_max);
                    if (!blocked) {
                        float tt, uu, vv;
                        uint32_t tri2, inst2;
                        if (two_level)
                            blocked = traverse_instances<true>(sc, st, sr.o, sr.d, T_MIN, sr.t_max, &tt, &uu, &vv,
                                                               &tri2, &inst2, &visits, &tests, &overflow);
                        else
                            blocked = traverse<true>(sc, st, 0, sr.o, sr.d, T_MIN, sr.t_max, &tt, &uu, &vv, &tri2,
                                                     &visits, &tests, &overflow);
                    }
                    if (!blocked) P.radiance = v3(P.radiance.x + sr.c.x, P.radiance.y + sr.c.y, P.radiance.z + sr.c.z);
                }
                if (!alive) break;
            }
            // RESOLVE's per-sample sums, in sample order
            sum = sum + P.radiance;
            sq = sq + P.radiance * P.radiance;
            al = al + f4xyz(g_al);
            nr = nr + f4xyz(g_nr);
            dp = dp + g_al.w;
        }
        float4* a = reinterpret_cast<float4*>(accum + pixel);
        float4 a0 = a[0], a1 = a[1], a2 = a[2], a3 = a[3];
        a0.x += sum.x; a0.y += sum.y; a0.z += sum.z; a0.w += (float)u.samples_per_launch;
        a1.x += al.x; a1.y += al.y; a1.z += al.z; a1.w += dp;
        a2.x += nr.x; a2.y += nr.y; a2.z += nr.z;
        a3.x += sq.x; a3.y += sq.y; a3.z += sq.z;
        a[0] = a0; a[1] = a1; a[2] = a2; a[3] = a3;
    }
    // ray counts into bounce 0's slots (the fold sums every bounce's slot)
    const uint64_t wr = wave_sum_u32(rays), ws = wave_sum_u32(shadow_rays);
    if ((threadIdx.x & 63) == 0) {
        if (wr) atomicAdd(&counters[0], (uint32_t)wr);
        if (ws) atomicAdd(&counters[SHADOW_AT], (uint32_t)ws);
    }
    flush_traversal_stats(counters, visits, tests, overflow, stats_on, 0);
}

// ---------------------------------------------------------------------------
// RESOLVE (resolve.wgsl:305-338): per pixel, sum the batch's samples in sample
// order and add them to the film.
// ---------------------------------------------------------------------------
// Pixel band.first + px of a render whose paths start at `base`: its samples
// in order.
__device__ inline void resolve_pixel(const GpuUniforms& u, const Planes& p, uint32_t base, uint32_t px,
                                     GpuAccum* accum, Band band) {
    V3 sum{0, 0, 0}, sq{0, 0, 0}, al{0, 0, 0}, nr{0, 0, 0};
    float dp = 0.0f;
    for (uint32_t j = 0; j < u.samples_per_launch; ++j) {
        const uint32_t q = base + px + j * band.count;
        const V3 r = f4xyz(p.rad[q]);
        const float4 ga = p.gal[q];
        sum = sum + r;
        sq = sq + r * r;
        al = al + f4xyz(ga);
        nr = nr + f4xyz(p.gnr[q]);
        dp = dp + ga.w;
    }
    float4* a = reinterpret_cast<float4*>(accum + band.first + px);
    float4 a0 = a[0], a1 = a[1], a2 = a[2], a3 = a[3];
    a0.x += sum.x; a0.y += sum.y; a0.z += sum.z; a0.w += (float)u.samples_per_launch;
    a1.x += al.x; a1.y += al.y; a1.z += al.z; a1.w += dp;
    a2.x += nr.x; a2.y += nr.y; a2.z += nr.z;
    a3.x += sq.x; a3.y += sq.y; a3.z += sq.z;
    a[0] = a0; a[1] = a1; a[2] = a2; a[3] = a3;
}

__global__ __launch_bounds__(256) void k_resolve(const GpuUniforms* __restrict__ U, Planes p, GpuAccum* accum,
                                                 Band band) {
    const GpuUniforms& u = *U;
    const uint32_t px = blockIdx.x * 256 + threadIdx.x;
    if (px >= band.count) return;
    resolve_pixel(u, p, 0u, px, accum, band);
}

// RESOLVE for one round of a batched render: every job's pixels, each into its film.
__global__ __launch_bounds__(256) void k_resolve_batch(const BatchJob* __restrict__ jobs, uint32_t njobs,
                                                       uint32_t total_pixels, Planes p) {
    const uint32_t g = blockIdx.x * 256 + threadIdx.x;
    if (g >= total_pixels) return;
    const uint32_t jid = find_job(njobs, g, [&](uint32_t k) { return jobs[k].pixel_base; });
    for (;;) {
        const uint32_t j = first_job(jid);
        if (jid != j) continue;
        resolve_pixel(jobs[j].u, p, jobs[j].path_base, g - jobs[j].pixel_base, jobs[j].film, Band{0u, jobs[j].pixels});
        break;
    }
}

// Adds one batch's counters into the render's running totals (u64 at word
// TOTALS_AT): extension rays, shadow rays, closest-hit visits and tests,
// any-hit visits and tests, stack overflows.
__global__ void k_fold_counters(uint32_t* counters, uint32_t depth) {
    if (threadIdx.x != 0) return;
    uint64_t* tot = reinterpret_cast<uint64_t*>(counters + TOTALS_AT);
    const uint64_t* batch = reinterpret_cast<const uint64_t*>(counters + STATS_AT);
    uint64_t ext = 0, sh = 0;
    for (uint32_t b = 0; b < depth; ++b) {
        ext += counters[b];
        sh += counters[SHADOW_AT + b];
    }
    tot[0] += ext;
    tot[1] += sh;
    for (int k = 0; k < 7; ++k) {
        uint64_t v = 0;
        for (uint32_t c = 0; c < STATS_STRIPES; ++c) v += batch[8 * c + k];
        tot[2 + k] += v;
    }
}

// Clears the per-bounce counters with a kernel (not a memset) so the batch can
// be captured in a graph and replayed.
__global__ void k_clear_counters(uint32_t* counters, uint32_t n) {
    for (uint32_t k = threadIdx.x; k < n; k += blockDim.x) counters[k] = 0;
}

// Device wall-clock stamp (constant-rate counter) after everything queued before it.
__global__ void k_stamp(unsigned long long* out, uint32_t slot) {
    if (threadIdx.x == 0) out[slot] = (unsigned long long)wall_clock64();
}

// bsdf.rs:1200-1227 energy::compute_table on the device: 32 x 32 entries,
// each the mean over a 128 x 128 stratified grid of the white-Fresnel
// single-scattering weight. One workgroup per entry, summed in f64 as the
// reference does.
__global__ __launch_bounds__(256) void k_ggx_table(float* table) {
    const int ri = blockIdx.x / 32, ci = blockIdx.x % 32;
    const float roughness = (float)ri / 31.0f;
    const float alpha = roughness_to_alpha(roughness);
    const float root_mu = (float)ci / 31.0f;
    const float mu = fmaxf(root_mu * root_mu, 1.0e-4f);
    const float sin_t = sqrtf(fmaxf(1.0f - mu * mu, 0.0f));
    const V3 wo{sin_t, 0.0f, mu};
    const float n = 128.0f;
    double sum = 0.0;
    for (int k = threadIdx.x; k < 128 * 128; k += 256) {
        const int a = k / 128, b = k % 128;
        const float ux = ((float)a + 0.5f) / n, uy = ((float)b + 0.5f) / n;
        sum += (double)specular_single_weight_white(alpha, wo, ux, uy);
    }
    __shared__ double red[256];
    red[threadIdx.x] = sum;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) table[blockIdx.x] = (float)(red[0] / (128.0 * 128.0));
}

inline uint32_t grid_for(size_t n, int block) { return (uint32_t)((n + block - 1) / block); }

// Per analytic primitive: a quad's w (quad_w); zero for spheres.
__global__ void k_prim_planes(const GpuPrimitive* __restrict__ prims, size_t n, float4* __restrict__ w) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const V3 q = prims[i].kind == LUCIDA_PRIM_KIND_QUAD ? quad_w(prims[i]) : V3{0.0f, 0.0f, 0.0f};
    w[i] = make_float4(q.x, q.y, q.z, 0.0f);
}

// The box around every analytic primitive (a quad by its four corners, a
// sphere by center +- radius) at w[n] (low) and w[n + 1] (high), then the box
// of each run of PRIM_CHUNK consecutive primitives g at w[n + 2 + 2g] and
// w[n + 3 + 2g], all padded by 1e-3 of the largest coordinate magnitude
// (plus 1e-6): far beyond the rounding of the primitive tests and of the
// slab test, so a ray that a primitive test accepts always hits its boxes
// (a flat set of quads gets a thickness).
__device__ inline void prim_bounds(const GpuPrimitive& q, float* lo, float* hi) {
    for (int a = 0; a < 3; ++a) {
        if (q.kind == LUCIDA_PRIM_KIND_QUAD) {
            const float p = q.position[a], u = q.edge_u[a], v = q.edge_v[a];
            lo[a] = fminf(fminf(p, p + u), fminf(p + v, p + u + v));
            hi[a] = fmaxf(fmaxf(p, p + u), fmaxf(p + v, p + u + v));
        } else {
            lo[a] = q.position[a] - fabsf(q.radius);
            hi[a] = q.position[a] + fabsf(q.radius);
        }
    }
}

__global__ __launch_bounds__(256) void k_prim_box(const GpuPrimitive* __restrict__ prims, size_t n,
                                                  float4* __restrict__ w) {
    __shared__ float red[6][256];
    float lo[3] = {3.0e38f, 3.0e38f, 3.0e38f}, hi[3] = {-3.0e38f, -3.0e38f, -3.0e38f};
    for (size_t i = threadIdx.x; i < n; i += 256) {
        float l[3], h[3];
        prim_bounds(prims[i], l, h);
        for (int a = 0; a < 3; ++a) {
            lo[a] = fminf(lo[a], l[a]);
            hi[a] = fmaxf(hi[a], h[a]);
        }
    }
    for (int a = 0; a < 3; ++a) {
        red[a][threadIdx.x] = lo[a];
        red[3 + a][threadIdx.x] = hi[a];
    }
    __syncthreads();
    for (int half = 128; half > 0; half >>= 1) {
        if ((int)threadIdx.x < half)
            for (int a = 0; a < 3; ++a) {
                red[a][threadIdx.x] = fminf(red[a][threadIdx.x], red[a][threadIdx.x + half]);
                red[3 + a][threadIdx.x] = fmaxf(red[3 + a][threadIdx.x], red[3 + a][threadIdx.x + half]);
            }
        __syncthreads();
    }
    float m = 0.0f;
    for (int a = 0; a < 6; ++a) m = fmaxf(m, fabsf(red[a][0]));
    const float pad = 1.0e-3f * m + 1.0e-6f;
    if (threadIdx.x == 0) {
        w[n] = make_float4(red[0][0] - pad, red[1][0] - pad, red[2][0] - pad, 0.0f);
        w[n + 1] = make_float4(red[3][0] + pad, red[4][0] + pad, red[5][0] + pad, 0.0f);
    }
    for (size_t g = threadIdx.x; g * PRIM_CHUNK < n; g += 256) {
        float l[3] = {3.0e38f, 3.0e38f, 3.0e38f}, h[3] = {-3.0e38f, -3.0e38f, -3.0e38f};
        for (size_t k = g * PRIM_CHUNK; k < n && k < g * PRIM_CHUNK + PRIM_CHUNK; ++k) {
            float pl[3], ph[3];
            prim_bounds(prims[k], pl, ph);
            for (int a = 0; a < 3; ++a) {
                l[a] = fminf(l[a], pl[a]);
                h[a] = fmaxf(h[a], ph[a]);
            }
        }
        w[n + 2 + 2 * g] = make_float4(l[0] - pad, l[1] - pad, l[2] - pad, 0.0f);
        w[n + 3 + 2 * g] = make_float4(h[0] + pad, h[1] + pad, h[2] + pad, 0.0f);
    }
}

// Per triangle (in the scene's leaf order): p0, p1 - p0, p2 - p0.
__global__ void k_tri_edges(const GpuTriangle* __restrict__ tris, const float4* __restrict__ pos, size_t n,
                            float4* __restrict__ edges) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const GpuTriangle t = tris[i];
        const V3 p0 = f4xyz(pos[t.i0]);
        const V3 e1 = f4xyz(pos[t.i1]) - p0;
        const V3 e2 = f4xyz(pos[t.i2]) - p0;
        edges[3 * i] = make_float4(p0.x, p0.y, p0.z, 0.0f);
        edges[3 * i + 1] = make_float4(e1.x, e1.y, e1.z, 0.0f);
        edges[3 * i + 2] = make_float4(e2.x, e2.y, e2.z, 0.0f);
    }
}

}  // namespace

lucida_status build_prim_planes(lucida_ctx* ctx, lucida_scene* scene) {
    const size_t n = scene->num_primitives;
    if (n == 0) return LUCIDA_OK;
    // per primitive its quad_w, then the primitives' boxes (k_prim_box)
    const size_t chunks = (n + PRIM_CHUNK - 1) / PRIM_CHUNK;
    if (!scene->prim_w) LUCIDA_HIP(hipMalloc(&scene->prim_w, (n + 2 + 2 * chunks) * sizeof(float4)), LUCIDA_OOM);
    hipLaunchKernelGGL(k_prim_planes, dim3(grid_for(n, 64)), dim3(64), 0, ctx->stream, scene->primitives, n,
                       scene->prim_w);
    hipLaunchKernelGGL(k_prim_box, dim3(1), dim3(256), 0, ctx->stream, scene->primitives, n, scene->prim_w);
    LUCIDA_HIP(hipGetLastError(), LUCIDA_SHADER);
    return LUCIDA_OK;
}

lucida_status build_tri_edges(lucida_ctx* ctx, lucida_scene* scene) {
    const size_t n = scene->num_triangles;
    if (!scene->tri_edges) LUCIDA_HIP(hipMalloc(&scene->tri_edges, n * 3 * sizeof(float4)), LUCIDA_OOM);
    const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_tri_edges, dim3(blocks), dim3(256), 0, ctx->stream, scene->triangles,
                       reinterpret_cast<const float4*>(scene->positions), n, scene->tri_edges);
    LUCIDA_HIP(hipGetLastError(), LUCIDA_SHADER);
    return LUCIDA_OK;
}

lucida_status ensure_ggx_table(lucida_ctx* ctx) {
    if (ctx->ggx_ready) return LUCIDA_OK;
    LUCIDA_HIP(ctx->ggx_table.ensure(1024 * sizeof(float)), LUCIDA_OOM);
    hipLaunchKernelGGL(k_ggx_table, dim3(1024), dim3(256), 0, ctx->stream, ctx->ggx_table.as<float>());
    LUCIDA_HIP(hipGetLastError(), LUCIDA_SHADER);
    LUCIDA_HIP(hipStreamSynchronize(ctx->stream), LUCIDA_VALIDATION);
    ctx->ggx_ready = true;
    return LUCIDA_OK;
}

// Points the path and shadow fields into the context's buffers (capacity C)
// as records or as arrays of fields.
static void set_layout(const WfPipe& w, Planes* out, bool records) {
    const size_t C = w.capacity;
    float4* pl = w.planes.as<float4>();
    PathField* path[7] = {&out->org, &out->dir, &out->thr, &out->prv, &out->rad, &out->gal, &out->gnr};
    for (int k = 0; k < 7; ++k)
        *path[k] = records ? PathField{pl + k, PATH_RECORD} : PathField{pl + k * C, 1u};
    float4* sh = w.shadow.as<float4>();
    ShadowField* shadow[3] = {&out->sro, &out->srd, &out->src};
    for (int k = 0; k < 3; ++k)
        *shadow[k] = records ? ShadowField{sh + k, SHADOW_RECORD} : ShadowField{sh + k * C, 1u};
}

// Device buffers of a wavefront pipeline, grown on demand and kept by the
// context; `records` picks the state layout (records for sorted queues).
static lucida_status ensure_wavefront(WfPipe& w, size_t P, Planes* out, GpuUniforms** dU, bool records) {
    if (P > w.capacity) {
        LUCIDA_HIP(hipStreamSynchronize(w.s), LUCIDA_VALIDATION);
        LUCIDA_HIP(hipStreamSynchronize(w.s2), LUCIDA_VALIDATION);
        LUCIDA_HIP(w.planes.ensure(P * 16 * PATH_RECORD), LUCIDA_OOM);
        LUCIDA_HIP(w.hits.ensure(P * 20), LUCIDA_OOM);
        LUCIDA_HIP(w.shadow.ensure(P * 16 * SHADOW_RECORD), LUCIDA_OOM);
        // q0, q1, job, the sorts' keys, orders and temporaries, then their scratch
        LUCIDA_HIP(w.queues.ensure((P * 10 + 2 * sort_counted_scratch_words(P)) * 4), LUCIDA_OOM);
        // spill columns for every lane of a traversal grid over P items
        LUCIDA_HIP(w.spill.ensure(2 * (P + TRACE_WG) * (INST_STACK - LDS_STACK) * 4), LUCIDA_OOM);
        w.capacity = P;
    }
    LUCIDA_HIP(w.counters.ensure(COUNTER_BYTES), LUCIDA_OOM);
    LUCIDA_HIP(w.uniforms.ensure(sizeof(GpuUniforms) * 2), LUCIDA_OOM);
    const size_t C = w.capacity;
    set_layout(w, out, records);
    uint32_t* q = w.queues.as<uint32_t>();
    out->hit = w.hits.as<float4>();
    out->hin = reinterpret_cast<uint32_t*>(w.hits.as<float4>() + C);
    out->q0 = q;
    out->q1 = q + C;
    out->job = q + 2 * C;
    out->qkey = q + 3 * C;
    out->qtmp = q + 4 * C;
    out->ktmp = q + 5 * C;
    out->skey = q + 6 * C;
    out->sidx = q + 7 * C;
    out->qtmp2 = q + 8 * C;
    out->ktmp2 = q + 9 * C;
    out->sort_scratch = q + 10 * C;
    out->sort_scratch2 = out->sort_scratch + sort_counted_scratch_words(C);
    out->spill = w.spill.as<uint32_t>();
    out->spill2 = out->spill + (C + TRACE_WG) * (INST_STACK - LDS_STACK);
    out->counters = w.counters.as<uint32_t>();
    *dU = w.uniforms.as<GpuUniforms>();
    return LUCIDA_OK;
}

}  // namespace lucida

using namespace lucida;

// The walks are binary: bvh_width 0 (default) or 2.
static bool binary_width(const lucida_wavefront_config& cfg) { return cfg.bvh_width == 0 || cfg.bvh_width == 2; }

// EXTEND / CONNECT for the traversal variant `mode` (TRAV_*).
template <bool BATCH>
static void launch_extend(int mode, dim3 grid, hipStream_t s, const SceneView& sc, const BatchJob* jobs,
                          const Planes& p, const uint32_t* queue, const uint32_t* count, uint32_t stats_on) {
    if (mode == TRAV_INSTANCED)
        hipLaunchKernelGGL((k_extend<BATCH, TRAV_INSTANCED>), grid, dim3(TRACE_WG), 0, s, sc, jobs, p, queue, count,
                           stats_on);
    else
        hipLaunchKernelGGL((k_extend<BATCH, TRAV_BINARY>), grid, dim3(TRACE_WG), 0, s, sc, jobs, p, queue, count,
                           stats_on);
}

template <bool BATCH>
static void launch_connect(int mode, dim3 grid, hipStream_t s, const SceneView& sc, const BatchJob* jobs,
                           const Planes& p, const uint32_t* count, const uint32_t* order, uint32_t stats_on) {
    if (mode == TRAV_INSTANCED)
        hipLaunchKernelGGL((k_connect<BATCH, TRAV_INSTANCED>), grid, dim3(TRACE_WG), 0, s, sc, jobs, p, count, order,
                           stats_on);
    else
        hipLaunchKernelGGL((k_connect<BATCH, TRAV_BINARY>), grid, dim3(TRACE_WG), 0, s, sc, jobs, p, count, order,
                           stats_on);
}


// Whether a render of `paths` paths per batch sorts its queues
// (lucida_wavefront_config.ray_order). The sorts cost a fixed ~0.1 ms per
// bounce plus their passes; they pay off when the walks are long (a large
// tree) and the queues long enough to group similar rays. On the 1.05M-
// triangle scene's restructured tree at 1024^2, 8 bounces (two bands of rows
// per one-batch render): 1 spp, 2^19 paths per band, 5.3 ms unsorted against
// 7.2 sorted; 2 spp, 2^20 per band, 8.6 against 9.5; 4 spp, 2^21 per band,
// 16.2 against 14.2.
static bool want_ray_order(const lucida_wavefront_config& cfg, const SceneView& sc, size_t paths) {
    if (sc.num_nodes == 0 || cfg.ray_order == 1) return false;
    if (cfg.ray_order == 2) return true;
    return sc.num_nodes >= 4096 && paths >= ((size_t)1 << 21);
}

// The sorts after SHADE of bounce b over a batch of P paths, by `bits` key
// bits, on pipeline `w`: the extension queue SHADE wrote to `qraw` on w.s
// (unless this was the last bounce; *next_queue receives the sorted queue)
// and, at the same time on `side` (w.s2, or w.s in a serialized render), the
// shadow rays (*shadow_order receives the record slots in CONNECT's order).
// Each stream then waits for the other's sort, so EXTEND and CONNECT start
// together, after both: a walk's long waves beside a sort would hold the CUs
// its short kernels need. The caller has recorded w.ev_shaded on w.s after
// SHADE. Every step is a launch or an event, so the sequence may be captured
// in a graph.
static lucida_status sort_queues(WfPipe& w, hipStream_t side, const Planes& p, uint32_t* qraw, size_t P, uint32_t b,
                                 bool extend, uint32_t bits, uint32_t** next_queue, uint32_t** shadow_order) {
    *next_queue = qraw;
    LUCIDA_HIP(hipStreamWaitEvent(side, w.ev_shaded, 0), LUCIDA_VALIDATION);
    lucida_status st = sort_pairs_counted(side, p.skey, p.sidx, p.ktmp2, p.qtmp2, P, p.counters + SHADOW_AT + b,
                                          bits, p.sort_scratch2, shadow_order);
    if (st != LUCIDA_OK) return st;
    if (extend) {
        st = sort_pairs_counted(w.s, p.qkey, qraw, p.ktmp, p.qtmp, P, p.counters + b + 1, bits, p.sort_scratch,
                                next_queue);
        if (st != LUCIDA_OK) return st;
    }
    LUCIDA_HIP(hipEventRecord(w.ev_shadow_sorted, side), LUCIDA_VALIDATION);
    LUCIDA_HIP(hipEventRecord(w.ev_sorted, w.s), LUCIDA_VALIDATION);
    LUCIDA_HIP(hipStreamWaitEvent(w.s, w.ev_shadow_sorted, 0), LUCIDA_VALIDATION);
    LUCIDA_HIP(hipStreamWaitEvent(side, w.ev_sorted, 0), LUCIDA_VALIDATION);
    return LUCIDA_OK;
}

// The kernels' view of a resident scene.
static SceneView make_view(lucida_ctx* ctx, const lucida_scene* scene) {
    SceneView sc;
    sc.materials = scene->materials;
    sc.prims = scene->primitives;
    sc.prim_w = scene->prim_w;
    sc.pos = reinterpret_cast<const float4*>(scene->positions);
    sc.attrs = scene->vertex_attrs;
    sc.tris = scene->triangles;
    sc.edges = scene->tri_edges;
    sc.nodes = scene->bvh_nodes;
    sc.insts = reinterpret_cast<const GpuInstance*>(scene->primitives);
    sc.num_instances = (uint32_t)scene->num_instances;
    sc.tlas_root = scene->tlas_root;
    sc.lights = scene->lights;
    sc.ggx = ctx->ggx_table.as<float>();
    sc.env.w = scene->env_width;
    sc.env.h = scene->env_height;
    sc.env.total_weight = scene->env_total_weight;
    sc.env.pixels = reinterpret_cast<const float4*>(scene->env_pixels);
    sc.env.conditional = scene->env_conditional;
    sc.env.marginal = scene->env_marginal;
    sc.num_prims = (uint32_t)scene->num_primitives;
    sc.num_tris = (uint32_t)scene->num_triangles;
    sc.num_nodes = (uint32_t)scene->num_bvh_nodes;
    sc.num_lights = (uint32_t)scene->num_lights;
    return sc;
}

extern "C" {

lucida_status lucida_energy_table(lucida_ctx* ctx, float* out) {
    if (!ctx || !out) return set_error(LUCIDA_INVALID_ARG, "lucida_energy_table: null argument");
    lucida_status st = ensure_ggx_table(ctx);
    if (st != LUCIDA_OK) return st;
    LUCIDA_HIP(hipMemcpyAsync(out, ctx->ggx_table.ptr, 1024 * sizeof(float), hipMemcpyDeviceToHost, ctx->stream),
               LUCIDA_READBACK);
    LUCIDA_HIP(hipStreamSynchronize(ctx->stream), LUCIDA_READBACK);
    return LUCIDA_OK;
}

lucida_status lucida_render_wavefront(lucida_ctx* ctx, const lucida_scene* scene, const GpuUniforms* uniforms,
                                      uint32_t samples, lucida_film* film, const lucida_wavefront_config* config,
                                      lucida_render_stats* stats) {
    if (!ctx || !scene || !uniforms || !film) return set_error(LUCIDA_INVALID_ARG, "lucida_render_wavefront: null argument");
    clear_error();
    if (uniforms->width != film->width || uniforms->height != film->height)
        return set_error(LUCIDA_INVALID_ARG, "uniforms width/height do not match the film");
    if (uniforms->max_depth > MAX_BOUNCES) return set_error(LUCIDA_INVALID_ARG, "max_depth above 64");
    if (scene->num_bvh_nodes >= (1u << 26) || scene->num_triangles + scene->num_primitives + scene->num_instances >= (1u << 26))
        return set_error(LUCIDA_INVALID_ARG, "scene too large: node and primitive indices are packed in 26 bits");
    if (stats) std::memset(stats, 0, sizeof *stats);
    if (config && !binary_width(*config)) return set_error(LUCIDA_INVALID_ARG, "bvh_width must be 0 or 2 (the binary walk)");
    if (samples == 0 || uniforms->max_depth == 0) return LUCIDA_OK;
    lucida_status st = ensure_ggx_table(ctx);
    if (st != LUCIDA_OK) return st;

    lucida_wavefront_config cfg{};
    if (config) cfg = *config;
    const size_t pixels = (size_t)film->width * film->height;
    uint32_t spb = cfg.samples_per_batch;
    if (spb == 0) {
        // Default: about 8M paths per batch (1 GB of path state), one such
        // batch in flight on each pipeline: larger queues sort into more
        // coherent waves (1024^2 x 16 spp: 236 Mpaths/s at 8 spp per batch,
        // 222 at 4, 205 at 2).
        size_t target = (size_t)1 << 23;
        spb = (uint32_t)std::max<size_t>(1, target / pixels);
    }
    spb = std::min(spb, samples);
    const size_t P = pixels * spb;
    if (P >= 0x7FFFFFFFull / 4) return set_error(LUCIDA_INVALID_ARG, "too many paths in one batch");

    const SceneView sc = make_view(ctx, scene);
    const int mode = scene->num_instances ? TRAV_INSTANCED : TRAV_BINARY;

    // The render's batches: spb samples of every pixel each. They alternate
    // between the context's two pipelines: while one runs its latency-bound
    // SHADE and sorts, the other's walks fill the device. A render of one
    // large batch is split into two bands of rows, one per pipeline.
    struct Item {
        uint32_t done, spb;  // first sample (relative to the render's) and count
        Band band;
    };
    std::vector<Item> items;
    const uint32_t width = film->width, height = film->height;
    if (samples <= spb && P >= ((size_t)1 << 19) && height >= 2) {
        const uint32_t top = height / 2;
        items.push_back({0u, samples, Band{0u, top * width}});
        items.push_back({0u, samples, Band{top * width, (height - top) * width}});
    } else {
        for (uint32_t done = 0; done < samples; done += spb)
            items.push_back({done, std::min(spb, samples - done), Band{0u, (uint32_t)pixels}});
    }
    size_t item_paths = 0;
    for (const Item& it : items) item_paths = std::max(item_paths, (size_t)it.band.count * it.spb);
    const bool sort_rays = want_ray_order(cfg, sc, item_paths);
    const uint32_t batches = (uint32_t)items.size();
    const int npipes = batches >= 2 && !cfg.serialize ? 2 : 1;
    Planes pl[2];
    GpuUniforms* dUs[2] = {nullptr, nullptr};
    for (int q = 0; q < npipes; ++q) {
        st = ensure_wavefront(ctx->wf[q], P, &pl[q], &dUs[q], sort_rays);
        if (st != LUCIDA_OK) return st;
    }

    const uint32_t depth = uniforms->max_depth;
    const uint32_t stats_on = (stats && cfg.count_traversal) ? 1u : 0u;
    const bool profile = stats && cfg.profile_kernels;

    // Optional per-kernel events: one start/stop pair per launch, destroyed
    // on every return path.
    struct EventSet {
        std::vector<hipEvent_t> e;
        ~EventSet() {
            for (hipEvent_t x : e)
                if (x) (void)hipEventDestroy(x);
        }
    } event_set;
    std::vector<hipEvent_t>& evs = event_set.e;
    std::vector<int> ev_kind;
    if (profile) {
        const size_t launches = (size_t)batches * (2 + 4 * depth);
        evs.resize(2 * launches, nullptr);
        for (auto& e : evs) LUCIDA_HIP(hipEventCreate(&e), LUCIDA_VALIDATION);
        ev_kind.reserve(launches);
    }
    size_t ev_next = 0;
    auto mark = [&](int kind, bool start, hipStream_t on) {
        if (!profile) return;
        if (start) ev_kind.push_back(kind);
        (void)hipEventRecord(evs[ev_next++], on);
    };

    hipStream_t s0 = ctx->wf[0].s;
    LUCIDA_HIP(hipEventRecord(ctx->ev0, s0), LUCIDA_VALIDATION);
    for (int q = 0; q < npipes; ++q) {
        if (q > 0) LUCIDA_HIP(hipStreamWaitEvent(ctx->wf[q].s, ctx->ev0, 0), LUCIDA_VALIDATION);
        hipLaunchKernelGGL(k_clear_counters, dim3(1), dim3(256), 0, ctx->wf[q].s, pl[q].counters,
                           (uint32_t)TOTALS_AT + 24u);
    }
    // each batch's uniforms, kept until the render has finished (their
    // uploads are asynchronous)
    std::vector<GpuUniforms> host_u(batches, *uniforms);
    for (uint32_t batch = 0; batch < batches; ++batch) {
        const Item& it = items[batch];
        WfPipe& w = ctx->wf[batch % npipes];
        const Planes& p = pl[batch % npipes];
        GpuUniforms* dU = dUs[batch % npipes];
        hipStream_t s = w.s, s2 = cfg.serialize ? w.s : w.s2;
        GpuUniforms& u = host_u[batch];
        u.sample_offset = uniforms->sample_offset + it.done;
        u.samples_per_launch = it.spb;
        const size_t Pb = (size_t)it.band.count * it.spb;
        LUCIDA_HIP(hipMemcpyAsync(dU, &u, sizeof u, hipMemcpyHostToDevice, s), LUCIDA_VALIDATION);
        hipLaunchKernelGGL(k_clear_counters, dim3(1), dim3(256), 0, s, p.counters, (uint32_t)TOTALS_AT);
        mark(0, true, s);
        hipLaunchKernelGGL(k_generate, dim3(grid_for(Pb, 256)), dim3(256), 0, s, dU, p, it.band);
        mark(0, false, s);
        // CONNECT of bounce b runs on the second stream beside EXTEND of
        // bounce b + 1 (shadow planes and radiance against queues, rays and
        // hits); SHADE of bounce b + 1 and RESOLVE wait for it (radiance, and
        // the shadow planes SHADE rewrites).
        uint32_t* qin = p.q0;  // GENERATE's queue
        for (uint32_t b = 0; b < depth; ++b) {
            uint32_t* qout = qin == p.q1 ? p.q0 : p.q1;  // SHADE's appends: any queue but the one it reads
            uint32_t* qnext = qout;
            uint32_t* order = nullptr;
            mark(1, true, s);
            launch_extend<false>(mode, dim3(grid_for(Pb, TRACE_WG)), s, sc, nullptr, p, qin, p.counters + b,
                                 stats_on);
            mark(1, false, s);
            if (b > 0) LUCIDA_HIP(hipStreamWaitEvent(s, w.ev_connected, 0), LUCIDA_VALIDATION);
            mark(2, true, s);
            hipLaunchKernelGGL(k_shade<false>, dim3(grid_for(Pb, SHADE_WG)), dim3(SHADE_WG), 0, s, sc, dU, nullptr, p,
                               qin, p.counters + b, qout, p.counters + b + 1, p.counters + SHADOW_AT + b,
                               sort_rays ? p.qkey : nullptr, it.band);
            mark(2, false, s);
            LUCIDA_HIP(hipEventRecord(w.ev_shaded, s), LUCIDA_VALIDATION);
            if (sort_rays) {
                mark(5, true, s);
                st = sort_queues(w, s2, p, qout, Pb, b, b + 1 < depth, RAY_KEY_BITS, &qnext, &order);
                if (st != LUCIDA_OK) return st;
                mark(5, false, s);
            } else {
                LUCIDA_HIP(hipStreamWaitEvent(s2, w.ev_shaded, 0), LUCIDA_VALIDATION);
            }
            mark(3, true, s2);
            launch_connect<false>(mode, dim3(grid_for(Pb, TRACE_WG)), s2, sc, nullptr, p, p.counters + SHADOW_AT + b,
                                  order, stats_on);
            mark(3, false, s2);
            LUCIDA_HIP(hipEventRecord(w.ev_connected, s2), LUCIDA_VALIDATION);
            qin = qnext;
        }
        LUCIDA_HIP(hipStreamWaitEvent(s, w.ev_connected, 0), LUCIDA_VALIDATION);
        // batches add into the film in order, whichever pipeline ran them
        if (batch > 0) LUCIDA_HIP(hipStreamWaitEvent(s, ctx->ev_resolved, 0), LUCIDA_VALIDATION);
        mark(4, true, s);
        hipLaunchKernelGGL(k_resolve, dim3(grid_for(it.band.count, 256)), dim3(256), 0, s, dU, p, film->accum,
                           it.band);
        mark(4, false, s);
        LUCIDA_HIP(hipEventRecord(ctx->ev_resolved, s), LUCIDA_VALIDATION);
        hipLaunchKernelGGL(k_fold_counters, dim3(1), dim3(64), 0, s, p.counters, depth);
        LUCIDA_HIP(hipGetLastError(), LUCIDA_SHADER);
    }
    for (int q = 1; q < npipes; ++q) {  // join the other pipeline
        LUCIDA_HIP(hipEventRecord(ctx->wf[q].ev_shaded, ctx->wf[q].s), LUCIDA_VALIDATION);
        LUCIDA_HIP(hipStreamWaitEvent(s0, ctx->wf[q].ev_shaded, 0), LUCIDA_VALIDATION);
    }
    LUCIDA_HIP(hipEventRecord(ctx->ev1, s0), LUCIDA_VALIDATION);
    LUCIDA_HIP(hipEventSynchronize(ctx->ev1), LUCIDA_VALIDATION);
    LUCIDA_HIP(hipGetLastError(), LUCIDA_VALIDATION);
    if (stats) {
        uint64_t totals[10] = {};
        for (int q = 0; q < npipes; ++q) {
            uint64_t t[10];
            LUCIDA_HIP(hipMemcpy(t, pl[q].counters + TOTALS_AT, sizeof t, hipMemcpyDeviceToHost), LUCIDA_READBACK);
            for (int k = 0; k < 10; ++k) totals[k] += t[k];
        }
        float ms = 0.0f;
        LUCIDA_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1), LUCIDA_VALIDATION);
        stats->extension_rays = totals[0];
        stats->shadow_rays = totals[1];
        stats->extend_node_visits = totals[2];
        stats->extend_triangle_tests = totals[3];
        stats->node_visits = totals[2] + totals[4];
        stats->triangle_tests = totals[3] + totals[5];
        stats->stack_overflows = totals[6];
        stats->extend_lane_slots = totals[7];
        stats->connect_lane_slots = totals[8];
        stats->elapsed_ms = ms;
        for (size_t k = 0; k < ev_kind.size(); ++k) {
            float t = 0.0f;
            LUCIDA_HIP(hipEventElapsedTime(&t, evs[2 * k], evs[2 * k + 1]), LUCIDA_VALIDATION);
            if (ev_kind[k] == 5) {
                stats->queue_sort_ms += t;
                stats->queue_sorts += 1;
            } else {
                stats->kernel_ms[ev_kind[k]] += t;
                stats->kernel_launches[ev_kind[k]] += 1;
            }
        }
    }
    return LUCIDA_OK;
}


lucida_status lucida_render_megakernel(lucida_ctx* ctx, const lucida_scene* scene, const GpuUniforms* uniforms,
                                       uint32_t samples, lucida_film* film, const lucida_wavefront_config* config,
                                       lucida_render_stats* stats) {
    if (!ctx || !scene || !uniforms || !film)
        return set_error(LUCIDA_INVALID_ARG, "lucida_render_megakernel: null argument");
    clear_error();
    if (uniforms->width != film->width || uniforms->height != film->height)
        return set_error(LUCIDA_INVALID_ARG, "uniforms width/height do not match the film");
    if (uniforms->max_depth > MAX_BOUNCES) return set_error(LUCIDA_INVALID_ARG, "max_depth above 64");
    if (scene->num_bvh_nodes >= (1u << 26) || scene->num_triangles + scene->num_primitives + scene->num_instances >= (1u << 26))
        return set_error(LUCIDA_INVALID_ARG, "scene too large: node and primitive indices are packed in 26 bits");
    if (stats) std::memset(stats, 0, sizeof *stats);
    if (config && !binary_width(*config)) return set_error(LUCIDA_INVALID_ARG, "bvh_width must be 0 or 2 (the binary walk)");
    if (samples == 0 || uniforms->max_depth == 0) return LUCIDA_OK;
    lucida_status st = ensure_ggx_table(ctx);
    if (st != LUCIDA_OK) return st;
    lucida_wavefront_config cfg{};
    if (config) cfg = *config;
    const size_t pixels = (size_t)film->width * film->height;
    // The wavefront's batches (about 2^23 paths, samples_per_batch if set):
    // each batch's per-pixel sums are added to the film as one, so the film
    // is summed exactly as lucida_render_wavefront sums it.
    uint32_t spb = cfg.samples_per_batch ? cfg.samples_per_batch
                                         : (uint32_t)std::max<size_t>(1, ((size_t)1 << 23) / pixels);
    spb = std::min(spb, samples);
    const uint32_t batches = (samples + spb - 1) / spb;
    const SceneView sc = make_view(ctx, scene);
    const int mode = scene->num_instances ? TRAV_INSTANCED : TRAV_BINARY;
    hipStream_t s = ctx->stream;
    LUCIDA_HIP(ctx->mk_spill.ensure((pixels + TRACE_WG) * (INST_STACK - LDS_STACK) * 4), LUCIDA_OOM);
    LUCIDA_HIP(ctx->mk_state.ensure(COUNTER_BYTES + (size_t)batches * sizeof(GpuUniforms)), LUCIDA_OOM);
    uint32_t* counters = ctx->mk_state.as<uint32_t>();
    GpuUniforms* dU = reinterpret_cast<GpuUniforms*>(ctx->mk_state.as<char>() + COUNTER_BYTES);
    std::vector<GpuUniforms> host_u(batches, *uniforms);
    for (uint32_t b = 0; b < batches; ++b) {
        host_u[b].sample_offset = uniforms->sample_offset + b * spb;
        host_u[b].samples_per_launch = std::min(spb, samples - b * spb);
    }
    LUCIDA_HIP(hipMemcpyAsync(dU, host_u.data(), batches * sizeof(GpuUniforms), hipMemcpyHostToDevice, s),
               LUCIDA_VALIDATION);
    const uint32_t stats_on = (stats && cfg.count_traversal) ? 1u : 0u;
    const Band whole{0u, (uint32_t)pixels};
    LUCIDA_HIP(hipEventRecord(ctx->ev0, s), LUCIDA_VALIDATION);
    hipLaunchKernelGGL(k_clear_counters, dim3(1), dim3(256), 0, s, counters, (uint32_t)TOTALS_AT + 24u);
    for (uint32_t b = 0; b < batches; ++b) {
        hipLaunchKernelGGL(k_clear_counters, dim3(1), dim3(256), 0, s, counters, (uint32_t)TOTALS_AT);
        const dim3 grid(grid_for(pixels, TRACE_WG)), block(TRACE_WG);
        if (mode == TRAV_INSTANCED)
            hipLaunchKernelGGL(k_megakernel<TRAV_INSTANCED>, grid, block, 0, s, sc, dU + b, film->accum, whole,
                               ctx->mk_spill.as<uint32_t>(), counters, stats_on);
        else
            hipLaunchKernelGGL(k_megakernel<TRAV_BINARY>, grid, block, 0, s, sc, dU + b, film->accum, whole,
                               ctx->mk_spill.as<uint32_t>(), counters, stats_on);
        hipLaunchKernelGGL(k_fold_counters, dim3(1), dim3(64), 0, s, counters, 1u);
        LUCIDA_HIP(hipGetLastError(), LUCIDA_SHADER);
    }
    LUCIDA_HIP(hipEventRecord(ctx->ev1, s), LUCIDA_VALIDATION);
    LUCIDA_HIP(hipEventSynchronize(ctx->ev1), LUCIDA_VALIDATION);
    LUCIDA_HIP(hipGetLastError(), LUCIDA_VALIDATION);
    if (stats) {
        uint64_t t[10];
        LUCIDA_HIP(hipMemcpy(t, counters + TOTALS_AT, sizeof t, hipMemcpyDeviceToHost), LUCIDA_READBACK);
        float ms = 0.0f;
        LUCIDA_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1), LUCIDA_VALIDATION);
        stats->extension_rays = t[0];
        stats->shadow_rays = t[1];
        stats->extend_node_visits = t[2];
        stats->extend_triangle_tests = t[3];
        stats->node_visits = t[2] + t[4];
        stats->triangle_tests = t[3] + t[5];
        stats->stack_overflows = t[6];
        stats->extend_lane_slots = t[7];
        stats->connect_lane_slots = t[8];
        stats->elapsed_ms = ms;
        stats->kernel_launches[1] = batches;
        stats->kernel_ms[1] = ms;
    }
    return LUCIDA_OK;
}

lucida_status lucida_scene_cast(lucida_ctx* ctx, const lucida_scene* scene, const float* rays, size_t n, int flags,
                                float* hits, uint32_t* instances) {
    const int any_hit = flags & LUCIDA_CAST_ANY_HIT;
    if (!ctx || !scene || (n && (!rays || !hits))) return set_error(LUCIDA_INVALID_ARG, "lucida_scene_cast: bad argument");
    if (flags & ~LUCIDA_CAST_ANY_HIT) return set_error(LUCIDA_INVALID_ARG, "lucida_scene_cast: unknown flag");
    if (n == 0) return LUCIDA_OK;
    if (n > (1u << 26)) return set_error(LUCIDA_INVALID_ARG, "lucida_scene_cast: more than 2^26 rays");
    clear_error();
    const SceneView sc = make_view(ctx, scene);
    DevBuf d_rays, d_hits, d_inst, d_spill;
    LUCIDA_HIP(d_rays.ensure(n * 32), LUCIDA_OOM);
    LUCIDA_HIP(d_hits.ensure(n * 16), LUCIDA_OOM);
    LUCIDA_HIP(d_inst.ensure(n * 4), LUCIDA_OOM);
    LUCIDA_HIP(d_spill.ensure((n + TRACE_WG) * (INST_STACK - LDS_STACK) * 4), LUCIDA_OOM);
    lucida_status st = LUCIDA_OK;
    if (hipMemcpyAsync(d_rays.ptr, rays, n * 32, hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
        st = set_error(LUCIDA_VALIDATION, "lucida_scene_cast: upload failed");
    if (st == LUCIDA_OK) {
        hipLaunchKernelGGL(k_cast, dim3(grid_for(n, TRACE_WG)), dim3(TRACE_WG), 0, ctx->stream, sc,
                           d_rays.as<float4>(), (uint32_t)n, any_hit ? 1u : 0u, d_hits.as<float4>(),
                           d_inst.as<uint32_t>(), d_spill.as<uint32_t>());
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(hits, d_hits.ptr, n * 16, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            (instances &&
             hipMemcpyAsync(instances, d_inst.ptr, n * 4, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess) ||
            hipStreamSynchronize(ctx->stream) != hipSuccess)
            st = set_error(LUCIDA_READBACK, "lucida_scene_cast: launch or readback failed");
    }
    (void)hipStreamSynchronize(ctx->stream);
    d_rays.release();
    d_hits.release();
    d_inst.release();
    d_spill.release();
    return st;
}

lucida_status lucida_render_batch(lucida_ctx* ctx, const lucida_job* jobs, size_t n,
                                  const lucida_wavefront_config* config, float* job_done_ms,
                                  lucida_render_stats* stats) {
    if (!ctx || (n && !jobs)) return set_error(LUCIDA_INVALID_ARG, "lucida_render_batch: null argument");
    clear_error();
    if (stats) std::memset(stats, 0, sizeof *stats);
    if (n == 0) return LUCIDA_OK;
    if (n > 65536) return set_error(LUCIDA_INVALID_ARG, "lucida_render_batch: more than 65,536 jobs");
    lucida_wavefront_config cfg{};
    if (config) cfg = *config;
    if (!binary_width(cfg)) return set_error(LUCIDA_INVALID_ARG, "lucida_render_batch: bvh_width must be 0 or 2");
    for (size_t j = 0; j < n; ++j) {
        const lucida_job& J = jobs[j];
        if (!J.scene || !J.film) return set_error(LUCIDA_INVALID_ARG, "lucida_render_batch: a job has no scene or film");
        if (J.uniforms.width != J.film->width || J.uniforms.height != J.film->height)
            return set_error(LUCIDA_INVALID_ARG, "lucida_render_batch: a job's uniforms do not match its film");
        if (J.uniforms.max_depth > MAX_BOUNCES) return set_error(LUCIDA_INVALID_ARG, "max_depth above 64");
        if (J.scene->num_bvh_nodes >= (1u << 26) ||
            J.scene->num_triangles + J.scene->num_primitives + J.scene->num_instances >= (1u << 26))
            return set_error(LUCIDA_INVALID_ARG, "scene too large: node and primitive indices are packed in 26 bits");
        for (size_t k = 0; k < j; ++k)
            if (jobs[k].film == J.film) return set_error(LUCIDA_INVALID_ARG, "lucida_render_batch: two jobs share a film");
    }
    lucida_status st = ensure_ggx_table(ctx);
    if (st != LUCIDA_OK) return st;

    // Each job is rendered in chunks of the single render's batch size, so
    // its film is summed exactly as a lucida_render_wavefront call would.
    std::vector<uint32_t> spb(n), chunks(n), next(n, 0), last_round(n, 0);
    std::vector<size_t> pixels(n);
    for (size_t j = 0; j < n; ++j) {
        pixels[j] = (size_t)jobs[j].uniforms.width * jobs[j].uniforms.height;
        uint32_t b = cfg.samples_per_batch ? cfg.samples_per_batch
                                           : (uint32_t)std::max<size_t>(1, ((size_t)1 << 23) / pixels[j]);
        spb[j] = std::max(1u, std::min(b, jobs[j].samples));
        chunks[j] = (jobs[j].samples == 0 || jobs[j].uniforms.max_depth == 0)
                        ? 0u
                        : (jobs[j].samples + spb[j] - 1) / spb[j];
    }
    const size_t cap = cfg.max_paths ? cfg.max_paths : ((size_t)1 << 24);
    struct Round {
        size_t first, count, paths, pixels;
        uint32_t depth;
        int solo_mode;  // -1: mixed round (batch kernels); else the traversal mode of its one chunk
        bool sort;      // a solo round that sorts its extension queues (ray_order)
    };
    std::vector<Round> rounds;
    std::vector<BatchJob> table;
    std::vector<SceneView> views(n);
    std::vector<int> solo_mode(n);
    bool any_instanced = false;
    for (size_t j = 0; j < n; ++j) {
        views[j] = make_view(ctx, jobs[j].scene);
        any_instanced = any_instanced || jobs[j].scene->num_instances > 0;
        solo_mode[j] = jobs[j].scene->num_instances ? TRAV_INSTANCED : TRAV_BINARY;
    }
    const int mode = any_instanced ? TRAV_INSTANCED : TRAV_BINARY;
    // A job whose chunks hold kSoloPaths paths or more fills the device by
    // itself: its chunks run as rounds of their own with the single-render
    // kernels (no per-path job lookup). The other jobs' chunks are packed
    // into mixed rounds, one chunk per job per round, and go first, so small
    // requests finish early; the large jobs follow, shortest first.
    const size_t kSoloPaths = (size_t)1 << 21;
    std::vector<char> solo(n, 0);
    for (size_t j = 0; j < n; ++j) solo[j] = chunks[j] > 0 && pixels[j] * spb[j] >= kSoloPaths;
    auto add_chunk = [&](size_t j, Round& r) {
        const uint32_t done = next[j] * spb[j];
        const uint32_t this_spb = std::min(spb[j], jobs[j].samples - done);
        const size_t paths = pixels[j] * this_spb;
        BatchJob b{};
        b.sc = views[j];
        b.u = jobs[j].uniforms;
        b.u.sample_offset = jobs[j].uniforms.sample_offset + done;
        b.u.samples_per_launch = this_spb;
        b.film = jobs[j].film->accum;
        b.path_base = (uint32_t)r.paths;
        b.pixel_base = (uint32_t)r.pixels;
        b.pixels = (uint32_t)pixels[j];
        table.push_back(b);
        r.paths += paths;
        r.pixels += pixels[j];
        r.depth = std::max(r.depth, jobs[j].uniforms.max_depth);
        ++r.count;
        ++next[j];
        if (next[j] == chunks[j]) last_round[j] = (uint32_t)rounds.size();
    };
    for (;;) {
        Round r{table.size(), 0, 0, 0, 0, -1, false};
        for (size_t j = 0; j < n; ++j) {
            if (solo[j] || next[j] >= chunks[j]) continue;
            const uint32_t this_spb = std::min(spb[j], jobs[j].samples - next[j] * spb[j]);
            if (r.count > 0 && r.paths + pixels[j] * this_spb > cap) continue;
            add_chunk(j, r);
        }
        if (r.count == 0) break;
        if (r.paths >= 0x7FFFFFFFull / 4) return set_error(LUCIDA_INVALID_ARG, "a job's chunk exceeds 2^29 paths");
        // sorted when some job's tree qualifies and the job index fits the key's 8 bits
        for (size_t k = r.first; k < r.first + r.count && r.count <= 256; ++k)
            r.sort = r.sort || want_ray_order(cfg, table[k].sc, r.paths);
        rounds.push_back(r);
    }
    std::vector<size_t> order;
    for (size_t j = 0; j < n; ++j)
        if (solo[j]) order.push_back(j);
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
        return pixels[a] * jobs[a].samples * jobs[a].uniforms.max_depth <
               pixels[b] * jobs[b].samples * jobs[b].uniforms.max_depth;
    });
    for (size_t j : order) {
        while (next[j] < chunks[j]) {
            Round r{table.size(), 0, 0, 0, 0, solo_mode[j], false};
            add_chunk(j, r);
            r.sort = want_ray_order(cfg, views[j], r.paths);
            if (r.paths >= 0x7FFFFFFFull / 4) return set_error(LUCIDA_INVALID_ARG, "a job's chunk exceeds 2^29 paths");
            rounds.push_back(r);
        }
    }
    if (rounds.empty()) return LUCIDA_OK;
    size_t max_paths = 0;
    for (const Round& r : rounds) max_paths = std::max(max_paths, r.paths);

    // Rounds alternate between the context's two pipelines (as the batches of
    // a single render do); their RESOLVEs run in round order, so each film
    // receives its job's chunks in order. A batch captured in a graph keeps
    // to one pipeline: captured with two, the process crashed (host
    // segmentation fault, cause not found).
    const int npipes = (rounds.size() >= 2 && !cfg.use_graph) ? 2 : 1;
    Planes p_fields[2], p_rec[2];  // state as arrays of fields, and as records (sorted rounds)
    for (int q = 0; q < npipes; ++q) {
        GpuUniforms* dU = nullptr;
        st = ensure_wavefront(ctx->wf[q], max_paths, &p_fields[q], &dU, false);
        if (st != LUCIDA_OK) return st;
        p_rec[q] = p_fields[q];
        set_layout(ctx->wf[q], &p_rec[q], true);
    }
    DevBuf d_table, d_stamps;
    LUCIDA_HIP(d_table.ensure(table.size() * sizeof(BatchJob)), LUCIDA_OOM);
    LUCIDA_HIP(d_stamps.ensure((rounds.size() + 1) * 8), LUCIDA_OOM);
    hipStream_t s0 = ctx->stream;  // pipeline 0's; the graph is captured from it
    st = LUCIDA_OK;
    if (hipMemcpyAsync(d_table.ptr, table.data(), table.size() * sizeof(BatchJob), hipMemcpyHostToDevice, s0) !=
        hipSuccess)
        st = set_error(LUCIDA_VALIDATION, "lucida_render_batch: job table upload failed");
    const BatchJob* dt = d_table.as<BatchJob>();
    unsigned long long* stamps = d_stamps.as<unsigned long long>();
    const uint32_t stats_on = (stats && cfg.count_traversal) ? 1u : 0u;
    const SceneView none{};

    // Queues every round; returns the first failing status (a HIP call, a
    // launch or a sort), so an error under capture comes back as a status.
    auto enqueue_all = [&]() -> lucida_status {
        hipError_t first = hipSuccess;
        auto chk = [&](hipError_t e) {
            if (e != hipSuccess && first == hipSuccess) first = e;
        };
        lucida_status sort_st = LUCIDA_OK;
        auto chk_sort = [&](lucida_status v) {
            if (v != LUCIDA_OK && sort_st == LUCIDA_OK) sort_st = v;
        };
        hipLaunchKernelGGL(k_stamp, dim3(1), dim3(64), 0, s0, stamps, 0u);
        for (int q = 0; q < npipes; ++q) {
            if (q > 0) {  // fork the other pipeline off pipeline 0's stream
                chk(hipEventRecord(ctx->wf[q].ev_sorted, s0));
                chk(hipStreamWaitEvent(ctx->wf[q].s, ctx->wf[q].ev_sorted, 0));
            }
            hipLaunchKernelGGL(k_clear_counters, dim3(1), dim3(256), 0, ctx->wf[q].s, p_fields[q].counters,
                               (uint32_t)TOTALS_AT + 24u);
        }
        for (size_t ri = 0; ri < rounds.size(); ++ri) {
            const Round& r = rounds[ri];
            const int q = (int)(ri % npipes);
            WfPipe& w = ctx->wf[q];
            hipStream_t s = w.s, s2 = w.s2;
            const Planes& p = r.sort ? p_rec[q] : p_fields[q];  // every round rewrites all path state
            const BatchJob* jr = dt + r.first;
            const uint32_t nj = (uint32_t)r.count, P = (uint32_t)r.paths;
            hipLaunchKernelGGL(k_clear_counters, dim3(1), dim3(256), 0, s, p.counters, (uint32_t)TOTALS_AT);
            if (r.solo_mode >= 0) {
                // One chunk: exactly the launches of a single render's batch,
                // with the chunk's uniforms read from its table entry.
                const GpuUniforms* uj =
                    reinterpret_cast<const GpuUniforms*>(reinterpret_cast<const char*>(jr) + offsetof(BatchJob, u));
                const SceneView& sc = table[r.first].sc;
                const Band whole{0u, (uint32_t)r.pixels};
                hipLaunchKernelGGL(k_generate, dim3(grid_for(P, 256)), dim3(256), 0, s, uj, p, whole);
                uint32_t* qin = p.q0;
                for (uint32_t b = 0; b < r.depth; ++b) {
                    uint32_t* qout = qin == p.q1 ? p.q0 : p.q1;
                    uint32_t* qnext = qout;
                    uint32_t* order = nullptr;
                    launch_extend<false>(r.solo_mode, dim3(grid_for(P, TRACE_WG)), s, sc, nullptr, p, qin,
                                         p.counters + b, stats_on);
                    if (b > 0) chk(hipStreamWaitEvent(s, w.ev_connected, 0));
                    hipLaunchKernelGGL(k_shade<false>, dim3(grid_for(P, SHADE_WG)), dim3(SHADE_WG), 0, s, sc, uj,
                                       nullptr, p, qin, p.counters + b, qout, p.counters + b + 1,
                                       p.counters + SHADOW_AT + b, r.sort ? p.qkey : nullptr, whole);
                    chk(hipEventRecord(w.ev_shaded, s));
                    if (r.sort)
                        chk_sort(sort_queues(w, w.s2, p, qout, P, b, b + 1 < r.depth, RAY_KEY_BITS, &qnext, &order));
                    else
                        chk(hipStreamWaitEvent(s2, w.ev_shaded, 0));
                    launch_connect<false>(r.solo_mode, dim3(grid_for(P, TRACE_WG)), s2, sc, nullptr, p,
                                          p.counters + SHADOW_AT + b, order, stats_on);
                    chk(hipEventRecord(w.ev_connected, s2));
                    qin = qnext;
                }
                if (r.depth > 0) chk(hipStreamWaitEvent(s, w.ev_connected, 0));
                if (npipes > 1 && ri > 0) chk(hipStreamWaitEvent(s, ctx->ev_resolved, 0));
                hipLaunchKernelGGL(k_resolve, dim3(grid_for(r.pixels, 256)), dim3(256), 0, s, uj, p,
                                   table[r.first].film, whole);
                if (npipes > 1) chk(hipEventRecord(ctx->ev_resolved, s));
                hipLaunchKernelGGL(k_fold_counters, dim3(1), dim3(64), 0, s, p.counters, r.depth);
                hipLaunchKernelGGL(k_stamp, dim3(1), dim3(64), 0, s, stamps, (uint32_t)(ri + 1));
                chk(hipGetLastError());
                continue;
            }
            hipLaunchKernelGGL(k_generate_batch, dim3(grid_for(P, 256)), dim3(256), 0, s, jr, nj, P, p);
            // a sorted mixed round orders by (job, ray key): 8 + 16 bits
            uint32_t* qin = p.q0;
            for (uint32_t b = 0; b < r.depth; ++b) {
                uint32_t* qout = qin == p.q1 ? p.q0 : p.q1;
                uint32_t* qnext = qout;
                uint32_t* order = nullptr;
                launch_extend<true>(mode, dim3(grid_for(P, TRACE_WG)), s, none, jr, p, qin, p.counters + b, stats_on);
                if (b > 0) chk(hipStreamWaitEvent(s, w.ev_connected, 0));
                hipLaunchKernelGGL(k_shade<true>, dim3(grid_for(P, SHADE_WG)), dim3(SHADE_WG), 0, s, none, nullptr, jr,
                                   p, qin, p.counters + b, qout, p.counters + b + 1, p.counters + SHADOW_AT + b,
                                   r.sort ? p.qkey : nullptr, Band{0u, 0u});
                chk(hipEventRecord(w.ev_shaded, s));
                if (r.sort)
                    chk_sort(sort_queues(w, w.s2, p, qout, P, b, b + 1 < r.depth, 24, &qnext, &order));
                else
                    chk(hipStreamWaitEvent(s2, w.ev_shaded, 0));
                launch_connect<true>(mode, dim3(grid_for(P, TRACE_WG)), s2, none, jr, p, p.counters + SHADOW_AT + b,
                                     order, stats_on);
                chk(hipEventRecord(w.ev_connected, s2));
                qin = qnext;
            }
            if (r.depth > 0) chk(hipStreamWaitEvent(s, w.ev_connected, 0));
            if (npipes > 1 && ri > 0) chk(hipStreamWaitEvent(s, ctx->ev_resolved, 0));
            hipLaunchKernelGGL(k_resolve_batch, dim3(grid_for(r.pixels, 256)), dim3(256), 0, s, jr, nj,
                               (uint32_t)r.pixels, p);
            if (npipes > 1) chk(hipEventRecord(ctx->ev_resolved, s));
            hipLaunchKernelGGL(k_fold_counters, dim3(1), dim3(64), 0, s, p.counters, r.depth);
            hipLaunchKernelGGL(k_stamp, dim3(1), dim3(64), 0, s, stamps, (uint32_t)(ri + 1));
            chk(hipGetLastError());
        }
        for (int q = 1; q < npipes; ++q) {  // join: pipeline 0's stream waits for the others
            chk(hipEventRecord(ctx->wf[q].ev_shaded, ctx->wf[q].s));
            chk(hipStreamWaitEvent(s0, ctx->wf[q].ev_shaded, 0));
        }
        if (sort_st != LUCIDA_OK) return sort_st;
        if (first != hipSuccess) {
            char msg[256];
            snprintf(msg, sizeof msg, "lucida_render_batch: queueing the rounds failed: %s", hipGetErrorString(first));
            return set_error(LUCIDA_VALIDATION, msg);
        }
        return LUCIDA_OK;
    };

    if (st == LUCIDA_OK) {
        LUCIDA_HIP(hipEventRecord(ctx->ev0, s0), LUCIDA_VALIDATION);
        if (cfg.use_graph) {
            // One graph for the whole batch: every launch above is a kernel
            // node (no memset, no host copy), replayed with one launch.
            hipGraph_t graph = nullptr;
            hipGraphExec_t exec = nullptr;
            if (hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal) != hipSuccess) {
                st = set_error(LUCIDA_VALIDATION, "lucida_render_batch: stream capture failed to start");
            } else {
                const lucida_status qs = enqueue_all();
                // the capture is ended whatever happened, so the stream leaves capture mode
                const hipError_t ec = hipStreamEndCapture(s0, &graph);
                if (qs != LUCIDA_OK)
                    st = qs;
                else if (ec != hipSuccess || !graph)
                    st = set_error(LUCIDA_VALIDATION, "lucida_render_batch: stream capture failed");
                else if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess)
                    st = set_error(LUCIDA_VALIDATION, "lucida_render_batch: graph instantiation failed");
                else if (hipGraphLaunch(exec, s0) != hipSuccess)
                    st = set_error(LUCIDA_VALIDATION, "lucida_render_batch: graph launch failed");
            }
            (void)hipStreamSynchronize(s0);
            if (exec) (void)hipGraphExecDestroy(exec);
            if (graph) (void)hipGraphDestroy(graph);
        } else {
            st = enqueue_all();
        }
        LUCIDA_HIP(hipEventRecord(ctx->ev1, s0), LUCIDA_VALIDATION);
        LUCIDA_HIP(hipEventSynchronize(ctx->ev1), LUCIDA_VALIDATION);
    }
    if (st == LUCIDA_OK) LUCIDA_HIP(hipGetLastError(), LUCIDA_VALIDATION);
    if (st == LUCIDA_OK && (stats || job_done_ms)) {
        std::vector<unsigned long long> h_stamps(rounds.size() + 1);
        uint64_t totals[10] = {};
        LUCIDA_HIP(hipMemcpy(h_stamps.data(), stamps, h_stamps.size() * 8, hipMemcpyDeviceToHost), LUCIDA_READBACK);
        for (int q = 0; q < npipes; ++q) {
            uint64_t t[10];
            LUCIDA_HIP(hipMemcpy(t, p_fields[q].counters + TOTALS_AT, sizeof t, hipMemcpyDeviceToHost),
                       LUCIDA_READBACK);
            for (int k = 0; k < 10; ++k) totals[k] += t[k];
        }
        int rate_khz = 0;
        LUCIDA_HIP(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, ctx->device), LUCIDA_VALIDATION);
        if (job_done_ms)
            for (size_t j = 0; j < n; ++j)
                job_done_ms[j] = chunks[j] == 0 || rate_khz <= 0
                                     ? 0.0f
                                     : (float)((double)(h_stamps[last_round[j] + 1] - h_stamps[0]) / rate_khz);
        if (stats) {
            float ms = 0.0f;
            LUCIDA_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1), LUCIDA_VALIDATION);
            stats->extension_rays = totals[0];
            stats->shadow_rays = totals[1];
            stats->extend_node_visits = totals[2];
            stats->extend_triangle_tests = totals[3];
            stats->node_visits = totals[2] + totals[4];
            stats->triangle_tests = totals[3] + totals[5];
            stats->stack_overflows = totals[6];
            stats->extend_lane_slots = totals[7];
            stats->connect_lane_slots = totals[8];
            stats->elapsed_ms = ms;
            stats->kernel_launches[0] = (uint32_t)rounds.size();
        }
    }
    d_table.release();
    d_stamps.release();
    return st;
}

}  // extern "C"
