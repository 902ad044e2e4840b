// Device-side shading for the wavefront tracer: counter-based samplers, frame
// math, GGX / rough-dielectric BSDF with energy compensation, area lights,
// environment map, and the analytic primitives. Every function follows the
// reference's CPU twin in operation order (the file is compiled with
// contraction off), so the GPU and the CPU draw identical random streams and
// differ only where a transcendental comes from a different libm.
//
// Reference: rng.rs:28-85, sobol.rs:147-361, math.rs:22-176,
// camera.rs:84-135, scene.rs:76-140 and 405-560, bsdf.rs:23-700 and
// 1178-1662, light.rs:150-335, envmap.rs:54-320.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/lucida.h"

namespace lucida {
namespace dev {

constexpr float PI = 3.14159265358979323846f;
constexpr float TAU = 6.28318530717958647692f;
constexpr float INV_PI = 0.318309886183790671538f;
constexpr float T_MAX = 1.0e30f;
constexpr float T_MIN = 1.0e-4f;

struct V3 {
    float x, y, z;
};
__device__ inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ inline V3 v3s(float s) { return V3{s, s, s}; }
__device__ inline V3 ld3(const float* p) { return V3{p[0], p[1], p[2]}; }
__device__ inline V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ inline V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline V3 operator*(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ inline V3 operator/(V3 a, V3 b) { return V3{a.x / b.x, a.y / b.y, a.z / b.z}; }
__device__ inline V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
__device__ inline V3 operator*(float s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }
__device__ inline V3 operator/(V3 a, float s) { return V3{a.x / s, a.y / s, a.z / s}; }
__device__ inline V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
__device__ inline bool is_zero(V3 a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }
__device__ inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline V3 cross(V3 a, V3 b) {
    return V3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
__device__ inline float length(V3 a) { return sqrtf(dot(a, a)); }
__device__ inline V3 normalize(V3 a) { return a * (1.0f / length(a)); }
__device__ inline V3 normalize_or(V3 a, V3 fb) {
    float r = 1.0f / length(a);
    return (isfinite(r) && r > 0.0f) ? a * r : fb;
}
__device__ inline V3 vmax(V3 a, V3 b) { return V3{fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)}; }
__device__ inline V3 vmin(V3 a, V3 b) { return V3{fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)}; }
__device__ inline float max3(V3 a) { return fmaxf(fmaxf(a.x, a.y), a.z); }
__device__ inline float clampf(float v, float lo, float hi) {
    if (v < lo) v = lo;
    if (v > hi) v = hi;
    return v;
}
__device__ inline float luminance(V3 c) { return dot(c, V3{0.2126f, 0.7152f, 0.0722f}); }
__device__ inline V3 f4xyz(float4 v) { return V3{v.x, v.y, v.z}; }

// --- samplers (rng.rs, sobol.rs) ------------------------------------------------

__device__ inline uint32_t pcg_hash(uint32_t input) {
    uint32_t state = input * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28) + 4u)) ^ state) * 277803737u;
    return (word >> 22) ^ word;
}
__device__ inline uint32_t rng_seed(uint32_t frame, uint32_t pixel, uint32_t sample) {
    uint32_t h = pcg_hash(frame);
    h = pcg_hash(h + pixel);
    return pcg_hash(h + sample);
}
__device__ inline uint32_t hash_u32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
__device__ inline uint32_t owen_scramble(uint32_t x, uint32_t seed) {
    x = __builtin_bitreverse32(x);
    x += seed;
    x ^= x * 0x6c50b47cu;
    x ^= x * 0xb82f1e52u;
    x ^= x * 0xc7afe638u;
    x ^= x * 0x8d22f6e6u;
    return __builtin_bitreverse32(x);
}
// Sobol dimension 1 (sobol.rs:88-140: degree-1 polynomial, m = {1}): its
// direction numbers are v_1 = 2^31, v_i = v_{i-1} ^ (v_{i-1} >> 1).
__device__ inline uint32_t sobol_dim1(uint32_t index) {
    uint32_t x = 0, v = 0x80000000u;
#pragma unroll
    for (int k = 0; k < 32; ++k) {
        if ((index >> k) & 1u) x ^= v;
        v ^= v >> 1;
    }
    return x;
}
__device__ inline float sobol_sample(uint32_t index, uint32_t dim, uint32_t seed) {
    uint32_t group = dim >> 1, base = dim & 1u;
    uint32_t group_seed = hash_u32(seed ^ (group * 0x9e3779b9u));
    uint32_t shuffled = owen_scramble(index, group_seed);
    uint32_t v = base == 0 ? __builtin_bitreverse32(shuffled) : sobol_dim1(shuffled);
    uint32_t s = owen_scramble(v, hash_u32(group_seed + base + 1u));
    return (float)(s >> 8) * (1.0f / 16777216.0f);
}

// One path's sampler: the independent PCG stream or Owen-scrambled Sobol, in
// the draw order of sobol.rs Sampler::{next_f32, next_vec2}.
struct Sampler {
    uint32_t state;  // PCG state (independent)
    uint32_t index;  // sample index (Sobol)
    uint32_t seed;   // per-pixel scramble (Sobol)
    uint32_t dim;    // next dimension (Sobol)
    bool sobol;
    __device__ inline float next_f32() {
        if (!sobol) {
            state = state * 747796405u + 2891336453u;
            uint32_t word = ((state >> ((state >> 28) + 4u)) ^ state) * 277803737u;
            uint32_t u = (word >> 22) ^ word;
            return (float)(u >> 8) * (1.0f / 16777216.0f);
        }
        float v = sobol_sample(index, dim, seed);
        dim += 1;
        return v;
    }
    __device__ inline void next_vec2(float* x, float* y) {
        if (!sobol) {
            *x = next_f32();
            *y = next_f32();
            return;
        }
        dim = (dim + 1u) & ~1u;
        *x = sobol_sample(index, dim, seed);
        *y = sobol_sample(index, dim + 1, seed);
        dim += 2;
    }
};
__device__ inline Sampler make_sampler(const GpuUniforms& u, uint32_t pixel, uint32_t sample) {
    Sampler s;
    s.sobol = u.sampler_kind == LUCIDA_SAMPLER_SOBOL;
    s.state = rng_seed(u.frame_seed, pixel, sample);
    s.index = sample;
    s.seed = hash_u32(pixel + u.frame_seed * 0x9e3779b9u);
    s.dim = 0;
    return s;
}

// --- math.rs --------------------------------------------------------------------

__device__ inline void onb(V3 n, V3* t, V3* b) {
    float sign = copysignf(1.0f, n.z);
    float a = -1.0f / (sign + n.z);
    float bb = n.x * n.y * a;
    *t = V3{1.0f + sign * n.x * n.x * a, sign * bb, -sign * n.x};
    *b = V3{bb, sign + n.y * n.y * a, -n.y};
}
__device__ inline V3 to_world(V3 l, V3 n) {
    V3 t, b;
    onb(n, &t, &b);
    return t * l.x + b * l.y + n * l.z;
}
__device__ inline V3 to_local(V3 w, V3 n) {
    V3 t, b;
    onb(n, &t, &b);
    return V3{dot(w, t), dot(w, b), dot(w, n)};
}
__device__ inline void concentric_disk(float ux, float uy, float* dx, float* dy) {
    float ox = 2.0f * ux - 1.0f, oy = 2.0f * uy - 1.0f;
    if (ox == 0.0f && oy == 0.0f) {
        *dx = 0.0f;
        *dy = 0.0f;
        return;
    }
    float r, theta;
    if (fabsf(ox) > fabsf(oy)) {
        r = ox;
        theta = (PI / 4.0f) * (oy / ox);
    } else {
        r = oy;
        theta = PI / 2.0f - (PI / 4.0f) * (ox / oy);
    }
    *dx = r * cosf(theta);
    *dy = r * sinf(theta);
}
__device__ inline V3 cosine_hemisphere(float ux, float uy) {
    float dx, dy;
    concentric_disk(ux, uy, &dx, &dy);
    float z = sqrtf(fmaxf(1.0f - dx * dx - dy * dy, 0.0f));
    return V3{dx, dy, z};
}
__device__ inline int32_t f2i_sat(float v) {
    if (v != v) return 0;
    if (v >= 2147483648.0f) return 2147483647;
    if (v <= -2147483648.0f) return (-2147483647 - 1);
    return (int32_t)v;
}
__device__ inline float bits_add(float p, int32_t of) {
    uint32_t b = __float_as_uint(p);
    b = b + (uint32_t)(p < 0.0f ? -of : of);
    return __uint_as_float(b);
}
// math.rs:118 — offset a ray origin off a surface by a few ULPs along n.
__device__ inline V3 offset_ray_origin(V3 p, V3 n) {
    const float ORIGIN = 1.0f / 32.0f, FLOAT_SCALE = 1.0f / 65536.0f, INT_SCALE = 256.0f;
    int32_t ox = f2i_sat(INT_SCALE * n.x), oy = f2i_sat(INT_SCALE * n.y), oz = f2i_sat(INT_SCALE * n.z);
    float px = bits_add(p.x, ox), py = bits_add(p.y, oy), pz = bits_add(p.z, oz);
    return V3{fabsf(p.x) < ORIGIN ? p.x + FLOAT_SCALE * n.x : px, fabsf(p.y) < ORIGIN ? p.y + FLOAT_SCALE * n.y : py,
              fabsf(p.z) < ORIGIN ? p.z + FLOAT_SCALE * n.z : pz};
}

// --- BSDF (bsdf.rs) ---------------------------------------------------------------

constexpr float MIN_ALPHA = 2.0e-3f;
__device__ inline float roughness_to_alpha(float r) { return fmaxf(r * r, MIN_ALPHA); }
__device__ inline float ggx_d(float alpha, float cos_m) {
    if (cos_m <= 0.0f) return 0.0f;
    float a2 = alpha * alpha, c2 = cos_m * cos_m;
    float sin2 = fmaxf(1.0f - c2, 0.0f);
    float denom = a2 * c2 + sin2;
    return a2 / (PI * denom * denom);
}
__device__ inline float smith_lambda(float alpha, float cos_theta) {
    float c = fabsf(cos_theta);
    if (c >= 1.0f) return 0.0f;
    float c2 = c * c;
    float tan2 = (1.0f - c2) / fmaxf(c2, 1.0e-12f);
    return 0.5f * (-1.0f + sqrtf(1.0f + alpha * alpha * tan2));
}
__device__ inline float smith_g1(float alpha, float c) { return 1.0f / (1.0f + smith_lambda(alpha, c)); }
__device__ inline float smith_g2(float alpha, float co, float ci) {
    return 1.0f / (1.0f + smith_lambda(alpha, co) + smith_lambda(alpha, ci));
}
__device__ inline V3 sample_ggx_vndf(V3 wo, float alpha, float ux, float uy) {
    V3 v = normalize(V3{alpha * wo.x, alpha * wo.y, wo.z});
    float lensq = v.x * v.x + v.y * v.y;
    V3 t1 = lensq > 0.0f ? V3{-v.y, v.x, 0.0f} / sqrtf(lensq) : V3{1.0f, 0.0f, 0.0f};
    V3 t2 = cross(v, t1);
    float r = sqrtf(ux);
    float phi = 2.0f * PI * uy;
    float p1 = r * cosf(phi);
    float p2 = r * sinf(phi);
    float s = 0.5f * (1.0f + v.z);
    p2 = (1.0f - s) * sqrtf(fmaxf(1.0f - p1 * p1, 0.0f)) + s * p2;
    V3 nh = p1 * t1 + p2 * t2 + sqrtf(fmaxf(1.0f - p1 * p1 - p2 * p2, 0.0f)) * v;
    return normalize(V3{alpha * nh.x, alpha * nh.y, fmaxf(nh.z, 0.0f)});
}
__device__ inline float ggx_vndf_pdf(float alpha, V3 wo, V3 m) {
    float c = wo.z;
    if (c <= 0.0f) return 0.0f;
    return smith_g1(alpha, c) * ggx_d(alpha, m.z) / (4.0f * c);
}
__device__ inline float ggx_vndf_half_pdf(float alpha, V3 wo, V3 m) {
    float co = wo.z, com = dot(wo, m);
    if (co <= 0.0f || com <= 0.0f) return 0.0f;
    return smith_g1(alpha, co) * ggx_d(alpha, fabsf(m.z)) * com / co;
}
__device__ inline V3 fresnel_schlick(V3 f0, float c) {
    float m = clampf(1.0f - c, 0.0f, 1.0f);
    float m2 = m * m;
    return f0 + (v3s(1.0f) - f0) * (m2 * m2 * m);
}
__device__ inline float fresnel_dielectric(float cos_i, float eta) {
    float c = fabsf(clampf(cos_i, -1.0f, 1.0f));
    float sin2_t = (1.0f - c * c) / (eta * eta);
    if (sin2_t >= 1.0f) return 1.0f;
    float cos_t = sqrtf(fmaxf(1.0f - sin2_t, 0.0f));
    float rpar = (eta * c - cos_t) / (eta * c + cos_t);
    float rper = (c - eta * cos_t) / (c + eta * cos_t);
    return 0.5f * (rpar * rpar + rper * rper);
}
__device__ inline float f0_from_ior(float ior) {
    float r = (ior - 1.0f) / (ior + 1.0f);
    return r * r;
}
__device__ inline V3 vsqrt_pos(V3 a) {
    return V3{powf(fmaxf(a.x, 0.0f), 0.5f), powf(fmaxf(a.y, 0.0f), 0.5f), powf(fmaxf(a.z, 0.0f), 0.5f)};
}
__device__ inline V3 fresnel_conductor(float cos_i, V3 eta, V3 k) {
    float c = clampf(cos_i, 0.0f, 1.0f);
    float c2 = c * c;
    float sin2 = 1.0f - c2;
    V3 eta2 = eta * eta, k2 = k * k;
    V3 t0 = eta2 - k2 - v3s(sin2);
    V3 a2pb2 = vsqrt_pos(t0 * t0 + 4.0f * eta2 * k2);
    V3 t1 = a2pb2 + v3s(c2);
    V3 a = vsqrt_pos(0.5f * (a2pb2 + t0));
    V3 t2 = 2.0f * c * a;
    V3 rs = (t1 - t2) / (t1 + t2);
    V3 t3 = c2 * a2pb2 + v3s(sin2 * sin2);
    V3 t4 = t2 * sin2;
    V3 rp = rs * (t3 - t4) / (t3 + t4);
    return 0.5f * (rs + rp);
}
__device__ inline V3 reflect(V3 v, V3 n) { return -v + 2.0f * dot(v, n) * n; }
__device__ inline bool refract(V3 wo, V3 m, float eta, V3* out) {
    float cos_i = dot(wo, m);
    float sin2_i = fmaxf(1.0f - cos_i * cos_i, 0.0f);
    float sin2_t = sin2_i / (eta * eta);
    if (sin2_t >= 1.0f) return false;
    float cos_t = sqrtf(1.0f - sin2_t);
    V3 wt = -wo / eta + (cos_i / eta - copysignf(cos_t, cos_i)) * m;
    *out = normalize_or(wt, v3s(0.0f));
    return true;
}

// The specular single-scattering sample used to build the energy table
// (bsdf.rs specular::sample_single with white Fresnel).
__device__ inline float specular_single_weight_white(float alpha, V3 wo, float ux, float uy) {
    if (wo.z <= 0.0f) return 0.0f;
    V3 m = sample_ggx_vndf(wo, alpha, ux, uy);
    V3 wi = reflect(wo, m);
    if (wi.z <= 0.0f) return 0.0f;
    float pdf = ggx_vndf_pdf(alpha, wo, m);
    if (pdf <= 0.0f) return 0.0f;
    V3 f = fresnel_schlick(v3s(1.0f), fmaxf(dot(wo, m), 0.0f));
    float g1 = smith_g1(alpha, wo.z);
    float g2 = smith_g2(alpha, wo.z, wi.z);
    return f.x * (g2 / g1);
}

// energy::lookup (bsdf.rs:1236)
__device__ inline float energy_lookup(const float* __restrict__ table, float roughness, float cos_o) {
    const int n = 32;
    float fx = clampf(roughness, 0.0f, 1.0f) * (float)(n - 1);
    float fy = sqrtf(clampf(cos_o, 0.0f, 1.0f)) * (float)(n - 1);
    int x0 = min((int)fx, n - 1), y0 = min((int)fy, n - 1);
    int x1 = min(x0 + 1, n - 1), y1 = min(y0 + 1, n - 1);
    float tx = fx - (float)x0, ty = fy - (float)y0;
    float a = table[x0 * n + y0], b = table[x0 * n + y1], c = table[x1 * n + y0], d = table[x1 * n + y1];
    float top = a + (b - a) * ty;
    float bot = c + (d - c) * ty;
    return clampf(top + (bot - top) * tx, 1.0e-3f, 1.0f);
}
__device__ inline V3 fresnel_average(V3 f0) { return f0 + (v3s(1.0f) - f0) * (1.0f / 21.0f); }

struct Surface {
    V3 diffuse_albedo, f0, eta, k, tint;
    float roughness, transmission, relative_ior;
    bool conductor;
    const float* table;

    __device__ inline V3 fresnel(float c) const {
        return conductor ? fresnel_conductor(c, eta, k) : fresnel_schlick(f0, c);
    }
    __device__ inline V3 normal_reflectance() const { return conductor ? fresnel_conductor(1.0f, eta, k) : f0; }
    __device__ inline V3 compensation(float cos_o, V3 f0n) const {
        float e = energy_lookup(table, roughness, cos_o);
        return v3s(1.0f) + fresnel_average(f0n) * ((1.0f - e) / e);
    }
    __device__ inline V3 specular_albedo(float cos_o) const {
        float e_ss = energy_lookup(table, roughness, fabsf(cos_o));
        V3 f = fresnel(fabsf(cos_o));
        V3 f_avg = fresnel_average(normal_reflectance());
        return vmin(vmax(f * e_ss + f_avg * (1.0f - e_ss), v3s(0.0f)), v3s(1.0f));
    }
    // specular_probability(cos_o), given sa = specular_albedo(cos_o)
    __device__ inline float specular_probability(float cos_o, V3 sa) const {
        if (transmission > 0.0f) {
            float f = fresnel_dielectric(fabsf(cos_o), relative_ior);
            if (f >= 1.0f) return 1.0f;
            return clampf(f, 0.05f, 0.95f);
        }
        float spec = luminance(sa);
        float diff = luminance(diffuse_albedo * (v3s(1.0f) - sa));
        if (spec + diff <= 0.0f) return 1.0f;
        return clampf(spec / (spec + diff), 0.1f, 0.9f);
    }
    // The terms that depend on the outgoing direction alone, computed once
    // per bounce and shared by every surface_eval_pdf / surface_sample of
    // that bounce (each must be given the wo bound here): the same values
    // those functions would otherwise compute per call.
    V3 wo_spec_albedo, wo_compensation;
    float wo_p_spec, alpha, wo_lambda;
    __device__ inline void bind(V3 wo) {
        alpha = roughness_to_alpha(roughness);
        wo_lambda = smith_lambda(alpha, wo.z);
        wo_spec_albedo = specular_albedo(wo.z);
        wo_p_spec = specular_probability(wo.z, wo_spec_albedo);
        wo_compensation = transmission > 0.0f ? v3s(1.0f) : compensation(wo.z, normal_reflectance());
    }
    __device__ inline V3 reflection_fresnel(float c) const {
        if (transmission > 0.0f && !conductor) return v3s(fresnel_dielectric(c, relative_ior));
        return fresnel(c);
    }
};

__device__ inline Surface surface_from_material(const GpuMaterial& m, bool front_face, const float* table) {
    V3 base = ld3(m.base_color);
    float metallic = clampf(m.metallic, 0.0f, 1.0f);
    Surface s;
    s.diffuse_albedo = base * (1.0f - metallic);
    V3 df0 = v3s(f0_from_ior(fmaxf(m.ior, 1.0f)));
    s.transmission = clampf(m.transmission, 0.0f, 1.0f) * (1.0f - metallic);
    float ior = fmaxf(m.ior, 1.0f);
    s.roughness = clampf(m.roughness, 0.0f, 1.0f);
    s.f0 = df0 * (1.0f - metallic) + base * metallic;
    s.eta = ld3(m.eta);
    s.k = ld3(m.k);
    s.conductor = m.k[0] > 0.0f || m.k[1] > 0.0f || m.k[2] > 0.0f;
    s.tint = is_zero(base) ? v3s(1.0f) : base;
    s.relative_ior = front_face ? ior : 1.0f / ior;
    s.table = table;
    return s;
}

__device__ inline bool dielectric_half(V3 wo, V3 wi, float eta, V3* h) {
    V3 v = normalize_or(wo + eta * wi, v3s(0.0f));
    if (is_zero(v)) return false;
    *h = v.z < 0.0f ? -v : v;
    return true;
}
__device__ inline float dielectric_eval(float alpha, float eta, V3 wo, V3 wi) {
    if (wo.z * wi.z >= 0.0f) return 0.0f;
    V3 h;
    if (!dielectric_half(wo, wi, eta, &h)) return 0.0f;
    float wo_h = dot(wo, h), wi_h = dot(wi, h);
    if (wo_h <= 0.0f || wi_h >= 0.0f) return 0.0f;
    float denom = wo_h + eta * wi_h;
    if (fabsf(denom) < 1.0e-9f) return 0.0f;
    float f = fresnel_dielectric(wo_h, eta);
    float d = ggx_d(alpha, fabsf(h.z));
    float g2 = smith_g2(alpha, fabsf(wo.z), fabsf(wi.z));
    return ((1.0f - f) * d * g2 * fabsf(wi_h * wo_h)) / (fabsf(wo.z) * fabsf(wi.z) * denom * denom);
}
__device__ inline float dielectric_pdf(float alpha, float eta, V3 wo, V3 wi) {
    if (wo.z * wi.z >= 0.0f) return 0.0f;
    V3 h;
    if (!dielectric_half(wo, wi, eta, &h)) return 0.0f;
    float wo_h = dot(wo, h), wi_h = dot(wi, h);
    if (wo_h <= 0.0f || wi_h >= 0.0f) return 0.0f;
    float denom = wo_h + eta * wi_h;
    if (fabsf(denom) < 1.0e-9f) return 0.0f;
    float jac = fabsf(eta * eta * wi_h) / (denom * denom);
    return ggx_vndf_half_pdf(alpha, wo, h) * jac;
}

// The BSDF value and its sampling pdf for one (wo, wi) (Bsdf::eval and
// Bsdf::pdf, bsdf.rs), computed together: the half vector and the GGX
// distribution once for both, the wo terms from Surface::bind (which must
// have been given this wo; so must surface_sample).
__device__ inline void surface_eval_pdf(const Surface& s, V3 wo, V3 wi, V3* f_out, float* pdf_out) {
    *f_out = v3s(0.0f);
    *pdf_out = 0.0f;
    if (wo.z <= 0.0f) return;
    const float alpha = s.alpha;
    const float p_spec = s.wo_p_spec;
    const float rest = 1.0f - p_spec;
    if (wi.z < 0.0f) {
        if (s.transmission <= 0.0f) return;
        const float t = dielectric_eval(alpha, s.relative_ior, wo, wi);
        *f_out = v3s(t) * s.transmission * s.tint;
        *pdf_out = rest * s.transmission * dielectric_pdf(alpha, s.relative_ior, wo, wi);
        return;
    }
    if (wi.z <= 0.0f) return;
    V3 transmitted = v3s(1.0f) - s.wo_spec_albedo;
    V3 f = s.diffuse_albedo * INV_PI * transmitted * (1.0f - s.transmission);
    V3 h = normalize_or(wo + wi, v3s(0.0f));
    float spec_pdf = 0.0f;
    if (!is_zero(h)) {
        const float d = ggx_d(alpha, h.z);
        const float g2 = 1.0f / (1.0f + s.wo_lambda + smith_lambda(alpha, wi.z));  // smith_g2(alpha, wo.z, wi.z)
        V3 fr = s.reflection_fresnel(fmaxf(dot(wo, h), 0.0f));
        V3 single = fr * (d * g2 / (4.0f * wo.z * wi.z));
        f = f + (s.transmission > 0.0f ? single : single * s.wo_compensation);
        spec_pdf = (1.0f / (1.0f + s.wo_lambda)) * d / (4.0f * wo.z);  // ggx_vndf_pdf(alpha, wo, h)
    }
    *f_out = f;
    *pdf_out = p_spec * spec_pdf + rest * (1.0f - s.transmission) * (wi.z * INV_PI);
}

// Returns pdf (0 when invalid); wi and weight through pointers.
__device__ inline float surface_sample(const Surface& s, V3 wo, float u_lobe, float ux, float uy, V3* wi_out,
                                       V3* weight) {
    if (wo.z <= 0.0f) return 0.0f;
    const float p_spec = s.wo_p_spec;
    const float alpha = s.alpha;
    V3 wi;
    if (u_lobe < p_spec) {
        V3 m = sample_ggx_vndf(wo, alpha, ux, uy);
        V3 r = reflect(wo, m);
        if (r.z <= 0.0f) return 0.0f;
        wi = r;
    } else if (u_lobe < p_spec + (1.0f - p_spec) * s.transmission) {
        V3 m = sample_ggx_vndf(wo, alpha, ux, uy);
        V3 wt;
        if (!refract(wo, m, s.relative_ior, &wt)) return 0.0f;
        if (wt.z >= 0.0f) return 0.0f;
        wi = wt;
    } else {
        wi = cosine_hemisphere(ux, uy);
    }
    if (wi.z == 0.0f || is_zero(wi)) return 0.0f;
    if (wi.z < 0.0f && s.transmission <= 0.0f) return 0.0f;
    V3 f;
    float pdf;
    surface_eval_pdf(s, wo, wi, &f, &pdf);
    if (!(pdf > 0.0f)) return 0.0f;
    *wi_out = wi;
    *weight = f * (fabsf(wi.z) / pdf);
    return pdf;
}

__device__ inline float power_heuristic(float a, float b) {
    if (a <= 0.0f) return 0.0f;
    float r = b / a;
    return 1.0f / (1.0f + r * r);
}

// --- analytic primitives (scene.rs:76-140) -------------------------------------------

__device__ inline bool intersect_sphere(const GpuPrimitive& s, V3 o, V3 d, float t_min, float t_max, float* out) {
    V3 oc = o - ld3(s.position);
    float a = dot(d, d);
    float h = dot(oc, d);
    float c = dot(oc, oc) - s.radius * s.radius;
    float disc = h * h - a * c;
    if (disc < 0.0f) return false;
    float sd = sqrtf(disc);
    float t = (-h - sd) / a;
    if (t < t_min || t > t_max) {
        t = (-h + sd) / a;
        if (t < t_min || t > t_max) return false;
    }
    *out = t;
    return true;
}
// The quad's w = n / |n|^2 (n = edge_u x edge_v), a constant of the quad: a
// scene computes it once (quad_w) with the same operations, so the test
// below sees the same value it would compute itself.
__device__ inline V3 quad_w(const GpuPrimitive& q) {
    const V3 n_raw = cross(ld3(q.edge_u), ld3(q.edge_v));
    return n_raw / dot(n_raw, n_raw);
}

// `w` is quad_w(q). With t_min > 0, a plane behind the origin (t's numerator
// and denominator of opposite signs, or a zero numerator) is rejected before
// the division: t would be negative or zero and fail t >= t_min anyway.
__device__ inline bool intersect_quad(const GpuPrimitive& q, V3 w, V3 o, V3 d, float t_min, float t_max,
                                      float* out) {
    V3 origin = ld3(q.position), eu = ld3(q.edge_u), ev = ld3(q.edge_v), normal = ld3(q.normal);
    float denom = dot(normal, d);
    if (fabsf(denom) < 1.0e-8f) return false;
    const float num = dot(normal, origin) - dot(normal, o);
    if (t_min > 0.0f && (num == 0.0f || signbit(num) != signbit(denom))) return false;
    // num * rcp(denom) is within a few ulp of num / denom: a plane far
    // beyond t_max needs no correctly rounded division
    if (num * __builtin_amdgcn_rcpf(denom) > t_max * 1.0001f) return false;
    float t = num / denom;
    if (t < t_min || t > t_max) return false;
    V3 p = o + t * d;
    V3 dd = p - origin;
    float a = dot(w, cross(dd, ev));
    float b = dot(w, cross(eu, dd));
    if (!(a >= 0.0f && a <= 1.0f) || !(b >= 0.0f && b <= 1.0f)) return false;
    *out = t;
    return true;
}

// --- environment map (envmap.rs) --------------------------------------------------------

struct EnvView {
    uint32_t w, h;
    float total_weight;
    const float4* __restrict__ pixels;
    const float* __restrict__ conditional;
    const float* __restrict__ marginal;
    __device__ inline bool present() const { return w > 0 && total_weight > 0.0f; }
};
__device__ inline V3 env_direction_from_uv(float u, float v) {
    float theta = v * PI, phi = u * TAU;
    float st = sinf(theta), ct = cosf(theta);
    return V3{st * cosf(phi), ct, st * sinf(phi)};
}
__device__ inline void env_texel(const EnvView& e, V3 d, uint32_t* i, uint32_t* j) {
    float theta = acosf(clampf(d.y, -1.0f, 1.0f));
    float phi = atan2f(d.z, d.x);
    float r = fmodf(phi, TAU);
    if (r < 0.0f) r = r + TAU;
    float u = r / TAU, v = theta / PI;
    float fi = u * (float)e.w, fj = v * (float)e.h;
    int64_t qi = fi != fi ? 0 : (int64_t)fi, qj = fj != fj ? 0 : (int64_t)fj;
    *i = (uint32_t)(qi < 0 ? 0 : (qi > (int64_t)e.w - 1 ? (int64_t)e.w - 1 : qi));
    *j = (uint32_t)(qj < 0 ? 0 : (qj > (int64_t)e.h - 1 ? (int64_t)e.h - 1 : qj));
}
__device__ inline V3 env_radiance(const EnvView& e, V3 d) {
    if (e.w == 0) return v3s(0.0f);
    uint32_t i, j;
    env_texel(e, d, &i, &j);
    return f4xyz(e.pixels[(size_t)j * e.w + i]);
}
__device__ inline float env_pdf(const EnvView& e, V3 d) {
    if (!e.present()) return 0.0f;
    uint32_t i, j;
    env_texel(e, d, &i, &j);
    float theta = ((float)j + 0.5f) / (float)e.h * PI;
    float st = sinf(theta);
    if (st <= 0.0f) return 0.0f;
    float weight = fmaxf(luminance(f4xyz(e.pixels[(size_t)j * e.w + i])), 0.0f) * st;
    float pdf_uv = weight / e.total_weight * (float)(e.w * e.h);
    return pdf_uv / (2.0f * PI * PI * st);
}
__device__ inline void env_sample_cdf(const float* __restrict__ cdf, size_t base, uint32_t n, float x, uint32_t* cell,
                                      float* frac) {
    uint32_t lo = 0, len = n;
    while (len > 0) {
        uint32_t half = len / 2, mid = lo + half;
        if (cdf[base + mid] <= x) {
            lo = mid + 1;
            len -= half + 1;
        } else {
            len = half;
        }
    }
    uint32_t c = min(lo > 0 ? lo - 1 : 0u, n - 1);
    float c0 = cdf[base + c], c1 = cdf[base + c + 1];
    float span = c1 - c0;
    *frac = span > 0.0f ? clampf((x - c0) / span, 0.0f, 1.0f) : 0.0f;
    *cell = c;
}
// Returns pdf (0: invalid).
__device__ inline float env_sample(const EnvView& e, float ux, float uy, V3* dir, V3* rad) {
    if (!e.present()) return 0.0f;
    uint32_t j, i;
    float dv, du;
    env_sample_cdf(e.marginal, 0, e.h, uy, &j, &dv);
    env_sample_cdf(e.conditional, (size_t)j * (e.w + 1), e.w, ux, &i, &du);
    float u = ((float)i + du) / (float)e.w, v = ((float)j + dv) / (float)e.h;
    V3 d = env_direction_from_uv(u, v);
    float theta = ((float)j + 0.5f) / (float)e.h * PI;
    float st = sinf(theta);
    if (st <= 0.0f) return 0.0f;
    V3 r = f4xyz(e.pixels[(size_t)j * e.w + i]);
    float weight = fmaxf(luminance(r), 0.0f) * st;
    float pdf_uv = weight / e.total_weight * (float)(e.w * e.h);
    float pdf = pdf_uv / (2.0f * PI * PI * st);
    if (pdf <= 0.0f) return 0.0f;
    *dir = d;
    *rad = r;
    return pdf;
}

}  // namespace dev
}  // namespace lucida
