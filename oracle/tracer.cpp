// Oracle: the CPU reference path tracer, restated from the reference's Rust
// twin, in operation order, with contraction off.
//
//   rng.rs:28-85            pcg_hash, Rng
//   sobol.rs:88-361         direction numbers, sobol, Owen scramble, Sampler
//   math.rs:22-176          onb, to_world/to_local, cosine hemisphere,
//                           concentric disk, offset_ray_origin, luminance
//   camera.rs:84-135        generate_ray
//   scene.rs:60-607         sphere/quad intersection, Scene::intersect,
//                           shade_triangle, shade_analytic, occluded
//   bsdf.rs:23-700,1178-1662  GGX, Smith, VNDF, Fresnel, dielectric BTDF,
//                           energy compensation, Surface eval/pdf/sample
//   light.rs:150-335        sample_lights, light_pdf_from_geometry,
//                           power_heuristic, unoccluded
//   envmap.rs:54-320        radiance, pdf, sample, sample_cdf
//   integrator.rs:209-699   radiance_with_guide, direct_light, render_pixel,
//                           render_with_progress (32-row bands)
//   denoise.rs:101-228      local_deviation, denoise
//   image.rs:158-225        compare
//
// TEST INFRASTRUCTURE: the checker for the HIP path and the CPU baseline;
// never linked into the product.
#include <atomic>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "bvh.hpp"
#include "oracle.h"
#include "tracer.h"

namespace oracle {

// --- rng.rs -------------------------------------------------------------------

static inline uint32_t pcg_hash(uint32_t input) {
    uint32_t state = input * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28) + 4u)) ^ state) * 277803737u;
    return (word >> 22) ^ word;
}

struct Rng {
    uint32_t state;
    Rng(uint32_t frame, uint32_t pixel, uint32_t sample) {
        uint32_t h = pcg_hash(frame);
        h = pcg_hash(h + pixel);
        h = pcg_hash(h + sample);
        state = h;
    }
    uint32_t next_u32() {
        state = state * 747796405u + 2891336453u;
        uint32_t word = ((state >> ((state >> 28) + 4u)) ^ state) * 277803737u;
        return (word >> 22) ^ word;
    }
    float next_f32() { return (float)(next_u32() >> 8) * (1.0f / 16777216.0f); }
};

// --- sobol.rs -------------------------------------------------------------------

static uint32_t SOBOL_V[2][32];
static std::once_flag sobol_once;

// sobol.rs:88-140: dimension 0 is van der Corput; dimension 1 from the
// degree-1 polynomial with m = {1}, by the recurrence on direction numbers.
static void init_sobol() {
    for (int k = 0; k < 32; ++k) SOBOL_V[0][k] = 1u << (31 - k);
    const int s = 1;
    uint32_t vv[33] = {0};
    vv[1] = 1u << 31;  // init[0] << (32 - 1)
    for (int i = s + 1; i <= 32; ++i) vv[i] = vv[i - s] ^ (vv[i - s] >> s);
    for (int k = 0; k < 32; ++k) SOBOL_V[1][k] = vv[k + 1];
}

uint32_t sobol(uint32_t index, uint32_t dim) {
    std::call_once(sobol_once, init_sobol);
    uint32_t x = 0;
    for (int k = 0; k < 32; ++k)
        if ((index >> k) & 1u) x ^= SOBOL_V[dim][k];
    return x;
}

static inline uint32_t reverse_bits(uint32_t x) {
    x = ((x >> 1) & 0x55555555u) | ((x & 0x55555555u) << 1);
    x = ((x >> 2) & 0x33333333u) | ((x & 0x33333333u) << 2);
    x = ((x >> 4) & 0x0F0F0F0Fu) | ((x & 0x0F0F0F0Fu) << 4);
    x = ((x >> 8) & 0x00FF00FFu) | ((x & 0x00FF00FFu) << 8);
    return (x >> 16) | (x << 16);
}

static inline uint32_t laine_karras(uint32_t x, uint32_t seed) {
    x += seed;
    x ^= x * 0x6c50b47cu;
    x ^= x * 0xb82f1e52u;
    x ^= x * 0xc7afe638u;
    x ^= x * 0x8d22f6e6u;
    return x;
}

uint32_t owen_scramble(uint32_t x, uint32_t seed) { return reverse_bits(laine_karras(reverse_bits(x), seed)); }

uint32_t hash_u32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

float sobol_sample(uint32_t index, uint32_t dim, uint32_t seed) {
    uint32_t group = dim / 2u;
    uint32_t base = dim % 2u;
    uint32_t group_seed = hash_u32(seed ^ (group * 0x9e3779b9u));
    uint32_t shuffled = owen_scramble(index, group_seed);
    uint32_t v = sobol(shuffled, base);
    uint32_t scrambled = owen_scramble(v, hash_u32(group_seed + base + 1u));
    return (float)(scrambled >> 8) * (1.0f / 16777216.0f);
}

struct Sampler {
    bool sobol_kind;
    Rng rng;
    uint32_t index, seed, dim;
    Sampler(bool sob, uint32_t frame, uint32_t pixel, uint32_t sample)
        : sobol_kind(sob), rng(frame, pixel, sample), index(sample), seed(hash_u32(pixel + frame * 0x9e3779b9u)),
          dim(0) {}
    float next_f32() {
        if (!sobol_kind) return rng.next_f32();
        float v = sobol_sample(index, dim, seed);
        dim += 1;
        return v;
    }
    Vec2 next_vec2() {
        if (!sobol_kind) {
            float x = rng.next_f32();
            float y = rng.next_f32();
            return Vec2{x, y};
        }
        dim = (dim + 1u) & ~1u;
        float x = sobol_sample(index, dim, seed);
        float y = sobol_sample(index, dim + 1, seed);
        dim += 2;
        return Vec2{x, y};
    }
};

// --- math.rs ------------------------------------------------------------------------

static inline void onb(Vec3 n, Vec3* t, Vec3* b) {
    float sign = std::copysign(1.0f, n.z);
    float a = -1.0f / (sign + n.z);
    float bb = n.x * n.y * a;
    *t = Vec3(1.0f + sign * n.x * n.x * a, sign * bb, -sign * n.x);
    *b = Vec3(bb, sign + n.y * n.y * a, -n.y);
}
static inline Vec3 to_world(Vec3 l, Vec3 n) {
    Vec3 t, b;
    onb(n, &t, &b);
    return t * l.x + b * l.y + n * l.z;
}
static inline Vec3 to_local(Vec3 w, Vec3 n) {
    Vec3 t, b;
    onb(n, &t, &b);
    return Vec3(dot(w, t), dot(w, b), dot(w, n));
}
static inline Vec2 concentric_sample_disk(Vec2 u) {
    float ox = 2.0f * u.x - 1.0f, oy = 2.0f * u.y - 1.0f;
    if (ox == 0.0f && oy == 0.0f) return Vec2{0, 0};
    float r, theta;
    if (std::fabs(ox) > std::fabs(oy)) {
        r = ox;
        theta = (PI / 4.0f) * (oy / ox);
    } else {
        r = oy;
        theta = PI / 2.0f - (PI / 4.0f) * (ox / oy);
    }
    return Vec2{r * std::cos(theta), r * std::sin(theta)};
}
static inline Vec3 sample_cosine_hemisphere(Vec2 u) {
    Vec2 d = concentric_sample_disk(u);
    float z = std::sqrt(std::fmax(1.0f - d.x * d.x - d.y * d.y, 0.0f));
    return Vec3(d.x, d.y, z);
}
static inline int32_t f2i_sat(float v) {
    if (v != v) return 0;
    if (v >= 2147483648.0f) return 2147483647;
    if (v <= -2147483648.0f) return (-2147483647 - 1);
    return (int32_t)v;
}
static inline float bits_add(float p, int32_t of) {
    uint32_t b;
    std::memcpy(&b, &p, 4);
    b = b + (uint32_t)(p < 0.0f ? -of : of);
    float r;
    std::memcpy(&r, &b, 4);
    return r;
}
Vec3 offset_ray_origin(Vec3 p, Vec3 n) {
    const float ORIGIN = 1.0f / 32.0f, FLOAT_SCALE = 1.0f / 65536.0f, INT_SCALE = 256.0f;
    int32_t ox = f2i_sat(INT_SCALE * n.x), oy = f2i_sat(INT_SCALE * n.y), oz = f2i_sat(INT_SCALE * n.z);
    float px = bits_add(p.x, ox), py = bits_add(p.y, oy), pz = bits_add(p.z, oz);
    return Vec3(std::fabs(p.x) < ORIGIN ? p.x + FLOAT_SCALE * n.x : px,
                std::fabs(p.y) < ORIGIN ? p.y + FLOAT_SCALE * n.y : py,
                std::fabs(p.z) < ORIGIN ? p.z + FLOAT_SCALE * n.z : pz);
}
static inline float luminance(Vec3 c) { return dot(c, Vec3(0.2126f, 0.7152f, 0.0722f)); }

// --- bsdf.rs ----------------------------------------------------------------------------

static const float MIN_ALPHA = 2.0e-3f;
static inline float roughness_to_alpha(float r) { return std::fmax(r * r, MIN_ALPHA); }

float ggx_d(float alpha, float cos_m) {
    if (cos_m <= 0.0f) return 0.0f;
    float a2 = alpha * alpha, c2 = cos_m * cos_m;
    float sin2 = std::fmax(1.0f - c2, 0.0f);
    float denom = a2 * c2 + sin2;
    return a2 / (PI * denom * denom);
}
float smith_lambda(float alpha, float cos_theta) {
    float c = std::fabs(cos_theta);
    if (c >= 1.0f) return 0.0f;
    float c2 = c * c;
    float tan2 = (1.0f - c2) / std::fmax(c2, 1.0e-12f);
    return 0.5f * (-1.0f + std::sqrt(1.0f + alpha * alpha * tan2));
}
static inline float smith_g1(float alpha, float c) { return 1.0f / (1.0f + smith_lambda(alpha, c)); }
static inline float smith_g2(float alpha, float co, float ci) {
    return 1.0f / (1.0f + smith_lambda(alpha, co) + smith_lambda(alpha, ci));
}
Vec3 sample_ggx_vndf(Vec3 wo, float alpha, Vec2 u) {
    Vec3 v = normalize(Vec3(alpha * wo.x, alpha * wo.y, wo.z));
    float lensq = v.x * v.x + v.y * v.y;
    Vec3 t1 = lensq > 0.0f ? Vec3(-v.y, v.x, 0.0f) / std::sqrt(lensq) : Vec3(1, 0, 0);
    Vec3 t2 = cross(v, t1);
    float r = std::sqrt(u.x);
    float phi = 2.0f * PI * u.y;
    float p1 = r * std::cos(phi);
    float p2 = r * std::sin(phi);
    float s = 0.5f * (1.0f + v.z);
    p2 = (1.0f - s) * std::sqrt(std::fmax(1.0f - p1 * p1, 0.0f)) + s * p2;
    Vec3 nh = p1 * t1 + p2 * t2 + std::sqrt(std::fmax(1.0f - p1 * p1 - p2 * p2, 0.0f)) * v;
    return normalize(Vec3(alpha * nh.x, alpha * nh.y, std::fmax(nh.z, 0.0f)));
}
static inline float ggx_vndf_pdf(float alpha, Vec3 wo, Vec3 m) {
    float cos_o = wo.z;
    if (cos_o <= 0.0f) return 0.0f;
    return smith_g1(alpha, cos_o) * ggx_d(alpha, m.z) / (4.0f * cos_o);
}
static inline float ggx_vndf_half_pdf(float alpha, Vec3 wo, Vec3 m) {
    float cos_o = wo.z, cos_om = dot(wo, m);
    if (cos_o <= 0.0f || cos_om <= 0.0f) return 0.0f;
    return smith_g1(alpha, cos_o) * ggx_d(alpha, std::fabs(m.z)) * cos_om / cos_o;
}
static inline Vec3 fresnel_schlick(Vec3 f0, float cos_theta) {
    float m = clampf(1.0f - cos_theta, 0.0f, 1.0f);
    float m2 = m * m;
    return f0 + (Vec3::splat(1.0f) - f0) * (m2 * m2 * m);
}
float fresnel_dielectric(float cos_i, float eta) {
    float c = std::fabs(clampf(cos_i, -1.0f, 1.0f));
    float sin2_t = (1.0f - c * c) / (eta * eta);
    if (sin2_t >= 1.0f) return 1.0f;
    float cos_t = std::sqrt(std::fmax(1.0f - sin2_t, 0.0f));
    float rpar = (eta * c - cos_t) / (eta * c + cos_t);
    float rper = (c - eta * cos_t) / (c + eta * cos_t);
    return 0.5f * (rpar * rpar + rper * rper);
}
static inline float f0_from_ior(float ior) {
    float r = (ior - 1.0f) / (ior + 1.0f);
    return r * r;
}
Vec3 fresnel_conductor(float cos_i, Vec3 eta, Vec3 k) {
    float c = clampf(cos_i, 0.0f, 1.0f);
    float c2 = c * c;
    float sin2 = 1.0f - c2;
    Vec3 eta2 = eta * eta, k2 = k * k;
    Vec3 t0 = eta2 - k2 - Vec3::splat(sin2);
    Vec3 a2pb2 = vpow(vmax(t0 * t0 + 4.0f * eta2 * k2, Vec3()), 0.5f);
    Vec3 t1 = a2pb2 + Vec3::splat(c2);
    Vec3 a = vpow(vmax(0.5f * (a2pb2 + t0), Vec3()), 0.5f);
    Vec3 t2 = 2.0f * c * a;
    Vec3 rs = (t1 - t2) / (t1 + t2);
    Vec3 t3 = c2 * a2pb2 + Vec3::splat(sin2 * sin2);
    Vec3 t4 = t2 * sin2;
    Vec3 rp = rs * (t3 - t4) / (t3 + t4);
    return 0.5f * (rs + rp);
}
static inline Vec3 reflect(Vec3 v, Vec3 n) { return -v + 2.0f * dot(v, n) * n; }
static inline bool refract(Vec3 wo, Vec3 m, float eta, Vec3* out) {
    float cos_i = dot(wo, m);
    float sin2_i = std::fmax(1.0f - cos_i * cos_i, 0.0f);
    float sin2_t = sin2_i / (eta * eta);
    if (sin2_t >= 1.0f) return false;
    float cos_t = std::sqrt(1.0f - sin2_t);
    Vec3 wt = -wo / eta + (cos_i / eta - std::copysign(cos_t, cos_i)) * m;
    *out = normalize_or_zero(wt);
    return true;
}

struct BsdfSample {
    Vec3 wi, weight;
    float pdf;
};
static const BsdfSample INVALID{Vec3(0, 0, 1), Vec3(), 0.0f};

// specular::sample_single (bsdf.rs:~ 485) used by the energy table.
static BsdfSample specular_sample_single(float alpha, Vec3 f0, Vec3 wo, Vec2 u) {
    if (wo.z <= 0.0f) return INVALID;
    Vec3 m = sample_ggx_vndf(wo, alpha, u);
    Vec3 wi = reflect(wo, m);
    if (wi.z <= 0.0f) return INVALID;
    float pdf = ggx_vndf_pdf(alpha, wo, m);
    if (pdf <= 0.0f) return INVALID;
    Vec3 f = fresnel_schlick(f0, std::fmax(dot(wo, m), 0.0f));
    float g1 = smith_g1(alpha, wo.z);
    float g2 = smith_g2(alpha, wo.z, wi.z);
    return BsdfSample{wi, f * (g2 / g1), pdf};
}

// bsdf.rs:1209-1227 specular_directional_albedo_grid + energy::compute_table.
static float directional_albedo_grid(float alpha, float cos_o, uint32_t grid) {
    float sin_t = std::sqrt(std::fmax(1.0f - cos_o * cos_o, 0.0f));
    Vec3 wo(sin_t, 0.0f, cos_o);
    double sum = 0.0;
    float n = (float)grid;
    for (uint32_t i = 0; i < grid; ++i)
        for (uint32_t j = 0; j < grid; ++j) {
            Vec2 u{((float)i + 0.5f) / n, ((float)j + 0.5f) / n};
            BsdfSample s = specular_sample_single(alpha, Vec3::splat(1.0f), wo, u);
            if (s.pdf > 0.0f) sum += (double)s.weight.x;
        }
    return (float)(sum / ((double)grid * (double)grid));
}

static float GGX_E[32 * 32];
static std::once_flag ggx_once;
static void compute_table() {
    std::vector<std::thread> th;
    for (int ri = 0; ri < 32; ++ri)
        th.emplace_back([ri] {
            float roughness = (float)ri / 31.0f;
            float alpha = roughness_to_alpha(roughness);
            for (int ci = 0; ci < 32; ++ci) {
                float root_mu = (float)ci / 31.0f;
                float mu = std::fmax(root_mu * root_mu, 1.0e-4f);
                GGX_E[ri * 32 + ci] = directional_albedo_grid(alpha, mu, 128);
            }
        });
    for (auto& t : th) t.join();
}
const float* ggx_table() {
    std::call_once(ggx_once, compute_table);
    return GGX_E;
}
static inline float energy_lookup(float roughness, float cos_o) {
    const float* table = ggx_table();
    const int n = 32;
    float fx = clampf(roughness, 0.0f, 1.0f) * (float)(n - 1);
    float fy = std::sqrt(clampf(cos_o, 0.0f, 1.0f)) * (float)(n - 1);
    int x0 = std::min((int)fx, n - 1), y0 = std::min((int)fy, n - 1);
    int x1 = std::min(x0 + 1, n - 1), y1 = std::min(y0 + 1, n - 1);
    float tx = fx - (float)x0, ty = fy - (float)y0;
    float a = table[x0 * n + y0], b = table[x0 * n + y1], c = table[x1 * n + y0], d = table[x1 * n + y1];
    float top = a + (b - a) * ty;
    float bot = c + (d - c) * ty;
    return clampf(top + (bot - top) * tx, 1.0e-3f, 1.0f);
}
static inline Vec3 fresnel_average(Vec3 f0) { return f0 + (Vec3::splat(1.0f) - f0) * (1.0f / 21.0f); }
static inline Vec3 compensation(float roughness, float cos_o, Vec3 f0) {
    float e = energy_lookup(roughness, cos_o);
    return Vec3::splat(1.0f) + fresnel_average(f0) * ((1.0f - e) / e);
}

static inline bool dielectric_half(Vec3 wo, Vec3 wi, float eta, Vec3* h) {
    Vec3 v = normalize_or_zero(wo + eta * wi);
    if (v == Vec3()) return false;
    *h = v.z < 0.0f ? -v : v;
    return true;
}
static float dielectric_eval(float alpha, float eta, Vec3 wo, Vec3 wi) {
    if (wo.z * wi.z >= 0.0f) return 0.0f;
    Vec3 h;
    if (!dielectric_half(wo, wi, eta, &h)) return 0.0f;
    float wo_h = dot(wo, h), wi_h = dot(wi, h);
    if (wo_h <= 0.0f || wi_h >= 0.0f) return 0.0f;
    float denom = wo_h + eta * wi_h;
    if (std::fabs(denom) < 1.0e-9f) return 0.0f;
    float f = fresnel_dielectric(wo_h, eta);
    float d = ggx_d(alpha, std::fabs(h.z));
    float g2 = smith_g2(alpha, std::fabs(wo.z), std::fabs(wi.z));
    return ((1.0f - f) * d * g2 * std::fabs(wi_h * wo_h)) / (std::fabs(wo.z) * std::fabs(wi.z) * denom * denom);
}
static float dielectric_pdf(float alpha, float eta, Vec3 wo, Vec3 wi) {
    if (wo.z * wi.z >= 0.0f) return 0.0f;
    Vec3 h;
    if (!dielectric_half(wo, wi, eta, &h)) return 0.0f;
    float wo_h = dot(wo, h), wi_h = dot(wi, h);
    if (wo_h <= 0.0f || wi_h >= 0.0f) return 0.0f;
    float denom = wo_h + eta * wi_h;
    if (std::fabs(denom) < 1.0e-9f) return 0.0f;
    float jac = std::fabs(eta * eta * wi_h) / (denom * denom);
    return ggx_vndf_half_pdf(alpha, wo, h) * jac;
}
static inline float specular_pdf(float alpha, Vec3 wo, Vec3 wi) {
    if (wo.z <= 0.0f || wi.z <= 0.0f) return 0.0f;
    Vec3 h = normalize_or_zero(wo + wi);
    if (h == Vec3()) return 0.0f;
    return ggx_vndf_pdf(alpha, wo, h);
}

struct Surface {
    Vec3 diffuse_albedo;
    float roughness;
    Vec3 f0, eta, k;
    bool conductor;
    float transmission;
    Vec3 tint;
    float relative_ior;

    static Surface from_material(const GpuMaterial& m, bool front_face) {
        Vec3 base = Vec3::from(m.base_color);
        float metallic = clampf(m.metallic, 0.0f, 1.0f);
        Surface s;
        s.diffuse_albedo = base * (1.0f - metallic);
        Vec3 dielectric_f0 = Vec3::splat(f0_from_ior(std::fmax(m.ior, 1.0f)));
        s.transmission = clampf(m.transmission, 0.0f, 1.0f) * (1.0f - metallic);
        float ior = std::fmax(m.ior, 1.0f);
        s.roughness = clampf(m.roughness, 0.0f, 1.0f);
        s.f0 = lerp(dielectric_f0, base, metallic);
        s.eta = Vec3::from(m.eta);
        s.k = Vec3::from(m.k);
        s.conductor = m.k[0] > 0.0f || m.k[1] > 0.0f || m.k[2] > 0.0f;
        s.tint = base == Vec3() ? Vec3::splat(1.0f) : base;
        s.relative_ior = front_face ? ior : 1.0f / ior;
        return s;
    }
    Vec3 fresnel(float c) const { return conductor ? fresnel_conductor(c, eta, k) : fresnel_schlick(f0, c); }
    Vec3 normal_reflectance() const { return conductor ? fresnel_conductor(1.0f, eta, k) : f0; }
    Vec3 specular_albedo(float cos_o) const {
        float e_ss = energy_lookup(roughness, std::fabs(cos_o));
        Vec3 f = fresnel(std::fabs(cos_o));
        Vec3 f_avg = fresnel_average(normal_reflectance());
        return clamp(f * e_ss + f_avg * (1.0f - e_ss), Vec3(), Vec3::splat(1.0f));
    }
    float specular_probability(float cos_o) const {
        if (transmission > 0.0f) {
            float f = fresnel_dielectric(std::fabs(cos_o), relative_ior);
            if (f >= 1.0f) return 1.0f;
            return clampf(f, 0.05f, 0.95f);
        }
        float spec = luminance(specular_albedo(cos_o));
        float diff = luminance(diffuse_albedo * (Vec3::splat(1.0f) - specular_albedo(cos_o)));
        if (spec + diff <= 0.0f) return 1.0f;
        return clampf(spec / (spec + diff), 0.1f, 0.9f);
    }
    Vec3 reflection_fresnel(float c) const {
        if (transmission > 0.0f && !conductor) return Vec3::splat(fresnel_dielectric(c, relative_ior));
        return fresnel(c);
    }
};

static Vec3 surface_eval(const Surface& s, Vec3 wo, Vec3 wi) {
    if (wo.z <= 0.0f) return Vec3();
    float alpha = roughness_to_alpha(s.roughness);
    if (wi.z < 0.0f) {
        if (s.transmission <= 0.0f) return Vec3();
        float t = dielectric_eval(alpha, s.relative_ior, wo, wi);
        return Vec3::splat(t) * s.transmission * s.tint;
    }
    if (wi.z <= 0.0f) return Vec3();
    Vec3 transmitted = Vec3::splat(1.0f) - s.specular_albedo(wo.z);
    Vec3 diffuse = wi.z <= 0.0f ? Vec3() : s.diffuse_albedo * INV_PI;
    Vec3 f = diffuse * transmitted * (1.0f - s.transmission);
    Vec3 h = normalize_or_zero(wo + wi);
    if (!(h == Vec3())) {
        float d = ggx_d(alpha, h.z);
        float g2 = smith_g2(alpha, wo.z, wi.z);
        Vec3 fr = s.reflection_fresnel(std::fmax(dot(wo, h), 0.0f));
        Vec3 single = fr * (d * g2 / (4.0f * wo.z * wi.z));
        f += s.transmission > 0.0f ? single : single * compensation(s.roughness, wo.z, s.normal_reflectance());
    }
    return f;
}

static float surface_pdf(const Surface& s, Vec3 wo, Vec3 wi) {
    if (wo.z <= 0.0f) return 0.0f;
    float p_spec = s.specular_probability(wo.z);
    float alpha = roughness_to_alpha(s.roughness);
    float rest = 1.0f - p_spec;
    if (wi.z < 0.0f) {
        if (s.transmission <= 0.0f) return 0.0f;
        return rest * s.transmission * dielectric_pdf(alpha, s.relative_ior, wo, wi);
    }
    if (wi.z <= 0.0f) return 0.0f;
    float dpdf = wi.z <= 0.0f ? 0.0f : wi.z * INV_PI;
    return p_spec * specular_pdf(alpha, wo, wi) + rest * (1.0f - s.transmission) * dpdf;
}

static BsdfSample surface_sample(const Surface& s, Vec3 wo, float u_lobe, Vec2 u) {
    if (wo.z <= 0.0f) return INVALID;
    float p_spec = s.specular_probability(wo.z);
    float alpha = roughness_to_alpha(s.roughness);
    Vec3 wi;
    if (u_lobe < p_spec) {
        Vec3 m = sample_ggx_vndf(wo, alpha, u);
        Vec3 r = reflect(wo, m);
        if (r.z <= 0.0f) return INVALID;
        wi = r;
    } else if (u_lobe < p_spec + (1.0f - p_spec) * s.transmission) {
        Vec3 m = sample_ggx_vndf(wo, alpha, u);
        Vec3 wt;
        if (!refract(wo, m, s.relative_ior, &wt)) return INVALID;
        if (wt.z >= 0.0f) return INVALID;
        wi = wt;
    } else {
        wi = sample_cosine_hemisphere(u);
    }
    if (wi.z == 0.0f || wi == Vec3()) return INVALID;
    if (wi.z < 0.0f && s.transmission <= 0.0f) return INVALID;
    Vec3 f = surface_eval(s, wo, wi);
    float pdf = surface_pdf(s, wo, wi);
    if (pdf <= 0.0f) return INVALID;
    return BsdfSample{wi, f * (std::fabs(wi.z) / pdf), pdf};
}

// --- scene.rs -----------------------------------------------------------------------

static const float T_MAX = 1.0e30f, T_MIN = 1.0e-4f;

static bool intersect_sphere(const GpuPrimitive& s, const Ray& ray, float t_min, float t_max, float* out) {
    Vec3 center = Vec3::from(s.position);
    Vec3 oc = ray.origin - center;
    float a = dot(ray.dir, ray.dir);
    float h = dot(oc, ray.dir);
    float c = dot(oc, oc) - s.radius * s.radius;
    float disc = h * h - a * c;
    if (disc < 0.0f) return false;
    float sd = std::sqrt(disc);
    float t = (-h - sd) / a;
    if (t < t_min || t > t_max) {
        t = (-h + sd) / a;
        if (t < t_min || t > t_max) return false;
    }
    *out = t;
    return true;
}

static bool intersect_quad(const GpuPrimitive& q, const Ray& ray, float t_min, float t_max, float* out) {
    Vec3 origin = Vec3::from(q.position), eu = Vec3::from(q.edge_u), ev = Vec3::from(q.edge_v),
         normal = Vec3::from(q.normal);
    float denom = dot(normal, ray.dir);
    if (std::fabs(denom) < 1.0e-8f) return false;
    float t = (dot(normal, origin) - dot(normal, ray.origin)) / denom;
    if (t < t_min || t > t_max) return false;
    Vec3 p = ray.origin + t * ray.dir;
    Vec3 d = p - origin;
    Vec3 n_raw = cross(eu, ev);
    Vec3 w = n_raw / dot(n_raw, n_raw);
    float a = dot(w, cross(d, ev));
    float b = dot(w, cross(eu, d));
    if (!(a >= 0.0f && a <= 1.0f) || !(b >= 0.0f && b <= 1.0f)) return false;
    *out = t;
    return true;
}

struct Hit {
    float t;
    uint32_t steps;
    Vec3 position, normal, geometric_normal;
    bool front_face;
    uint32_t material;
    float light_area;
};

struct EnvSampleR {
    Vec3 direction, radiance;
    float pdf;
};

struct Env {
    uint32_t w = 0, h = 0;
    const float* pixels = nullptr;
    const float* conditional = nullptr;
    const float* marginal = nullptr;
    float total_weight = 0.0f;
    bool empty() const { return w == 0 || total_weight <= 0.0f; }
    static Vec3 direction_from_uv(float u, float v) {
        float theta = v * PI, phi = u * TAU;
        float st = std::sin(theta), ct = std::cos(theta);
        return Vec3(st * std::cos(phi), ct, st * std::sin(phi));
    }
    static Vec2 uv_from_direction(Vec3 d) {
        float theta = std::acos(clampf(d.y, -1.0f, 1.0f));
        float phi = std::atan2(d.z, d.x);
        float r = std::fmod(phi, TAU);
        if (r < 0.0f) r = r + std::fabs(TAU);
        return Vec2{r / TAU, theta / PI};
    }
    void texel_of(Vec2 uv, size_t* i, size_t* j) const {
        auto cl = [](float v, int64_t hi) -> int64_t {
            int64_t q = v != v ? 0 : (v >= 9.2e18f ? INT64_MAX : (v <= -9.2e18f ? INT64_MIN : (int64_t)v));
            return q < 0 ? 0 : (q > hi ? hi : q);
        };
        *i = (size_t)cl(uv.x * (float)w, (int64_t)w - 1);
        *j = (size_t)cl(uv.y * (float)h, (int64_t)h - 1);
    }
    Vec3 px(size_t i, size_t j) const {
        const float* p = &pixels[4 * (j * w + i)];
        return Vec3(p[0], p[1], p[2]);
    }
    Vec3 radiance(Vec3 dir) const {
        if (w == 0) return Vec3();
        size_t i, j;
        texel_of(uv_from_direction(dir), &i, &j);
        return px(i, j);
    }
    float pdf(Vec3 dir) const {
        if (empty()) return 0.0f;
        size_t i, j;
        texel_of(uv_from_direction(dir), &i, &j);
        float theta = ((float)j + 0.5f) / (float)h * PI;
        float st = std::sin(theta);
        if (st <= 0.0f) return 0.0f;
        float weight = std::fmax(luminance(px(i, j)), 0.0f) * st;
        float pdf_uv = weight / total_weight * (float)(w * h);
        return pdf_uv / (2.0f * PI * PI * st);
    }
    static void sample_cdf(const float* cdf, size_t base, size_t n, float x, size_t* cell_out, float* frac_out) {
        size_t lo = 0, len = n;
        while (len > 0) {
            size_t half = len / 2, mid = lo + half;
            if (cdf[base + mid] <= x) {
                lo = mid + 1;
                len -= half + 1;
            } else {
                len = half;
            }
        }
        size_t cell = std::min(lo > 0 ? lo - 1 : 0, n - 1);
        float c0 = cdf[base + cell], c1 = cdf[base + cell + 1];
        float span = c1 - c0;
        *frac_out = span > 0.0f ? clampf((x - c0) / span, 0.0f, 1.0f) : 0.0f;
        *cell_out = cell;
    }
    EnvSampleR sample(Vec2 u) const {
        EnvSampleR inv{Vec3(0, 1, 0), Vec3(), 0.0f};
        if (empty()) return inv;
        size_t j, i;
        float dv, du;
        sample_cdf(marginal, 0, h, u.y, &j, &dv);
        sample_cdf(conditional, j * (w + 1), w, u.x, &i, &du);
        Vec2 uv{((float)i + du) / (float)w, ((float)j + dv) / (float)h};
        Vec3 direction = direction_from_uv(uv.x, uv.y);
        float theta = ((float)j + 0.5f) / (float)h * PI;
        float st = std::sin(theta);
        if (st <= 0.0f) return inv;
        Vec3 rad = px(i, j);
        float weight = std::fmax(luminance(rad), 0.0f) * st;
        float pdf_uv = weight / total_weight * (float)(w * h);
        float pdf = pdf_uv / (2.0f * PI * PI * st);
        if (pdf <= 0.0f) return inv;
        return EnvSampleR{direction, rad, pdf};
    }
};

struct OScene {
    const GpuMaterial* materials;
    const GpuPrimitive* prims;
    size_t num_prims;
    const GpuInstance* instances = nullptr;  // instance k is primitive num_prims + k
    size_t num_instances = 0;
    uint32_t tlas_root = 0;
    BvhView bvh;
    const GpuVertexAttr* attrs;
    const GpuLight* lights;
    size_t num_lights;
    Env env;
    Vec3 background;
    uint32_t light_strategy_count() const { return (uint32_t)num_lights + (env.empty() ? 0u : 1u); }
};

static Hit shade_triangle(const OScene& sc, const TriHit& th, const Ray& ray) {
    const GpuTriangle& tri = sc.bvh.triangles[th.triangle];
    Vec3 p0 = pos3(sc.bvh.positions, tri.i0), p1 = pos3(sc.bvh.positions, tri.i1),
         p2 = pos3(sc.bvh.positions, tri.i2);
    float w = 1.0f - th.u - th.v;
    auto attr = [&](uint32_t i) { return Vec3::from(sc.attrs[i].normal); };
    Vec3 shading = normalize_or(w * attr(tri.i0) + th.u * attr(tri.i1) + th.v * attr(tri.i2), Vec3(0, 0, 1));
    Vec3 cr = cross(p1 - p0, p2 - p0);
    Vec3 geometric = normalize_or(cr, Vec3(0, 0, 1));
    bool front = dot(geometric, ray.dir) < 0.0f;
    float flip = front ? 1.0f : -1.0f;
    Hit h;
    h.t = th.t;
    h.steps = 0;
    h.position = ray.origin + th.t * ray.dir;
    h.normal = shading * flip;
    h.geometric_normal = geometric * flip;
    h.front_face = front;
    h.material = tri.material;
    h.light_area = 0.5f * length(cr);
    return h;
}

static bool shade_analytic(const OScene& sc, long best, float best_t, const Ray& ray, Hit* out) {
    if (best < 0) return false;
    const GpuPrimitive& prim = sc.prims[best];
    Vec3 position = ray.origin + best_t * ray.dir;
    bool quad = prim.kind == LUCIDA_PRIM_KIND_QUAD;
    float light_area = quad ? length(cross(Vec3::from(prim.edge_u), Vec3::from(prim.edge_v))) : 0.0f;
    Vec3 gn = quad ? Vec3::from(prim.normal) : (position - Vec3::from(prim.position)) / prim.radius;
    bool front = dot(gn, ray.dir) < 0.0f;
    Vec3 oriented = front ? gn : -gn;
    *out = Hit{best_t, 0, position, oriented, oriented, front, prim.material, light_area};
    return true;
}

// --- instancing (instance.rs:120-160, :342-410; scene.rs:375-402) -------------------

static Vec3 inst_point(const GpuInstance& I, Vec3 p) {
    return Vec3(dot(Vec3::from(I.inv_row0), p) + I.inv_translation[0],
                dot(Vec3::from(I.inv_row1), p) + I.inv_translation[1],
                dot(Vec3::from(I.inv_row2), p) + I.inv_translation[2]);
}
static Vec3 inst_vector(const GpuInstance& I, Vec3 v) {
    return Vec3(dot(Vec3::from(I.inv_row0), v), dot(Vec3::from(I.inv_row1), v), dot(Vec3::from(I.inv_row2), v));
}
static Vec3 inst_normal(const GpuInstance& I, Vec3 n) {
    const float* r0 = I.inv_row0;
    const float* r1 = I.inv_row1;
    const float* r2 = I.inv_row2;
    return normalize_or(Vec3(r0[0] * n.x + r1[0] * n.y + r2[0] * n.z, r0[1] * n.x + r1[1] * n.y + r2[1] * n.z,
                             r0[2] * n.x + r1[2] * n.y + r2[2] * n.z),
                        Vec3(0, 0, 1));
}

// InstanceSet::intersect (instance.rs:342): the TLAS walked in world space,
// every node's own box tested, left child first; an instance leaf sends the
// unnormalised object-space ray down its BLAS with `closest` carried over.
// With any_hit set it answers occlusion through both levels.
static bool instances_intersect(const OScene& sc, const Ray& ray, float t_min, float t_max, bool any_hit,
                                TriHit* out, uint32_t* inst_out, TraversalStats* stats) {
    const Vec3 inv = safe_inv_dir(ray.dir);
    float closest = t_max;
    bool found = false;
    uint32_t stack[32];
    size_t sp = 0;
    uint32_t node = sc.tlas_root;
    for (;;) {
        const GpuBvhNode& n = sc.bvh.nodes[node];
        if (stats) stats->node_visits += 1;
        Aabb box;
        box.min = Vec3::from(n.bounds_min);
        box.max = Vec3::from(n.bounds_max);
        float entry;
        if (box.hit(ray.origin, inv, t_min, closest, &entry)) {
            if (n.count > 0) {
                for (uint32_t k = 0; k < n.count; ++k) {
                    const uint32_t ii = n.left_first + k;
                    const GpuInstance& I = sc.instances[ii - sc.num_prims];
                    const Ray local{inst_point(I, ray.origin), inst_vector(I, ray.dir)};
                    if (any_hit) {
                        if (bvh_occluded(sc.bvh, I.blas_root, local, t_min, closest)) return true;
                        continue;
                    }
                    TriHit h;
                    if (bvh_intersect(sc.bvh, I.blas_root, local, t_min, closest, &h, stats)) {
                        closest = h.t;
                        *out = h;
                        *inst_out = ii;
                        found = true;
                    }
                }
            } else {
                if (sp + 1 < 32) stack[sp++] = n.left_first + 1;
                node = n.left_first;
                continue;
            }
        }
        if (sp == 0) break;
        node = stack[--sp];
    }
    return found;
}

static Hit shade_triangle(const OScene& sc, const TriHit& th, const Ray& ray);

static Hit shade_instanced_triangle(const OScene& sc, const TriHit& th, uint32_t ii, const Ray& ray) {
    const GpuInstance& I = sc.instances[ii - sc.num_prims];
    Ray local{inst_point(I, ray.origin), inst_vector(I, ray.dir)};
    Hit h = shade_triangle(sc, th, local);
    h.normal = inst_normal(I, h.normal);
    h.geometric_normal = inst_normal(I, h.geometric_normal);
    h.position = ray.origin + th.t * ray.dir;
    h.front_face = dot(h.geometric_normal, ray.dir) < 0.0f;
    const float flip = h.front_face ? 1.0f : -1.0f;
    h.normal = h.normal * flip;
    h.geometric_normal = h.geometric_normal * flip;
    if (I.material != 0xFFFFFFFFu) h.material = I.material;
    h.light_area = 0.0f;
    return h;
}

static bool scene_intersect(const OScene& sc, const Ray& ray, Hit* out, TraversalStats* stats) {
    float best_t = T_MAX;
    long best = -1;
    for (size_t i = 0; i < sc.num_prims; ++i) {
        const GpuPrimitive& p = sc.prims[i];
        float t;
        bool hit = p.kind == LUCIDA_PRIM_KIND_QUAD ? intersect_quad(p, ray, T_MIN, best_t, &t)
                                                   : intersect_sphere(p, ray, T_MIN, best_t, &t);
        if (hit) {
            best_t = t;
            best = (long)i;
        }
    }
    TriHit th;
    if (sc.num_instances) {
        // instead of the single-level path (scene.rs:475-497)
        uint32_t ii = 0;
        if (instances_intersect(sc, ray, T_MIN, best_t, false, &th, &ii, stats)) {
            *out = shade_instanced_triangle(sc, th, ii, ray);
            return true;
        }
        return shade_analytic(sc, best, best_t, ray, out);
    }
    bool tri_hit = sc.bvh.num_nodes == 0 ? brute_force_intersect(sc.bvh, ray, T_MIN, best_t, &th)
                                         : bvh_intersect(sc.bvh, 0, ray, T_MIN, best_t, &th, stats);
    if (tri_hit) {
        *out = shade_triangle(sc, th, ray);
        return true;
    }
    return shade_analytic(sc, best, best_t, ray, out);
}

static bool scene_occluded(const OScene& sc, const Ray& ray, float t_max) {
    for (size_t i = 0; i < sc.num_prims; ++i) {
        const GpuPrimitive& p = sc.prims[i];
        float t;
        bool hit = p.kind == LUCIDA_PRIM_KIND_QUAD ? intersect_quad(p, ray, T_MIN, t_max, &t)
                                                   : intersect_sphere(p, ray, T_MIN, t_max, &t);
        if (hit) return true;
    }
    if (sc.num_instances) {
        // Two-level any-hit. The reference's CPU twin ignores instances here
        // (scene.rs:577-606) and its GPU walks only the tree at node 0
        // (connect.wgsl:192-212); both levels are the occluders a shadow ray
        // meets in the instanced scene, and the device path does the same.
        TriHit th;
        uint32_t ii;
        return instances_intersect(sc, ray, T_MIN, t_max, true, &th, &ii, nullptr);
    }
    if (sc.bvh.num_nodes == 0) {
        TriHit th;
        return brute_force_intersect(sc.bvh, ray, T_MIN, t_max, &th);
    }
    return bvh_occluded(sc.bvh, 0, ray, T_MIN, t_max);
}

// --- light.rs -------------------------------------------------------------------------

static inline float power_heuristic(float a, float b) {
    if (a <= 0.0f) return 0.0f;
    float r = b / a;
    return 1.0f / (1.0f + r * r);
}
static inline float light_pdf_from_geometry(uint32_t num_lights, float area, Vec3 ln, Vec3 sp, Vec3 hp) {
    if (num_lights == 0 || area <= 0.0f) return 0.0f;
    Vec3 to_light = hp - sp;
    float dist_sq = length_squared(to_light);
    if (dist_sq <= 0.0f) return 0.0f;
    Vec3 wi = to_light / std::sqrt(dist_sq);
    float cos_light = dot(ln, -wi);
    if (cos_light <= 1.0e-6f) return 0.0f;
    float pdf_area = 1.0f / ((float)num_lights * area);
    return pdf_area * dist_sq / cos_light;
}
struct LightSampleR {
    Vec3 wi;
    float distance;
    Vec3 radiance;
    float pdf;
};
static bool sample_lights(const OScene& sc, Vec3 sp, float u_select, Vec2 u_area, LightSampleR* out) {
    if (sc.num_lights == 0) return false;
    size_t n = sc.num_lights;
    size_t index = std::min((size_t)f2u_sat(u_select * (float)n), n - 1);
    const GpuLight& L = sc.lights[index];
    Vec3 origin = Vec3::from(L.origin), eu = Vec3::from(L.edge_u), ev = Vec3::from(L.edge_v);
    Vec3 point;
    if (L.kind == LUCIDA_LIGHT_KIND_TRIANGLE) {
        float su = std::sqrt(u_area.x);
        float b1 = 1.0f - su, b2 = u_area.y * su;
        point = origin + b1 * eu + b2 * ev;
    } else {
        point = origin + u_area.x * eu + u_area.y * ev;
    }
    Vec3 to_light = point - sp;
    float dist_sq = length_squared(to_light);
    if (dist_sq <= 0.0f) return false;
    float distance = std::sqrt(dist_sq);
    Vec3 wi = to_light / distance;
    float cos_light = dot(Vec3::from(L.normal), -wi);
    if (cos_light <= 1.0e-6f) return false;
    float pdf_area = 1.0f / ((float)n * L.area);
    float pdf = pdf_area * dist_sq / cos_light;
    if (!std::isfinite(pdf) || pdf <= 0.0f) return false;
    *out = LightSampleR{wi, distance, Vec3::from(sc.materials[L.material].emissive), pdf};
    return true;
}

// --- integrator.rs ----------------------------------------------------------------------

static Vec3 direct_light(const OScene& sc, const Surface& surf, const Hit& hit, Vec3 wo, Sampler& rng, uint32_t mode,
                         uint64_t* shadow_rays) {
    float u_select = rng.next_f32();
    Vec2 u_area = rng.next_vec2();
    uint32_t n_lights = (uint32_t)sc.num_lights;
    uint32_t total = sc.light_strategy_count();
    if (total == 0) return Vec3();
    uint32_t pick = std::min(f2u_sat(u_select * (float)total), total - 1);
    Vec3 wi_world, radiance;
    float distance, pdf;
    if (pick < n_lights) {
        float u_light = ((float)pick + 0.5f) / (float)n_lights;
        LightSampleR ls;
        if (!sample_lights(sc, hit.position, u_light, u_area, &ls)) return Vec3();
        float scale = (float)n_lights / (float)total;
        wi_world = ls.wi;
        radiance = ls.radiance;
        distance = ls.distance;
        pdf = ls.pdf * scale;
    } else {
        EnvSampleR es = sc.env.sample(u_area);
        if (es.pdf <= 0.0f) return Vec3();
        wi_world = es.direction;
        radiance = es.radiance;
        distance = std::numeric_limits<float>::infinity();
        pdf = es.pdf / (float)total;
    }
    if (pdf <= 0.0f) return Vec3();
    Vec3 wi = to_local(wi_world, hit.normal);
    if (wi.z == 0.0f || (wi.z < 0.0f && surf.transmission <= 0.0f)) return Vec3();
    Vec3 f = surface_eval(surf, wo, wi);
    if (max_element(f) <= 0.0f) return Vec3();
    float mis = mode == LUCIDA_SAMPLING_MIS ? power_heuristic(pdf, surface_pdf(surf, wo, wi)) : 1.0f;
    if (mis <= 0.0f) return Vec3();
    Vec3 origin = offset_ray_origin(hit.position, wi.z < 0.0f ? -hit.geometric_normal : hit.geometric_normal);
    if (shadow_rays) *shadow_rays += 1;
    if (scene_occluded(sc, Ray{origin, wi_world}, distance * (1.0f - 1.0e-3f))) return Vec3();
    return f * std::fabs(wi.z) * radiance * (mis / pdf);
}

struct Guide {
    Vec3 albedo, normal;
    float depth = 0.0f, steps = 0.0f;
};

static Vec3 radiance_with_guide(const OScene& sc, Ray ray, Sampler& rng, uint32_t max_depth, uint32_t mode,
                                Guide* guide, RenderCounts* counts) {
    Vec3 radiance, throughput = Vec3::splat(1.0f);
    float prev_bsdf_pdf = 0.0f;
    Vec3 prev_position = ray.origin;
    for (uint32_t depth = 0; depth < max_depth; ++depth) {
        Hit hit;
        TraversalStats ts;
        if (counts) counts->extension_rays += 1;
        bool got = scene_intersect(sc, ray, &hit, &ts);
        if (counts) {
            counts->node_visits += ts.node_visits;
            counts->triangle_tests += ts.triangle_tests;
        }
        if (!got) {
            if (!sc.env.empty()) {
                Vec3 l = sc.env.radiance(ray.dir);
                float weight = 1.0f;
                if (depth > 0) {
                    if (mode == LUCIDA_SAMPLING_NEE) weight = 0.0f;
                    else if (mode == LUCIDA_SAMPLING_MIS) {
                        uint32_t total = std::max(sc.light_strategy_count(), 1u);
                        float pdf_env = sc.env.pdf(ray.dir) / (float)total;
                        weight = power_heuristic(prev_bsdf_pdf, pdf_env);
                    }
                }
                if (weight > 0.0f) radiance += throughput * l * weight;
            } else {
                radiance += throughput * sc.background;
            }
            break;
        }
        hit.steps = ts.node_visits;
        const GpuMaterial& m = sc.materials[hit.material];
        Vec3 emissive = Vec3::from(m.emissive);
        if (hit.front_face && max_element(emissive) > 0.0f) {
            float weight = 1.0f;
            if (depth > 0) {
                if (mode == LUCIDA_SAMPLING_NEE) weight = 0.0f;
                else if (mode == LUCIDA_SAMPLING_MIS) {
                    uint32_t n = (uint32_t)sc.num_lights;
                    uint32_t total = sc.light_strategy_count();
                    float pdf_light = light_pdf_from_geometry(n, hit.light_area, hit.geometric_normal, prev_position,
                                                              hit.position) *
                                      ((float)n / (float)std::max(total, 1u));
                    weight = power_heuristic(prev_bsdf_pdf, pdf_light);
                }
            }
            if (weight > 0.0f) radiance += throughput * emissive * weight;
        }
        Surface surf = Surface::from_material(m, hit.front_face);
        if (depth == 0) {
            guide->albedo = vmax(surf.diffuse_albedo + surf.f0 + emissive, Vec3::splat(1.0e-3f));
            guide->normal = hit.normal;
            guide->depth = hit.t;
            guide->steps = (float)hit.steps;
        }
        Vec3 wo = to_local(-ray.dir, hit.normal);
        if (wo.z <= 0.0f) break;
        bool has_env = !sc.env.empty();
        if (mode != LUCIDA_SAMPLING_BSDF && depth + 1 < max_depth && (sc.num_lights > 0 || has_env))
            radiance += throughput * direct_light(sc, surf, hit, wo, rng, mode, counts ? &counts->shadow_rays : nullptr);
        float u_lobe = rng.next_f32();
        Vec2 u = rng.next_vec2();
        BsdfSample s = surface_sample(surf, wo, u_lobe, u);
        if (!(s.pdf > 0.0f)) break;
        throughput *= s.weight;
        if (max_element(throughput) <= 0.0f) break;
        prev_bsdf_pdf = s.pdf;
        prev_position = hit.position;
        Vec3 dir = to_world(s.wi, hit.normal);
        ray = Ray{offset_ray_origin(hit.position, s.wi.z < 0.0f ? -hit.geometric_normal : hit.geometric_normal), dir};
    }
    return radiance;
}

static Vec3 generate_ray_dir(const GpuUniforms& u, uint32_t x, uint32_t y, Vec2 puv, Vec2 luv, Vec3* origin) {
    float s = ((float)x + puv.x) / (float)u.width;
    float t = ((float)y + puv.y) / (float)u.height;
    Vec3 o = Vec3::from(u.cam_origin);
    Vec3 target = Vec3::from(u.cam_upper_left) + s * Vec3::from(u.cam_horizontal) + t * Vec3::from(u.cam_vertical);
    if (u.lens_radius <= 0.0f) {
        *origin = o;
        return normalize(target - o);
    }
    Vec2 d = concentric_sample_disk(luv);
    d.x *= u.lens_radius;
    d.y *= u.lens_radius;
    Vec3 right = normalize(Vec3::from(u.cam_horizontal)), down = normalize(Vec3::from(u.cam_vertical));
    Vec3 off = right * d.x + down * d.y;
    o = o + off;
    *origin = o;
    return normalize(target - o);
}

static void render_pixel(const OScene& sc, const GpuUniforms& u, uint32_t x, uint32_t y, float* rgb, float* guide_out,
                         RenderCounts* counts) {
    uint32_t pixel = y * u.width + x;
    Vec3 sum;
    Guide acc;
    for (uint32_t s = 0; s < u.samples_per_launch; ++s) {
        Sampler rng(u.sampler_kind == LUCIDA_SAMPLER_SOBOL, u.frame_seed, pixel, u.sample_offset + s);
        Vec2 puv = rng.next_vec2();
        Vec2 luv = rng.next_vec2();
        Ray ray;
        ray.dir = generate_ray_dir(u, x, y, puv, luv, &ray.origin);
        uint32_t mode = u.sampling_mode > 2 ? 2 : u.sampling_mode;
        Guide g;
        sum += radiance_with_guide(sc, ray, rng, u.max_depth, mode, &g, counts);
        acc.albedo += g.albedo;
        acc.normal += g.normal;
        acc.depth += g.depth;
        acc.steps += g.steps;
    }
    float inv = 1.0f / (float)u.samples_per_launch;
    Vec3 r = sum * inv;
    r.store(rgb);
    if (guide_out) {
        (acc.albedo * inv).store(guide_out);
        (acc.normal * inv).store(guide_out + 3);
        guide_out[6] = acc.depth * inv;
        guide_out[7] = acc.steps * inv;
    }
}

// integrator.rs:635-699 — 32-row bands handed out through an atomic counter.
static OScene make_scene(const lucida_scene_desc& d, const float background[3]) {
    OScene sc;
    sc.materials = d.materials;
    sc.prims = d.primitives;
    sc.num_prims = d.num_primitives;
    sc.instances = d.instances;
    sc.num_instances = d.num_instances;
    sc.tlas_root = d.tlas_root;
    sc.bvh.nodes = d.bvh_nodes;
    sc.bvh.num_nodes = d.num_bvh_nodes;
    sc.bvh.prim_indices = nullptr;
    sc.bvh.triangles = d.triangles;
    sc.bvh.num_triangles = d.num_triangles;
    sc.bvh.positions = d.positions;
    sc.attrs = d.vertex_attrs;
    sc.lights = d.lights;
    sc.num_lights = d.num_lights;
    sc.env.w = d.env_width;
    sc.env.h = d.env_height;
    sc.env.pixels = d.env_pixels;
    sc.env.conditional = d.env_conditional;
    sc.env.marginal = d.env_marginal;
    sc.env.total_weight = d.env_total_weight;
    sc.background = Vec3::from(background);
    return sc;
}

// Closest hit (or occlusion) of one ray against the scene, without shading:
// the analytic primitives first, then the triangles through the instance
// hierarchy or the single-level BVH, as scene_intersect / scene_occluded.
// id = triangle, 0x80000000 | primitive, or 0xFFFFFFFF.
static void cast_one(const OScene& sc, const Ray& ray, float t_max, bool any_hit, float out[4], uint32_t* inst) {
    float best_t = t_max;
    long best = -1;
    *inst = 0;
    for (size_t i = 0; i < sc.num_prims; ++i) {
        const GpuPrimitive& p = sc.prims[i];
        float t;
        bool hit = p.kind == LUCIDA_PRIM_KIND_QUAD ? intersect_quad(p, ray, T_MIN, best_t, &t)
                                                   : intersect_sphere(p, ray, T_MIN, best_t, &t);
        if (hit) {
            best_t = t;
            best = (long)i;
            if (any_hit) break;
        }
    }
    uint32_t id = best >= 0 ? (0x80000000u | (uint32_t)best) : 0xFFFFFFFFu;
    float t = best >= 0 ? best_t : 0.0f, u = 0.0f, v = 0.0f;
    if (!(any_hit && best >= 0)) {
        TriHit th;
        bool got;
        if (sc.num_instances)
            got = instances_intersect(sc, ray, T_MIN, best_t, any_hit, &th, inst, nullptr);
        else if (any_hit)
            got = sc.bvh.num_nodes ? bvh_occluded(sc.bvh, 0, ray, T_MIN, best_t)
                                   : brute_force_intersect(sc.bvh, ray, T_MIN, best_t, &th);
        else
            got = sc.bvh.num_nodes ? bvh_intersect(sc.bvh, 0, ray, T_MIN, best_t, &th, nullptr)
                                   : brute_force_intersect(sc.bvh, ray, T_MIN, best_t, &th);
        if (got) {
            if (any_hit) {
                id = 0;
                t = 0.0f;
            } else {
                id = th.triangle;
                t = th.t;
                u = th.u;
                v = th.v;
            }
        }
    }
    out[0] = t;
    out[1] = u;
    out[2] = v;
    std::memcpy(&out[3], &id, 4);
}

void scene_cast(const lucida_scene_desc& d, const float* rays, size_t n, int any_hit, float* hits, uint32_t* insts) {
    const float bg[3] = {0, 0, 0};
    const OScene sc = make_scene(d, bg);
    for (size_t i = 0; i < n; ++i) {
        const float* r = rays + 8 * i;
        uint32_t inst = 0;
        cast_one(sc, Ray{Vec3(r[0], r[1], r[2]), Vec3(r[4], r[5], r[6])}, r[7], any_hit != 0, hits + 4 * i, &inst);
        if (insts) insts[i] = inst;
    }
}

void render(const lucida_scene_desc& d, const GpuUniforms& u, const float background[3], uint32_t threads,
            uint32_t row_begin, uint32_t row_end, float* rgb, float* guides, RenderCounts* counts) {
    const OScene sc = make_scene(d, background);
    ggx_table();
    if (threads == 0) threads = std::max(1u, std::thread::hardware_concurrency());
    const uint32_t BAND = 32;
    if (row_end > u.height) row_end = u.height;
    uint32_t bands = (row_end - row_begin + BAND - 1) / BAND;
    std::atomic<uint32_t> next{0};
    std::mutex mu;
    auto worker = [&]() {
        RenderCounts local{};
        for (;;) {
            uint32_t b = next.fetch_add(1);
            if (b >= bands) break;
            uint32_t y0 = row_begin + b * BAND, y1 = std::min(y0 + BAND, row_end);
            for (uint32_t y = y0; y < y1; ++y)
                for (uint32_t x = 0; x < u.width; ++x) {
                    size_t i = (size_t)(y - row_begin) * u.width + x;
                    render_pixel(sc, u, x, y, rgb + 3 * i, guides ? guides + 8 * i : nullptr, counts ? &local : nullptr);
                }
        }
        if (counts) {
            std::lock_guard<std::mutex> g(mu);
            counts->extension_rays += local.extension_rays;
            counts->shadow_rays += local.shadow_rays;
            counts->node_visits += local.node_visits;
            counts->triangle_tests += local.triangle_tests;
        }
    };
    std::vector<std::thread> pool;
    for (uint32_t t = 0; t < threads; ++t) pool.emplace_back(worker);
    for (auto& t : pool) t.join();
}

// The rows listed (each row one work item for the threads), written
// compactly: output row k is image row rows[k]. For a bounded CPU sample of
// an image that keeps every thread busy on rows spread over the frame.
void render_rows(const lucida_scene_desc& d, const GpuUniforms& u, const float background[3], uint32_t threads,
                 const uint32_t* rows, uint32_t n_rows, float* rgb, RenderCounts* counts) {
    const OScene sc = make_scene(d, background);
    ggx_table();
    if (threads == 0) threads = std::max(1u, std::thread::hardware_concurrency());
    std::atomic<uint32_t> next{0};
    std::mutex mu;
    auto worker = [&]() {
        RenderCounts local{};
        for (;;) {
            uint32_t k = next.fetch_add(1);
            if (k >= n_rows) break;
            const uint32_t y = rows[k];
            if (y >= u.height) continue;
            for (uint32_t x = 0; x < u.width; ++x)
                render_pixel(sc, u, x, y, rgb + 3 * ((size_t)k * u.width + x), nullptr, counts ? &local : nullptr);
        }
        if (counts) {
            std::lock_guard<std::mutex> g(mu);
            counts->extension_rays += local.extension_rays;
            counts->shadow_rays += local.shadow_rays;
            counts->node_visits += local.node_visits;
            counts->triangle_tests += local.triangle_tests;
        }
    };
    std::vector<std::thread> pool;
    for (uint32_t t = 0; t < std::min(threads, n_rows); ++t) pool.emplace_back(worker);
    for (auto& t : pool) t.join();
}

// --- denoise.rs ---------------------------------------------------------------------------

void denoise(const float* film, const float* albedo, const float* normal, const float* depth, uint32_t width,
             uint32_t height, uint32_t iterations, float sn, float sd, float sl, float sa, float* out) {
    const int w = (int)width, h = (int)height;
    const size_t n = (size_t)w * h;
    auto V = [](const float* p, size_t i) { return Vec3(p[3 * i], p[3 * i + 1], p[3 * i + 2]); };
    std::vector<Vec3> cur(n), next(n);
    for (size_t i = 0; i < n; ++i) cur[i] = V(film, i) / vmax(V(albedo, i), Vec3::splat(1.0e-3f));
    next = cur;
    std::vector<float> dev(n);
    for (size_t i = 0; i < n; ++i) {
        int x = (int)(i % width), y = (int)(i / width);
        float sum = 0, sum_sq = 0, cnt = 0;
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                int px = x + dx, py = y + dy;
                if (px < 0 || py < 0 || px >= w || py >= h) continue;
                float l = luminance(cur[(size_t)py * w + px]);
                sum += l;
                sum_sq += l * l;
                cnt += 1.0f;
            }
        float mean = sum / cnt;
        dev[i] = std::sqrt(std::fmax(sum_sq / cnt - mean * mean, 0.0f));
    }
    const float K[5] = {1.0f / 16.0f, 1.0f / 4.0f, 3.0f / 8.0f, 1.0f / 4.0f, 1.0f / 16.0f};
    for (uint32_t level = 0; level < iterations; ++level) {
        int stride = 1 << level;
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < w; ++x) {
                size_t ci = (size_t)y * w + x;
                Vec3 cn = V(normal, ci);
                float cd = depth[ci];
                float cl = luminance(cur[ci]);
                float ca = std::fmax(luminance(V(albedo, ci)), 1.0e-3f);
                float lum_scale = sl * std::fmax(dev[ci], 1.0e-4f);
                Vec3 sum;
                float wsum = 0.0f;
                for (int ky = 0; ky < 5; ++ky)
                    for (int kx = 0; kx < 5; ++kx) {
                        int px = x + (kx - 2) * stride, py = y + (ky - 2) * stride;
                        if (px < 0 || py < 0 || px >= w || py >= h) continue;
                        size_t qi = (size_t)py * w + px;
                        float nd = std::fmax(dot(cn, V(normal, qi)), 0.0f);
                        float w_n = std::pow(nd, sn);
                        float dd = std::fabs(cd - depth[qi]);
                        float w_d = std::exp(-dd / (sd * std::fmax(std::fabs(cd), 1.0e-3f)));
                        float dl = std::fabs(cl - luminance(cur[qi]));
                        float w_l = std::exp(-dl / lum_scale);
                        float da = std::fabs(ca - luminance(V(albedo, qi)));
                        float w_a = std::exp(-da / (sa * ca));
                        float weight = K[kx] * K[ky] * w_n * w_d * w_l * w_a;
                        sum += cur[qi] * weight;
                        wsum += weight;
                    }
                next[ci] = wsum > 0.0f ? sum / wsum : cur[ci];
            }
        std::swap(cur, next);
    }
    for (size_t i = 0; i < n; ++i) (cur[i] * vmax(V(albedo, i), Vec3::splat(1.0e-3f))).store(out + 3 * i);
}

}  // namespace oracle

using namespace oracle;

extern "C" {

uint32_t oracle_sobol(uint32_t index, uint32_t dim) { return sobol(index, dim); }
uint32_t oracle_owen_scramble(uint32_t x, uint32_t seed) { return owen_scramble(x, seed); }
uint32_t oracle_hash_u32(uint32_t x) { return hash_u32(x); }
float oracle_sobol_sample(uint32_t index, uint32_t dim, uint32_t seed) { return sobol_sample(index, dim, seed); }

void oracle_sampler_draws(uint32_t kind, uint32_t frame, uint32_t pixel, uint32_t sample, const uint32_t* pattern,
                          size_t n, float* out) {
    Sampler s(kind == LUCIDA_SAMPLER_SOBOL, frame, pixel, sample);
    size_t k = 0;
    for (size_t i = 0; i < n; ++i) {
        if (pattern[i] == 2) {
            Vec2 v = s.next_vec2();
            out[k++] = v.x;
            out[k++] = v.y;
        } else {
            out[k++] = s.next_f32();
        }
    }
}

void oracle_ggx_table(float* out) { std::memcpy(out, ggx_table(), sizeof(float) * 1024); }

float oracle_directional_albedo_grid(float roughness, float cos_o, uint32_t grid) {
    return directional_albedo_grid(roughness_to_alpha(roughness), cos_o, grid);
}

void oracle_offset_ray_origin(const float p[3], const float n[3], float out[3]) {
    offset_ray_origin(Vec3::from(p), Vec3::from(n)).store(out);
}

void oracle_render(const lucida_scene_desc* d, const GpuUniforms* u, const float background[3], uint32_t threads,
                   uint32_t row_begin, uint32_t row_end, float* rgb, float* guides, RenderCounts* counts) {
    render(*d, *u, background, threads, row_begin, row_end, rgb, guides, counts);
}

void oracle_render_rows(const lucida_scene_desc* d, const GpuUniforms* u, const float background[3], uint32_t threads,
                        const uint32_t* rows, uint32_t n_rows, float* rgb, RenderCounts* counts) {
    oracle::render_rows(*d, *u, background, threads, rows, n_rows, rgb, counts);
}

void oracle_scene_cast(const lucida_scene_desc* d, const float* rays, size_t n, int any_hit, float* hits,
                       uint32_t* instances) {
    scene_cast(*d, rays, n, any_hit, hits, instances);
}

void oracle_denoise(const float* film, const float* albedo, const float* normal, const float* depth, uint32_t width,
                    uint32_t height, const lucida_denoise_params* p, float* out) {
    denoise(film, albedo, normal, depth, width, height, p->iterations, p->sigma_normal, p->sigma_depth,
            p->sigma_luminance, p->sigma_albedo, out);
}

void oracle_compare(const float* a, const float* b, size_t pixels, double* out6) {
    double n = (double)pixels * 3.0;
    double sum_sq = 0, sum_rel = 0, sum_a = 0, sum_b = 0, max_abs = 0, sum_abs = 0;
    for (size_t i = 0; i < pixels * 3; ++i) {
        double va = a[i], vb = b[i];
        double d = std::fabs(va - vb);
        sum_sq += d * d;
        sum_rel += d / (std::fabs(va) + std::fabs(vb) + 1e-4);
        sum_a += va;
        sum_b += vb;
        sum_abs += d;
        if (d > max_abs) max_abs = d;
    }
    out6[0] = std::sqrt(sum_sq / n);  // rmse
    out6[1] = max_abs;
    out6[2] = sum_rel / n;  // mean_rel
    out6[3] = sum_a / n;
    out6[4] = sum_b / n;
    out6[5] = sum_abs / n;  // mean |a - b| per channel
}

}  // extern "C"

extern "C" {

float oracle_bsdf_sample(const GpuMaterial* m, uint32_t front, const float wo[3], float u_lobe, float ux, float uy,
                         float out[6]) {
    Surface s = Surface::from_material(*m, front != 0);
    BsdfSample b = surface_sample(s, Vec3::from(wo), u_lobe, Vec2{ux, uy});
    b.wi.store(out);
    b.weight.store(out + 3);
    return b.pdf;
}

float oracle_bsdf_eval_pdf(const GpuMaterial* m, uint32_t front, const float wo[3], const float wi[3], float f_out[3]) {
    Surface s = Surface::from_material(*m, front != 0);
    surface_eval(s, Vec3::from(wo), Vec3::from(wi)).store(f_out);
    return surface_pdf(s, Vec3::from(wo), Vec3::from(wi));
}

}  // extern "C"

extern "C" {

// --- closed-form checks and the environment sampler (bsdf.rs, envmap.rs) ---

float oracle_fresnel_dielectric(float cos_i, float eta) { return fresnel_dielectric(cos_i, eta); }

void oracle_fresnel_conductor(float cos_i, const float eta[3], const float k[3], float out[3]) {
    fresnel_conductor(cos_i, Vec3::from(eta), Vec3::from(k)).store(out);
}

// bsdf.rs:555 refract: 1 and the transmitted direction, or 0 under total
// internal reflection.
int oracle_refract(const float wo[3], const float n[3], float eta, float out[3]) {
    Vec3 wt;
    if (!refract(Vec3::from(wo), Vec3::from(n), eta, &wt)) return 0;
    wt.store(out);
    return 1;
}

// The reflected and transmitted halves of a surface's directional albedo:
// the mean over n samples of surface::sample's luminance weight, split by the
// side of the surface wi lands on (crates/core/tests/dielectric.rs
// total_albedo), with a PCG stream Rng::new(seed, 0, 0) drawing u_lobe then
// the 2D sample.
void oracle_bsdf_split_albedo(const GpuMaterial* m, uint32_t front, const float wo[3], uint32_t n, uint32_t seed,
                              double out[2]) {
    Surface s = Surface::from_material(*m, front != 0);
    Rng rng(seed, 0, 0);
    double up = 0.0, down = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
        float ul = rng.next_f32();
        float ux = rng.next_f32();
        float uy = rng.next_f32();
        BsdfSample b = surface_sample(s, Vec3::from(wo), ul, Vec2{ux, uy});
        if (!(b.pdf > 0.0f)) continue;
        double w = (double)luminance(b.weight);
        if (b.wi.z > 0.0f) up += w; else down += w;
    }
    out[0] = up / (double)n;
    out[1] = down / (double)n;
}

// EnvMap::new (envmap.rs:112): the oracle's own build of the conditional
// (w + 1 per row) and marginal (h + 1) CDFs over luminance * sin(theta at the
// row centre); returns total_weight.
float oracle_env_build(uint32_t width, uint32_t height, const float* pixels, float* conditional, float* marginal) {
    const size_t w = width, h = height;
    std::fill(conditional, conditional + (w + 1) * h, 0.0f);
    std::fill(marginal, marginal + h + 1, 0.0f);
    std::vector<float> row_weight(h, 0.0f);
    for (size_t j = 0; j < h; ++j) {
        const float sin_theta = std::sin(((float)j + 0.5f) / (float)h * PI);
        const size_t base = j * (w + 1);
        float running = 0.0f;
        for (size_t i = 0; i < w; ++i) {
            const float* p = &pixels[4 * (j * w + i)];
            running += std::fmax(luminance(Vec3(p[0], p[1], p[2])), 0.0f) * sin_theta;
            conditional[base + i + 1] = running;
        }
        row_weight[j] = running;
        if (running > 0.0f) {
            const float inv = 1.0f / running;
            for (size_t k = 1; k <= w; ++k) conditional[base + k] *= inv;
            conditional[base + w] = 1.0f;
        }
    }
    float total = 0.0f;
    for (size_t j = 0; j < h; ++j) {
        total += row_weight[j];
        marginal[j + 1] = total;
    }
    if (total > 0.0f) {
        const float inv = 1.0f / total;
        for (size_t k = 1; k <= h; ++k) marginal[k] *= inv;
        marginal[h] = 1.0f;
    }
    return total;
}

static Env env_view(uint32_t w, uint32_t h, const float* pixels, const float* cond, const float* marg, float total) {
    Env e;
    e.w = w;
    e.h = h;
    e.pixels = pixels;
    e.conditional = cond;
    e.marginal = marg;
    e.total_weight = total;
    return e;
}

// EnvMap::pdf / radiance (envmap.rs:211 / :188) at a direction.
float oracle_env_pdf(uint32_t w, uint32_t h, const float* pixels, const float* cond, const float* marg, float total,
                     const float dir[3], float radiance_out[3]) {
    Env e = env_view(w, h, pixels, cond, marg, total);
    if (radiance_out) e.radiance(Vec3::from(dir)).store(radiance_out);
    return e.pdf(Vec3::from(dir));
}

// EnvMap::sample (envmap.rs:235): out = direction[3], radiance[3]; returns pdf.
float oracle_env_sample(uint32_t w, uint32_t h, const float* pixels, const float* cond, const float* marg, float total,
                        float u, float v, float out[6]) {
    Env e = env_view(w, h, pixels, cond, marg, total);
    EnvSampleR r = e.sample(Vec2{u, v});
    r.direction.store(out);
    r.radiance.store(out + 3);
    return r.pdf;
}

// EnvMap::direction_from_uv / uv_from_direction (envmap.rs:89 / :98).
void oracle_env_direction_from_uv(float u, float v, float out[3]) { Env::direction_from_uv(u, v).store(out); }
void oracle_env_uv_from_direction(const float d[3], float out[2]) {
    Vec2 uv = Env::uv_from_direction(Vec3::from(d));
    out[0] = uv.x;
    out[1] = uv.y;
}

// convergence.wgsl (the reduce / finish passes of crates/gpu/src/stats.rs)
// restated per pixel in f32, summed in double: the mean over pixels with
// N >= 2 and a mean luminance above 1e-3 of sqrt(luminance(var) / N) /
// luminance(mean), var = radiance_sq / N - mean^2 clamped at zero; -1 when
// no pixel qualifies. accum holds 16 floats (one GpuAccum) per pixel.
double oracle_convergence(const float* accum, size_t pixels) {
    double err = 0.0, cnt = 0.0;
    for (size_t i = 0; i < pixels; ++i) {
        const float* a = accum + 16 * i;
        const float n = a[3];
        if (!(n >= 2.0f)) continue;
        const Vec3 mean = Vec3(a[0], a[1], a[2]) / n;
        const float mean_l = luminance(mean);
        if (!(mean_l > 1.0e-3f)) continue;
        const Vec3 var = vmax(Vec3(a[12], a[13], a[14]) / n - mean * mean, Vec3());
        err += (double)(std::sqrt(luminance(var) / n) / mean_l);
        cnt += 1.0;
    }
    return cnt > 0.0 ? err / cnt : -1.0;
}

// Directional albedo of a material's full surface BSDF at cos_o, estimated
// from n one-sample draws of surface::sample (bsdf.rs:1571) with a PCG stream.
// A white, energy-compensated lobe integrates to 1 (the white furnace).
double oracle_bsdf_albedo(const GpuMaterial* m, float cos_o, uint32_t n, uint32_t seed) {
    Surface s = Surface::from_material(*m, true);
    float sin_o = std::sqrt(std::fmax(1.0f - cos_o * cos_o, 0.0f));
    Vec3 wo(sin_o, 0.0f, cos_o);
    Rng rng(seed, 0, 0);
    double sum = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
        float ul = rng.next_f32();
        float ux = rng.next_f32();
        float uy = rng.next_f32();
        BsdfSample b = surface_sample(s, wo, ul, Vec2{ux, uy});
        if (b.pdf > 0.0f) sum += (double)b.weight.x;
    }
    return sum / (double)n;
}

}  // extern "C"
