// Host-side scene generation: the inputs of the hot path, laid out per the
// reference's crates/core/src/gpu_layout.rs.
//
// Restated from the reference: GpuMaterial constructors (gpu_layout.rs:74-160),
// GpuPrimitive::quad/sphere (gpu_layout.rs:214-240), MeshBuilder::append_to,
// quad and icosphere (mesh.rs:20-249), build_lights (light.rs:62-120),
// Camera::look_at / basis / write_uniforms (camera.rs:28-80), EnvMap::new
// (envmap.rs:112-180), procedural_sky (envmap.rs:343), the reference's nine
// built-in scenes (scenes.rs:56-815) and the rng.rs PCG stream. The two
// benchmark scenes ("box-c1", "sphere-c2") are new generators written in the
// same idioms.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/lucida_scenes.h"
#include "blob.hpp"
#include "hvec.hpp"
#include "status.hpp"

namespace lucida {

// --- materials and primitives (gpu_layout.rs) ------------------------------

static GpuMaterial mat_diffuse(f3 c) {
    GpuMaterial m{};
    c.store(m.base_color);
    m.roughness = 1.0f;
    m.ior = 1.5f;
    return m;
}
static GpuMaterial mat_glossy(f3 c, float roughness) {
    GpuMaterial m{};
    c.store(m.base_color);
    m.roughness = roughness;
    m.ior = 1.5f;
    return m;
}
static GpuMaterial mat_emissive(f3 radiance) {
    GpuMaterial m{};
    radiance.store(m.emissive);
    m.roughness = 1.0f;
    m.ior = 1.0f;
    return m;
}
static GpuMaterial mat_metal(f3 c, float roughness) {
    GpuMaterial m{};
    c.store(m.base_color);
    m.metallic = 1.0f;
    m.roughness = roughness;
    m.ior = 1.5f;
    return m;
}
static GpuMaterial mat_glass(f3 tint, float roughness, float ior) {
    GpuMaterial m{};
    tint.store(m.base_color);
    m.roughness = roughness;
    m.ior = ior;
    m.transmission = 1.0f;
    return m;
}
// bsdf.rs:358 fresnel_conductor, needed on the host for the base colour the
// conductor constructor stores (its normal-incidence reflectance).
static f3 fresnel_conductor_host(float cos_i, f3 eta, f3 k) {
    const float c = std::min(std::max(cos_i, 0.0f), 1.0f);
    const float c2 = c * c, sin2 = 1.0f - c2;
    f3 out;
    for (int a = 0; a < 3; ++a) {
        const float e = eta[a], kk = k[a];
        const float eta2 = e * e, k2 = kk * kk;
        const float t0 = eta2 - k2 - sin2;
        const float a2pb2 = std::pow(std::max(t0 * t0 + 4.0f * eta2 * k2, 0.0f), 0.5f);
        const float t1 = a2pb2 + c2;
        const float aa = std::pow(std::max(0.5f * (a2pb2 + t0), 0.0f), 0.5f);
        const float t2 = 2.0f * c * aa;
        const float rs = (t1 - t2) / (t1 + t2);
        const float t3 = c2 * a2pb2 + sin2 * sin2;
        const float t4 = t2 * sin2;
        const float rp = rs * (t3 - t4) / (t3 + t4);
        (a == 0 ? out.x : (a == 1 ? out.y : out.z)) = 0.5f * (rs + rp);
    }
    return out;
}
static GpuMaterial mat_conductor(f3 eta, f3 k, float roughness) {
    GpuMaterial m{};
    fresnel_conductor_host(1.0f, eta, k).store(m.base_color);
    m.metallic = 1.0f;
    m.roughness = roughness;
    m.ior = 1.5f;
    eta.store(m.eta);
    k.store(m.k);
    return m;
}
static GpuPrimitive prim_quad(f3 origin, f3 eu, f3 ev, uint32_t material) {
    GpuPrimitive p{};
    origin.store(p.position);
    eu.store(p.edge_u);
    p.kind = LUCIDA_PRIM_KIND_QUAD;
    ev.store(p.edge_v);
    p.material = material;
    normalize(cross(eu, ev)).store(p.normal);
    return p;
}
static GpuPrimitive prim_sphere(f3 c, float r, uint32_t material) {
    GpuPrimitive p{};
    c.store(p.position);
    p.radius = r;
    p.kind = LUCIDA_PRIM_KIND_SPHERE;
    p.material = material;
    return p;
}

// --- meshes (mesh.rs) ---------------------------------------------------------

struct Mesh {
    std::vector<f3> positions, normals;
    std::vector<float> uvs;  // 2 per vertex
    std::vector<uint32_t> indices;  // 3 per triangle
    uint32_t push_vertex(f3 p, f3 n, float u, float v) {
        positions.push_back(p);
        normals.push_back(n);
        uvs.push_back(u);
        uvs.push_back(v);
        return (uint32_t)positions.size() - 1;
    }
    void append_to(Blob& b, uint32_t material) const {
        uint32_t base = (uint32_t)(b.positions.size() / 4);
        for (size_t i = 0; i < positions.size(); ++i) {
            b.positions.insert(b.positions.end(), {positions[i].x, positions[i].y, positions[i].z, 0.0f});
            GpuVertexAttr a{};
            normals[i].store(a.normal);
            a.uv[0] = uvs[2 * i];
            a.uv[1] = uvs[2 * i + 1];
            b.vertex_attrs.push_back(a);
        }
        for (size_t t = 0; t < indices.size(); t += 3)
            b.triangles.push_back(GpuTriangle{base + indices[t], base + indices[t + 1], base + indices[t + 2], material});
    }
};

// mesh.rs:80-95
static Mesh mesh_quad(f3 origin, f3 eu, f3 ev) {
    f3 n = normalize(cross(eu, ev));
    Mesh m;
    m.push_vertex(origin, n, 0, 0);
    m.push_vertex(origin + eu, n, 1, 0);
    m.push_vertex(origin + eu + ev, n, 1, 1);
    m.push_vertex(origin + ev, n, 0, 1);
    m.indices = {0, 1, 2, 0, 2, 3};
    return m;
}

// mesh.rs:169-249 — subdivided icosahedron, midpoints deduplicated per level.
static Mesh mesh_icosphere(f3 center, float radius, uint32_t levels) {
    const float t = (1.0f + std::sqrt(5.0f)) / 2.0f;
    const f3 base[12] = {f3(-1, t, 0), f3(1, t, 0),  f3(-1, -t, 0), f3(1, -t, 0), f3(0, -1, t), f3(0, 1, t),
                         f3(0, -1, -t), f3(0, 1, -t), f3(t, 0, -1),  f3(t, 0, 1),  f3(-t, 0, -1), f3(-t, 0, 1)};
    static const uint32_t faces0[20][3] = {{0, 11, 5}, {0, 5, 1},  {0, 1, 7},   {0, 7, 10}, {0, 10, 11},
                                           {1, 5, 9},  {5, 11, 4}, {11, 10, 2}, {10, 7, 6}, {7, 1, 8},
                                           {3, 9, 4},  {3, 4, 2},  {3, 2, 6},   {3, 6, 8},  {3, 8, 9},
                                           {4, 9, 5},  {2, 4, 11}, {6, 2, 10},  {8, 6, 7},  {9, 8, 1}};
    std::vector<f3> unit;
    for (auto& v : base) unit.push_back(normalize(v));
    std::vector<uint32_t> faces;
    for (auto& f : faces0) faces.insert(faces.end(), {f[0], f[1], f[2]});
    for (uint32_t lv = 0; lv < levels; ++lv) {
        std::unordered_map<uint64_t, uint32_t> cache;
        std::vector<uint32_t> next;
        next.reserve(faces.size() * 4);
        auto midpoint = [&](uint32_t a, uint32_t b) -> uint32_t {
            uint64_t key = a < b ? ((uint64_t)a << 32 | b) : ((uint64_t)b << 32 | a);
            auto it = cache.find(key);
            if (it != cache.end()) return it->second;
            f3 m = normalize((unit[a] + unit[b]) * 0.5f);
            unit.push_back(m);
            uint32_t i = (uint32_t)unit.size() - 1;
            cache.emplace(key, i);
            return i;
        };
        for (size_t f = 0; f < faces.size(); f += 3) {
            uint32_t f0 = faces[f], f1 = faces[f + 1], f2 = faces[f + 2];
            uint32_t a = midpoint(f0, f1);
            uint32_t b = midpoint(f1, f2);
            uint32_t c = midpoint(f2, f0);
            next.insert(next.end(), {f0, a, c, f1, b, a, f2, c, b, a, b, c});
        }
        faces.swap(next);
    }
    Mesh m;
    const float TAU_F = 6.28318530717958647692f, PI_F = 3.14159265358979323846f;
    for (auto& u : unit) {
        float uu = 0.5f + std::atan2(u.z, u.x) / TAU_F;
        float vv = 0.5f - std::asin(u.y) / PI_F;
        m.push_vertex(center + u * radius, u, uu, vv);
    }
    m.indices = faces;
    return m;
}

// --- lights (light.rs:62-120) -----------------------------------------------

static void build_lights(Blob& b) {
    b.lights.clear();
    auto emits = [&](uint32_t m) {
        const float* e = b.materials[m].emissive;
        return e[0] > 0.0f || e[1] > 0.0f || e[2] > 0.0f;
    };
    for (const auto& p : b.primitives) {
        if (p.kind != LUCIDA_PRIM_KIND_QUAD || !emits(p.material)) continue;
        f3 eu(p.edge_u[0], p.edge_u[1], p.edge_u[2]), ev(p.edge_v[0], p.edge_v[1], p.edge_v[2]);
        GpuLight l{};
        std::memcpy(l.origin, p.position, 12);
        l.area = length(cross(eu, ev));
        std::memcpy(l.edge_u, p.edge_u, 12);
        l.kind = LUCIDA_LIGHT_KIND_QUAD;
        std::memcpy(l.edge_v, p.edge_v, 12);
        l.material = p.material;
        std::memcpy(l.normal, p.normal, 12);
        b.lights.push_back(l);
    }
    for (const auto& t : b.triangles) {
        if (!emits(t.material)) continue;
        auto P = [&](uint32_t i) { return f3(b.positions[4 * i], b.positions[4 * i + 1], b.positions[4 * i + 2]); };
        f3 p0 = P(t.i0), p1 = P(t.i1), p2 = P(t.i2);
        f3 e1 = p1 - p0, e2 = p2 - p0;
        f3 c = cross(e1, e2);
        float len = length(c);
        if (len <= 0.0f) continue;
        GpuLight l{};
        p0.store(l.origin);
        l.area = 0.5f * len;
        e1.store(l.edge_u);
        l.kind = LUCIDA_LIGHT_KIND_TRIANGLE;
        e2.store(l.edge_v);
        l.material = t.material;
        (c / len).store(l.normal);
        b.lights.push_back(l);
    }
}

// --- camera (camera.rs:28-80) -------------------------------------------------

static void camera_look_at(Blob& b, f3 eye, f3 look_at, float vfov_deg) {
    b.cam_eye = eye;
    b.cam_look_at = look_at;
    b.cam_up = f3(0, 1, 0);
    b.cam_vfov_deg = vfov_deg;
    b.cam_aperture = 0.0f;
    b.cam_focus_distance = length(look_at - eye);
}

// Camera::basis + Camera::write_uniforms (camera.rs:44-80).
static void write_camera(f3 eye, f3 look_at, f3 up_hint, float vfov_deg, float aperture, float focus_distance,
                         GpuUniforms& u, float aspect) {
    f3 forward = normalize(look_at - eye);
    f3 right = normalize(cross(forward, up_hint));
    f3 up = cross(right, forward);
    f3 down = -up;
    const float DEG = 3.14159265358979323846f / 180.0f;
    float half_height = std::tan(vfov_deg * DEG * 0.5f) * focus_distance;
    float half_width = half_height * aspect;
    f3 horizontal = right * (2.0f * half_width);
    f3 vertical = down * (2.0f * half_height);
    f3 upper_left = eye + forward * focus_distance - right * half_width - down * half_height;
    eye.store(u.cam_origin);
    upper_left.store(u.cam_upper_left);
    horizontal.store(u.cam_horizontal);
    vertical.store(u.cam_vertical);
    u.lens_radius = aperture * 0.5f;
    u.focus_distance = focus_distance;
}

static void camera_write_uniforms(const Blob& b, GpuUniforms& u, float aspect) {
    write_camera(b.cam_eye, b.cam_look_at, b.cam_up, b.cam_vfov_deg, b.cam_aperture, b.cam_focus_distance, u, aspect);
}

// --- environment map (envmap.rs:112-180) --------------------------------------

static void env_build(Blob& b, uint32_t width, uint32_t height, const float* rgba) {
    size_t w = width, h = height;
    b.env_width = width;
    b.env_height = height;
    b.env_pixels.assign(rgba, rgba + 4 * w * h);
    b.env_conditional.assign((w + 1) * h, 0.0f);
    b.env_marginal.assign(h + 1, 0.0f);
    std::vector<float> row_weight(h, 0.0f);
    const float PI_F = 3.14159265358979323846f;
    for (size_t j = 0; j < h; ++j) {
        float theta = ((float)j + 0.5f) / (float)h * PI_F;
        float sin_theta = std::sin(theta);
        size_t base = j * (w + 1);
        float running = 0.0f;
        for (size_t i = 0; i < w; ++i) {
            const float* p = &rgba[4 * (j * w + i)];
            float lum = p[0] * 0.2126f + p[1] * 0.7152f + p[2] * 0.0722f;
            float weight = std::fmax(lum, 0.0f) * sin_theta;
            running += weight;
            b.env_conditional[base + i + 1] = running;
        }
        row_weight[j] = running;
        if (running > 0.0f) {
            float inv = 1.0f / running;
            for (size_t k = 1; k <= w; ++k) b.env_conditional[base + k] *= inv;
            b.env_conditional[base + w] = 1.0f;
        }
    }
    float running = 0.0f;
    for (size_t j = 0; j < h; ++j) {
        running += row_weight[j];
        b.env_marginal[j + 1] = running;
    }
    b.env_total_weight = running;
    if (running > 0.0f) {
        float inv = 1.0f / running;
        for (size_t k = 1; k <= h; ++k) b.env_marginal[k] *= inv;
        b.env_marginal[h] = 1.0f;
    }
}

// procedural_sky (envmap.rs:343): a sky gradient over a dim ground with a sun
// disc widened to at least one texel and dimmed by the area ratio, then the
// CDF build. Pixels' fourth channel is 0, as in the reference.
static void procedural_sky(Blob& b, uint32_t width, uint32_t height, f3 sun_direction, float sun_intensity,
                           float sun_radius_degrees) {
    const float PI_F = 3.14159265358979323846f, TAU_F = 6.28318530717958647692f;
    const f3 sun = normalize_or(sun_direction, f3(0, 1, 0));
    const float texel = std::max(PI_F / (float)height, TAU_F / (float)width);
    const float requested = sun_radius_degrees * (PI_F / 180.0f);
    const float radius = std::max(requested, texel);
    const float area_ratio = (1.0f - std::cos(requested)) / std::max(1.0f - std::cos(radius), 1e-20f);
    const float intensity = sun_intensity * std::min(area_ratio, 1.0f);
    const float cos_sun_radius = std::cos(radius);
    std::vector<float> px;
    px.reserve((size_t)width * height * 4);
    for (uint32_t j = 0; j < height; ++j) {
        const float v = ((float)j + 0.5f) / (float)height;
        for (uint32_t i = 0; i < width; ++i) {
            const float u = ((float)i + 0.5f) / (float)width;
            const float theta = v * PI_F, phi = u * TAU_F;
            const float st = std::sin(theta), ct = std::cos(theta);
            const f3 d(st * std::cos(phi), ct, st * std::sin(phi));
            f3 c;
            if (d.y >= 0.0f) {
                const float t = std::pow(d.y, 0.45f);
                c = f3(0.75f, 0.86f, 1.00f) * t + f3(0.85f, 0.70f, 0.55f) * (1.0f - t);
            } else {
                c = f3::splat(0.12f);
            }
            c = c * 0.40f;
            if (dot(d, sun) >= cos_sun_radius) c = c + f3(1.0f, 0.92f, 0.78f) * intensity;
            px.insert(px.end(), {c.x, c.y, c.z, 0.0f});
        }
    }
    env_build(b, width, height, px.data());
}

// rng.rs:28-85 PCG-RXS-M-XS, counter-seeded; the instance scatter draws from it.
struct Pcg32 {
    uint32_t state;
    static uint32_t hash(uint32_t x) {
        uint32_t st = x * 747796405u + 2891336453u;
        uint32_t w = ((st >> ((st >> 28) + 4u)) ^ st) * 277803737u;
        return (w >> 22) ^ w;
    }
    Pcg32(uint32_t frame, uint32_t pixel, uint32_t sample) {
        uint32_t h = hash(frame);
        h = hash(h + pixel);
        state = hash(h + sample);
    }
    uint32_t next_u32() {
        state = state * 747796405u + 2891336453u;
        uint32_t w = ((state >> ((state >> 28) + 4u)) ^ state) * 277803737u;
        return (w >> 22) ^ w;
    }
    float next_f32() { return (float)(next_u32() >> 8) * (1.0f / 16777216.0f); }
};

// --- reference scenes (scenes.rs:56-815) ---------------------------------------

static void cornell_box_ref(Blob& b) {
    const float S = 555.0f;
    b.materials = {mat_diffuse(f3(0.725f, 0.710f, 0.680f)), mat_diffuse(f3(0.630f, 0.065f, 0.050f)),
                   mat_diffuse(f3(0.140f, 0.450f, 0.091f)), mat_emissive(f3(18.4f, 15.6f, 8.0f)),
                   mat_diffuse(f3::splat(0.800f)),          mat_diffuse(f3(0.200f, 0.380f, 0.800f))};
    b.primitives = {prim_quad(f3(0, 0, 0), f3(0, 0, S), f3(S, 0, 0), 0),
                    prim_quad(f3(0, S, 0), f3(S, 0, 0), f3(0, 0, S), 0),
                    prim_quad(f3(0, 0, S), f3(0, S, 0), f3(S, 0, 0), 0),
                    prim_quad(f3(S, 0, 0), f3(0, 0, S), f3(0, S, 0), 1),
                    prim_quad(f3(0, 0, 0), f3(0, S, 0), f3(0, 0, S), 2),
                    prim_quad(f3(213.0f, S - 0.1f, 227.0f), f3(130.0f, 0, 0), f3(0, 0, 105.0f), 3),
                    prim_sphere(f3(185.0f, 120.0f, 175.0f), 120.0f, 4),
                    prim_sphere(f3(370.0f, 90.0f, 350.0f), 90.0f, 5)};
    camera_look_at(b, f3(278.0f, 278.0f, -800.0f), f3(278.0f, 278.0f, 0.0f), 37.0f);
    b.background = f3(0, 0, 0);
}

static void cornell_mesh_ref(Blob& b) {
    cornell_box_ref(b);
    std::vector<GpuPrimitive> prims;
    prims.swap(b.primitives);
    for (const auto& p : prims)
        if (p.kind == LUCIDA_PRIM_KIND_QUAD)
            mesh_quad(f3(p.position[0], p.position[1], p.position[2]), f3(p.edge_u[0], p.edge_u[1], p.edge_u[2]),
                      f3(p.edge_v[0], p.edge_v[1], p.edge_v[2]))
                .append_to(b, p.material);
    const uint32_t sphere_mats[2] = {4, 5};
    int k = 0;
    for (const auto& p : prims) {
        if (p.kind == LUCIDA_PRIM_KIND_QUAD) continue;
        if (k >= 2) break;
        mesh_icosphere(f3(p.position[0], p.position[1], p.position[2]), p.radius, 4).append_to(b, sphere_mats[k]);
        ++k;
    }
}

// bvh_stress (scenes.rs:218): the Cornell box's quads with a 6 x 2 x 6 lattice
// of level-4 icospheres (radius 30) in two alternating materials.
static void bvh_stress_ref(Blob& b) {
    const float S = 555.0f;
    cornell_box_ref(b);
    std::vector<GpuPrimitive> quads;
    for (const auto& p : b.primitives)
        if (p.kind == LUCIDA_PRIM_KIND_QUAD) quads.push_back(p);
    b.primitives.swap(quads);
    const uint32_t nx = 6, ny = 2, nz = 6;
    auto f = [&](uint32_t i, uint32_t count) { return ((float)i + 1.0f) / ((float)count + 1.0f) * S; };
    for (uint32_t ix = 0; ix < nx; ++ix)
        for (uint32_t iy = 0; iy < ny; ++iy)
            for (uint32_t iz = 0; iz < nz; ++iz) {
                const f3 centre(f(ix, nx), f(iy, ny) * 0.62f + 95.0f, f(iz, nz));
                const uint32_t material = (ix + iy + iz) % 2 == 0 ? 4u : 5u;
                mesh_icosphere(centre, 30.0f, 4).append_to(b, material);
            }
}

// furnace_test (scenes.rs:275): five white metal spheres, roughness 0..1, in
// a uniform environment of radiance 1.
static void furnace_test_ref(Blob& b) {
    const int N = 5;
    for (int i = 0; i < N; ++i) {
        const float roughness = (float)i / (float)(N - 1);
        b.materials.push_back(mat_metal(f3::splat(1.0f), roughness));
        const float x = ((float)i - (float)(N - 1) * 0.5f) * 2.4f;
        b.primitives.push_back(prim_sphere(f3(x, 0.0f, 0.0f), 1.0f, (uint32_t)i));
    }
    camera_look_at(b, f3(0.0f, 0.8f, -12.0f), f3(0, 0, 0), 45.0f);
    b.background = f3::splat(1.0f);
}

// metal_sweep (scenes.rs:327): copper and gold rows (measured complex IOR,
// bsdf.rs conductors) of increasing roughness over a floor.
static void metal_sweep_ref(Blob& b) {
    const int N = 6;
    const f3 COPPER_ETA(0.200438f, 0.924033f, 1.102212f), COPPER_K(3.9124f, 2.44753f, 2.13764f);
    const f3 GOLD_ETA(0.143245f, 0.375346f, 1.442386f), GOLD_K(3.98312f, 2.385704f, 1.603575f);
    b.materials.push_back(mat_diffuse(f3::splat(0.35f)));
    for (int row = 0; row < 2; ++row) {
        const f3 eta = row == 0 ? COPPER_ETA : GOLD_ETA, k = row == 0 ? COPPER_K : GOLD_K;
        for (int i = 0; i < N; ++i) {
            const float t = (float)i / (float)(N - 1);
            const float roughness = 0.02f + 0.78f * t * t;
            b.materials.push_back(mat_conductor(eta, k, roughness));
            const float x = ((float)(N - 1) * 0.5f - (float)i) * 2.3f;
            b.primitives.push_back(
                prim_sphere(f3(x, 1.0f, (float)row * 2.6f), 1.0f, (uint32_t)b.materials.size() - 1));
        }
    }
    b.primitives.push_back(prim_quad(f3(-40.0f, 0.0f, -40.0f), f3(0.0f, 0.0f, 80.0f), f3(80.0f, 0.0f, 0.0f), 0));
    camera_look_at(b, f3(0.0f, 3.4f, -13.0f), f3(0.0f, 1.0f, 1.3f), 40.0f);
    b.background = f3::splat(0.75f);
}

// mis_scene (scenes.rs:410): Veach's plates of increasing roughness, each
// tilted to reflect the emitter row to the camera, and four emitters of equal
// power and increasing size.
static void mis_scene_ref(Blob& b) {
    b.materials.push_back(mat_diffuse(f3::splat(0.25f)));
    const f3 eye(0.0f, 3.0f, -8.5f);
    const float light_y = 5.4f, light_z = 1.2f, depth = 1.5f;
    const float rough[4] = {0.045f, 0.09f, 0.18f, 0.35f};
    for (int i = 0; i < 4; ++i) {
        b.materials.push_back(mat_metal(f3::splat(0.85f), rough[i]));
        const f3 centre(0.0f, 0.62f * (float)i, 0.6f + (float)i * 1.9f);
        const f3 to_camera = normalize(eye - centre);
        const f3 to_light = normalize(f3(0.0f, light_y, light_z) - centre);
        const f3 n = normalize(to_camera + to_light);
        const f3 edge_u(0.0f, -depth * n.z, depth * n.y);
        const f3 edge_v(10.0f, 0.0f, 0.0f);
        const f3 origin = centre - 0.5f * edge_u - 0.5f * edge_v;
        b.primitives.push_back(prim_quad(origin, edge_u, edge_v, (uint32_t)b.materials.size() - 1));
    }
    const float POWER = 46.0f;
    const float half[4] = {0.055f, 0.16f, 0.46f, 1.15f};
    for (int j = 0; j < 4; ++j) {
        const float hw = half[j];
        const float x = 3.7f - (float)j * 2.45f;
        const float area = 4.0f * hw * hw;
        b.materials.push_back(mat_emissive(f3::splat(POWER / area)));
        b.primitives.push_back(prim_quad(f3(x - hw, light_y, light_z - hw), f3(2.0f * hw, 0.0f, 0.0f),
                                         f3(0.0f, 0.0f, 2.0f * hw), (uint32_t)b.materials.size() - 1));
    }
    b.primitives.push_back(prim_quad(f3(-16.0f, -3.0f, 10.5f), f3(0.0f, 16.0f, 0.0f), f3(32.0f, 0.0f, 0.0f), 0));
    camera_look_at(b, eye, f3(0.0f, 1.35f, 3.4f), 42.0f);
    b.background = f3::splat(0.015f);
}

// glass_box (scenes.rs:536): the Cornell walls with a wider light, clear,
// diamond and frosted glass spheres and a tinted pane.
static void glass_box_ref(Blob& b) {
    const float S = 555.0f;
    b.materials = {mat_diffuse(f3(0.725f, 0.710f, 0.680f)), mat_diffuse(f3(0.630f, 0.065f, 0.050f)),
                   mat_diffuse(f3(0.140f, 0.450f, 0.091f)), mat_emissive(f3(30.0f, 26.0f, 18.0f)),
                   mat_glass(f3::splat(1.0f), 0.0f, 1.5f),  mat_glass(f3::splat(1.0f), 0.0f, 2.42f),
                   mat_glass(f3::splat(1.0f), 0.25f, 1.5f), mat_glass(f3(0.35f, 0.75f, 0.55f), 0.02f, 1.5f)};
    b.primitives = {prim_quad(f3(0, 0, 0), f3(0, 0, S), f3(S, 0, 0), 0),
                    prim_quad(f3(0, S, 0), f3(S, 0, 0), f3(0, 0, S), 0),
                    prim_quad(f3(0, 0, S), f3(0, S, 0), f3(S, 0, 0), 0),
                    prim_quad(f3(S, 0, 0), f3(0, 0, S), f3(0, S, 0), 1),
                    prim_quad(f3(0, 0, 0), f3(0, S, 0), f3(0, 0, S), 2),
                    prim_quad(f3(163.0f, S - 0.1f, 177.0f), f3(230.0f, 0, 0), f3(0, 0, 205.0f), 3),
                    prim_quad(f3(60.0f, 40.0f, 430.0f), f3(0, 300.0f, 0), f3(230.0f, 0, 0), 7),
                    prim_sphere(f3(150.0f, 110.0f, 160.0f), 110.0f, 4),
                    prim_sphere(f3(390.0f, 95.0f, 140.0f), 95.0f, 5),
                    prim_sphere(f3(300.0f, 85.0f, 330.0f), 85.0f, 6)};
    camera_look_at(b, f3(278.0f, 278.0f, -800.0f), f3(278.0f, 278.0f, 0.0f), 37.0f);
    b.background = f3(0, 0, 0);
}

// sunset (scenes.rs:651): a diffuse, a near-mirror and a glass sphere on a
// ground quad, lit only by a 256 x 128 procedural sky with a low sun.
static void sunset_ref(Blob& b) {
    b.materials = {mat_diffuse(f3(0.38f, 0.35f, 0.30f)), mat_diffuse(f3::splat(0.78f)),
                   mat_metal(f3(0.95f, 0.93f, 0.88f), 0.06f), mat_glass(f3::splat(1.0f), 0.0f, 1.5f)};
    b.primitives = {prim_quad(f3(-40.0f, 0.0f, -40.0f), f3(80.0f, 0.0f, 0.0f), f3(0.0f, 0.0f, 80.0f), 0),
                    prim_sphere(f3(-2.4f, 1.0f, 0.0f), 1.0f, 1), prim_sphere(f3(0.0f, 1.0f, 0.0f), 1.0f, 2),
                    prim_sphere(f3(2.4f, 1.0f, 0.0f), 1.0f, 3)};
    procedural_sky(b, 256, 128, f3(-0.45f, 0.22f, -0.86f), 6000.0f, 1.0f);
    camera_look_at(b, f3(0.0f, 1.6f, 7.5f), f3(0.0f, 0.9f, 0.0f), 42.0f);
    b.background = f3(0, 0, 0);
}

// instance_forest (scenes.rs:733): 49 placements of one level-2 icosphere on
// a 7 x 7 jittered grid, rotated about y and scaled non-uniformly (PCG seed
// 0x1057), every third one gold, under a procedural sky. The placement matrix
// is translation * rotation_y * scale, whose entries are single products.
static void instance_forest_ref(Blob& b) {
    b.materials = {mat_diffuse(f3(0.34f, 0.32f, 0.29f)), mat_glossy(f3(0.70f, 0.35f, 0.25f), 0.35f),
                   mat_metal(f3(0.95f, 0.80f, 0.40f), 0.18f)};
    b.primitives = {prim_quad(f3(-30.0f, 0.0f, -30.0f), f3(60.0f, 0.0f, 0.0f), f3(0.0f, 0.0f, 60.0f), 0)};
    mesh_icosphere(f3(0, 0, 0), 0.5f, 2).append_to(b, 1);
    b.meshes.push_back(lucida_mesh_range{0, (uint32_t)b.triangles.size()});
    Pcg32 rng(0x1057u, 0u, 0u);
    const float TAU_F = 6.28318530717958647692f;
    for (int gx = 0; gx < 7; ++gx)
        for (int gz = 0; gz < 7; ++gz) {
            const float x = (float)gx * 1.6f - 4.8f + (rng.next_f32() - 0.5f) * 0.5f;
            const float z = (float)gz * 1.6f - 4.8f + (rng.next_f32() - 0.5f) * 0.5f;
            const float sx = 0.7f + rng.next_f32() * 0.7f;
            const float sy = 0.5f + rng.next_f32() * 1.1f;
            const float sz = 0.7f + rng.next_f32() * 0.7f;
            const float spin = rng.next_f32() * TAU_F;
            const float y = 0.5f * sy;
            const float sn = std::sin(spin), cs = std::cos(spin);
            lucida_instance_placement pl{};
            const float m[12] = {cs * sx, 0.0f, sn * sz, x, 0.0f, sy, 0.0f, y, -sn * sx, 0.0f, cs * sz, z};
            std::memcpy(pl.object_to_world, m, sizeof m);
            pl.mesh = 0;
            pl.material = (gx + gz) % 3 == 0 ? 2u : 0xFFFFFFFFu;
            b.placements.push_back(pl);
        }
    procedural_sky(b, 256, 128, f3(-0.4f, 0.35f, -0.85f), 6000.0f, 1.0f);
    camera_look_at(b, f3(0.0f, 3.2f, -11.0f), f3(0.0f, 0.7f, 0.0f), 40.0f);
    b.background = f3(0, 0, 0);
}

// --- benchmark scene 1: the 1x1x1 triangle-quad box ---------------------------
//
// White 0.73 floor/ceiling/back, red (0.63,0.065,0.05) left, green
// (0.14,0.45,0.09) right, all as two-triangle quads with the front open; two
// analytic spheres (r = 0.15 and 0.2, diffuse 0.73) and a 0.25 x 0.25 analytic
// ceiling light of radiance 15 facing down. The camera looks into the open
// front with the reference Cornell camera's proportions (eye at 800/555 of the
// box size, vfov 37).
static void box_c1(Blob& b) {
    b.materials = {mat_diffuse(f3::splat(0.73f)), mat_diffuse(f3(0.63f, 0.065f, 0.05f)),
                   mat_diffuse(f3(0.14f, 0.45f, 0.09f)), mat_emissive(f3::splat(15.0f)),
                   mat_diffuse(f3::splat(0.73f))};
    const float S = 1.0f;
    mesh_quad(f3(0, 0, 0), f3(0, 0, S), f3(S, 0, 0)).append_to(b, 0);  // floor, +y
    mesh_quad(f3(0, S, 0), f3(S, 0, 0), f3(0, 0, S)).append_to(b, 0);  // ceiling, -y
    mesh_quad(f3(0, 0, S), f3(0, S, 0), f3(S, 0, 0)).append_to(b, 0);  // back, -z
    mesh_quad(f3(S, 0, 0), f3(0, 0, S), f3(0, S, 0)).append_to(b, 1);  // x = 1 wall, red
    mesh_quad(f3(0, 0, 0), f3(0, S, 0), f3(0, 0, S)).append_to(b, 2);  // x = 0 wall, green
    b.primitives = {prim_quad(f3(0.375f, S - 2.0e-4f, 0.375f), f3(0.25f, 0, 0), f3(0, 0, 0.25f), 3),
                    prim_sphere(f3(0.33f, 0.15f, 0.32f), 0.15f, 4), prim_sphere(f3(0.66f, 0.20f, 0.63f), 0.20f, 4)};
    camera_look_at(b, f3(0.5f, 0.5f, -800.0f / 555.0f), f3(0.5f, 0.5f, 0.0f), 37.0f);
    b.background = f3(0, 0, 0);
}

// --- benchmark scene 2: displaced geodesic sphere in a room --------------------

static inline uint32_t pcg_hash(uint32_t x) {
    uint32_t state = x * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28) + 4u)) ^ state) * 277803737u;
    return (word >> 22) ^ word;
}

// Value noise on the integer lattice: lattice values from a sequential PCG hash
// of (seed, ix, iy, iz), trilinear interpolation with a smoothstep fade.
static double value_noise(double x, double y, double z, uint32_t seed) {
    double fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
    int ix = (int)fx, iy = (int)fy, iz = (int)fz;
    double tx = x - fx, ty = y - fy, tz = z - fz;
    auto fade = [](double t) { return t * t * (3.0 - 2.0 * t); };
    tx = fade(tx);
    ty = fade(ty);
    tz = fade(tz);
    auto lat = [&](int a, int b2, int c) {
        uint32_t h = pcg_hash(pcg_hash(pcg_hash(pcg_hash(seed) + (uint32_t)a) + (uint32_t)b2) + (uint32_t)c);
        return (double)h * (2.0 / 4294967296.0) - 1.0;
    };
    double v000 = lat(ix, iy, iz), v100 = lat(ix + 1, iy, iz), v010 = lat(ix, iy + 1, iz), v110 = lat(ix + 1, iy + 1, iz);
    double v001 = lat(ix, iy, iz + 1), v101 = lat(ix + 1, iy, iz + 1), v011 = lat(ix, iy + 1, iz + 1),
           v111 = lat(ix + 1, iy + 1, iz + 1);
    double x00 = v000 + (v100 - v000) * tx, x10 = v010 + (v110 - v010) * tx;
    double x01 = v001 + (v101 - v001) * tx, x11 = v011 + (v111 - v011) * tx;
    double y0 = x00 + (x10 - x00) * ty, y1 = x01 + (x11 - x01) * ty;
    return y0 + (y1 - y0) * tz;
}

// Three octaves (lattice frequency 2, 4, 8 on the unit sphere, amplitude 1,
// 1/2, 1/4), normalised to [-1, 1].
static double noise3(double x, double y, double z, uint32_t seed) {
    double sum = 0.0, amp = 1.0, freq = 2.0, norm = 0.0;
    for (int o = 0; o < 3; ++o) {
        sum += amp * value_noise(x * freq, y * freq, z * freq, seed + (uint32_t)o * 0x9e3779b9u);
        norm += amp;
        amp *= 0.5;
        freq *= 2.0;
    }
    return sum / norm;
}

static inline float rng_f32(uint32_t& state) {
    state = state * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28) + 4u)) ^ state) * 277803737u;
    uint32_t u = (word >> 22) ^ word;
    return (float)(u >> 8) * (1.0f / 16777216.0f);
}

// Frequency-f geodesic icosphere (20 f^2 triangles, 10 f^2 + 2 vertices):
// every icosahedron face is cut into an f x f barycentric grid and the grid
// points are projected onto the sphere. Vertices on shared edges and corners
// are created once. `group_of` receives a material group per triangle: the
// base face and which quarter of the grid rows it sits in (80 groups).
static void geodesic(uint32_t f, std::vector<f3>& unit, std::vector<uint32_t>& tris, std::vector<uint32_t>& group_of) {
    const float t = (1.0f + std::sqrt(5.0f)) / 2.0f;
    const f3 base[12] = {f3(-1, t, 0), f3(1, t, 0),  f3(-1, -t, 0), f3(1, -t, 0), f3(0, -1, t), f3(0, 1, t),
                         f3(0, -1, -t), f3(0, 1, -t), f3(t, 0, -1),  f3(t, 0, 1),  f3(-t, 0, -1), f3(-t, 0, 1)};
    static const uint32_t faces0[20][3] = {{0, 11, 5}, {0, 5, 1},  {0, 1, 7},   {0, 7, 10}, {0, 10, 11},
                                           {1, 5, 9},  {5, 11, 4}, {11, 10, 2}, {10, 7, 6}, {7, 1, 8},
                                           {3, 9, 4},  {3, 4, 2},  {3, 2, 6},   {3, 6, 8},  {3, 8, 9},
                                           {4, 9, 5},  {2, 4, 11}, {6, 2, 10},  {8, 6, 7},  {9, 8, 1}};
    double B[12][3];
    for (int i = 0; i < 12; ++i) {
        double l = std::sqrt((double)base[i].x * base[i].x + (double)base[i].y * base[i].y + (double)base[i].z * base[i].z);
        B[i][0] = base[i].x / l;
        B[i][1] = base[i].y / l;
        B[i][2] = base[i].z / l;
    }
    unit.clear();
    tris.clear();
    group_of.clear();
    auto make = [&](double wa, int a, double wb, int b2, double wc, int c) {
        double p[3];
        for (int k = 0; k < 3; ++k) p[k] = wa * B[a][k] + wb * B[b2][k] + wc * B[c][k];
        double l = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
        unit.push_back(f3((float)(p[0] / l), (float)(p[1] / l), (float)(p[2] / l)));
        return (uint32_t)unit.size() - 1;
    };
    std::vector<uint32_t> corner(12, 0xFFFFFFFFu);
    std::unordered_map<uint64_t, uint32_t> edge_vertex;
    std::vector<uint32_t> grid((size_t)(f + 1) * (f + 1));
    for (int fi = 0; fi < 20; ++fi) {
        int a = (int)faces0[fi][0], b2 = (int)faces0[fi][1], c = (int)faces0[fi][2];
        for (uint32_t i = 0; i <= f; ++i) {
            for (uint32_t j = 0; i + j <= f; ++j) {
                uint32_t k = f - i - j;  // weights: a gets k, b gets i, c gets j
                uint32_t id;
                int nz = (k > 0) + (i > 0) + (j > 0);
                if (nz == 1) {
                    int v = k > 0 ? a : (i > 0 ? b2 : c);
                    if (corner[v] == 0xFFFFFFFFu) corner[v] = make(v == a, a, v == b2, b2, v == c, c);
                    id = corner[v];
                } else if (nz == 2) {
                    int u0, u1;
                    uint32_t w1;
                    if (k == 0) {
                        u0 = b2; u1 = c; w1 = j;
                    } else if (i == 0) {
                        u0 = a; u1 = c; w1 = j;
                    } else {
                        u0 = a; u1 = b2; w1 = i;
                    }
                    // Canonical key: smaller corner id first, weight of the larger one.
                    uint32_t wl = u0 < u1 ? w1 : f - w1;
                    int lo = u0 < u1 ? u0 : u1, hi = u0 < u1 ? u1 : u0;
                    uint64_t key = ((uint64_t)lo * 12 + hi) * (f + 1) + wl;
                    auto it = edge_vertex.find(key);
                    if (it == edge_vertex.end()) {
                        id = make((double)(f - wl) / f, lo, (double)wl / f, hi, 0.0, hi);
                        edge_vertex.emplace(key, id);
                    } else {
                        id = it->second;
                    }
                } else {
                    id = make((double)k / f, a, (double)i / f, b2, (double)j / f, c);
                }
                grid[(size_t)i * (f + 1) + j] = id;
            }
        }
        auto G = [&](uint32_t i, uint32_t j) { return grid[(size_t)i * (f + 1) + j]; };
        for (uint32_t i = 0; i < f; ++i) {
            uint32_t band = (uint32_t)fi * 4 + (i * 4) / f;
            for (uint32_t j = 0; i + j < f; ++j) {
                tris.insert(tris.end(), {G(i, j), G(i + 1, j), G(i, j + 1)});
                group_of.push_back(band);
                if (i + j + 1 < f) {
                    tris.insert(tris.end(), {G(i + 1, j), G(i + 1, j + 1), G(i, j + 1)});
                    group_of.push_back(band);
                }
            }
        }
    }
}

// A face of the room: an nu x nv grid of two-triangle quads spanning eu x ev
// from `origin`, facing along cross(eu, ev).
static void room_face(Blob& b, f3 origin, f3 eu, f3 ev, uint32_t nu, uint32_t nv, uint32_t material) {
    for (uint32_t a = 0; a < nu; ++a)
        for (uint32_t c = 0; c < nv; ++c) {
            f3 o = origin + eu * ((float)a / (float)nu) + ev * ((float)c / (float)nv);
            mesh_quad(o, eu / (float)nu, ev / (float)nv).append_to(b, material);
        }
}

// Unit geodesic sphere of frequency f displaced radially by amplitude *
// noise3 (seeded by `seed`), with area-weighted smooth normals. Triangle
// materials: `material` + the triangle's group (80 groups) when `by_group`,
// else `material`.
static void append_displaced_sphere(Blob& b, uint32_t f, float amplitude, uint32_t seed, uint32_t material,
                                    bool by_group) {
    std::vector<f3> unit;
    std::vector<uint32_t> tris, groups;
    geodesic(f, unit, tris, groups);
    std::vector<f3> pos(unit.size());
    for (size_t i = 0; i < unit.size(); ++i) {
        double n = noise3(unit[i].x, unit[i].y, unit[i].z, seed * 2u + 2u);
        float r = (float)(1.0 + (double)amplitude * n);
        pos[i] = unit[i] * r;
    }
    // Smooth normals: area-weighted sum of the incident face normals.
    std::vector<double> acc(unit.size() * 3, 0.0);
    for (size_t t = 0; t < tris.size(); t += 3) {
        f3 p0 = pos[tris[t]], p1 = pos[tris[t + 1]], p2 = pos[tris[t + 2]];
        f3 c = cross(p1 - p0, p2 - p0);
        for (int k = 0; k < 3; ++k) {
            acc[3 * tris[t + k] + 0] += c.x;
            acc[3 * tris[t + k] + 1] += c.y;
            acc[3 * tris[t + k] + 2] += c.z;
        }
    }
    uint32_t base = (uint32_t)(b.positions.size() / 4);
    for (size_t i = 0; i < unit.size(); ++i) {
        b.positions.insert(b.positions.end(), {pos[i].x, pos[i].y, pos[i].z, 0.0f});
        double nx = acc[3 * i], ny = acc[3 * i + 1], nz = acc[3 * i + 2];
        double l = std::sqrt(nx * nx + ny * ny + nz * nz);
        GpuVertexAttr a{};
        if (l > 0) {
            a.normal[0] = (float)(nx / l);
            a.normal[1] = (float)(ny / l);
            a.normal[2] = (float)(nz / l);
        } else {
            unit[i].store(a.normal);
        }
        b.vertex_attrs.push_back(a);
    }
    for (size_t t = 0, g = 0; t < tris.size(); t += 3, ++g)
        b.triangles.push_back(GpuTriangle{base + tris[t], base + tris[t + 1], base + tris[t + 2], by_group ? material + groups[g] : material});

}

// Room [-2.5, 2.5] x [-1.5, 2.5] x [-2.5, 2.5], closed, inward-facing:
// floor and ceiling 12 x 12 quads, four walls 8 x 7 quads (512 quads, 1,024
// triangles). Unit sphere at the origin, geodesic frequency f (229 gives
// 1,048,820 triangles), displaced radially by amplitude * noise. Materials:
// GGX over a diffuse base, roughness U[0.05, 1], albedo U[0.1, 0.9]^3, one per
// room face and one per sphere group (80), drawn from a PCG stream seeded by
// `seed`. Sixteen 0.25 x 0.25 analytic quads under the ceiling emit 20.
static void sphere_c2(Blob& b, uint32_t f, float amplitude, uint32_t seed) {
    // materials: 0..5 room faces, 6..85 sphere groups, 86 emitter
    uint32_t st = pcg_hash(seed * 0x9e3779b9u + 3u);
    for (int m = 0; m < 86; ++m) {
        float r = 0.05f + 0.95f * rng_f32(st);
        float cr = 0.1f + 0.8f * rng_f32(st);
        float cg = 0.1f + 0.8f * rng_f32(st);
        float cb = 0.1f + 0.8f * rng_f32(st);
        b.materials.push_back(mat_glossy(f3(cr, cg, cb), r));
    }
    b.materials.push_back(mat_emissive(f3(20.0f, 18.0f, 15.0f)));
    const uint32_t LIGHT = 86;

    const float X0 = -2.5f, X1 = 2.5f, Y0 = -1.5f, Y1 = 2.5f, Z0 = -2.5f, Z1 = 2.5f;
    room_face(b, f3(X0, Y0, Z0), f3(0, 0, Z1 - Z0), f3(X1 - X0, 0, 0), 12, 12, 0);  // floor, +y
    room_face(b, f3(X0, Y1, Z0), f3(X1 - X0, 0, 0), f3(0, 0, Z1 - Z0), 12, 12, 1);  // ceiling, -y
    room_face(b, f3(X0, Y0, Z1), f3(0, Y1 - Y0, 0), f3(X1 - X0, 0, 0), 8, 7, 2);    // z = Z1, -z
    room_face(b, f3(X0, Y0, Z0), f3(X1 - X0, 0, 0), f3(0, Y1 - Y0, 0), 8, 7, 3);    // z = Z0, +z
    room_face(b, f3(X1, Y0, Z0), f3(0, 0, Z1 - Z0), f3(0, Y1 - Y0, 0), 8, 7, 4);    // x = X1, -x
    room_face(b, f3(X0, Y0, Z0), f3(0, Y1 - Y0, 0), f3(0, 0, Z1 - Z0), 8, 7, 5);    // x = X0, +x

    append_displaced_sphere(b, f, amplitude, seed, 6u, true);

    for (int gx = 0; gx < 4; ++gx)
        for (int gz = 0; gz < 4; ++gz) {
            float cx = -1.5f + 1.0f * gx, cz = -1.5f + 1.0f * gz;
            b.primitives.push_back(
                prim_quad(f3(cx - 0.125f, Y1 - 1.0e-3f, cz - 0.125f), f3(0.25f, 0, 0), f3(0, 0, 0.25f), LIGHT));
        }
    camera_look_at(b, f3(0.0f, 0.45f, 2.3f), f3(0.0f, 0.0f, 0.0f), 60.0f);
    b.background = f3(0, 0, 0);
}

// Instanced field (the config-3 scene at count 4096, f 22, extent 100, env
// 2048 x 1024): one noise-displaced geodesic sphere mesh (20 f^2 triangles,
// amplitude 0.05, noise seed 4) placed `count` times with a uniformly random
// rotation (unit quaternion from three uniforms, Shoemake 1992), uniform scale
// U[0.5, 2] and position U[0, extent]^3. Each placement overrides the
// material: 30% rough dielectric (IOR 1.5, roughness U[0, 0.3]), the rest GGX
// (roughness U[0.05, 1], albedo U[0.1, 0.9]^3). Lit only by an environment
// map: log-normal radiance (mu 0, sigma 1.5) per texel plus three Gaussian
// hot spots of peak 1e4 (seed 5).
static void instance_field(Blob& b, uint32_t f, uint32_t count, float extent, uint32_t env_w, uint32_t env_h,
                           uint32_t seed) {
    b.materials.push_back(mat_glossy(f3(0.5f, 0.5f, 0.5f), 0.5f));  // the mesh's own (always overridden)
    const uint32_t first = (uint32_t)b.triangles.size();
    append_displaced_sphere(b, f, 0.05f, 4u, 0u, false);
    b.meshes.push_back(lucida_mesh_range{first, (uint32_t)b.triangles.size() - first});
    uint32_t st = pcg_hash(seed * 0x9e3779b9u + 7u);
    const float TWO_PI = 6.2831853071795864f;
    for (uint32_t i = 0; i < count; ++i) {
        const float u1 = rng_f32(st), u2 = rng_f32(st), u3 = rng_f32(st);
        const float a = std::sqrt(1.0f - u1), c = std::sqrt(u1);
        const float qx = a * std::sin(TWO_PI * u2), qy = a * std::cos(TWO_PI * u2);
        const float qz = c * std::sin(TWO_PI * u3), qw = c * std::cos(TWO_PI * u3);
        const float scale = 0.5f + 1.5f * rng_f32(st);
        const f3 pos(extent * rng_f32(st), extent * rng_f32(st), extent * rng_f32(st));
        const float R[3][3] = {{1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qz * qw), 2 * (qx * qz + qy * qw)},
                               {2 * (qx * qy + qz * qw), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qx * qw)},
                               {2 * (qx * qz - qy * qw), 2 * (qy * qz + qx * qw), 1 - 2 * (qx * qx + qy * qy)}};
        lucida_instance_placement pl{};
        for (int r = 0; r < 3; ++r) {
            for (int k = 0; k < 3; ++k) pl.object_to_world[4 * r + k] = scale * R[r][k];
        }
        pl.object_to_world[3] = pos.x;
        pl.object_to_world[7] = pos.y;
        pl.object_to_world[11] = pos.z;
        pl.mesh = 0;
        pl.material = (uint32_t)b.materials.size();
        if (rng_f32(st) < 0.3f) {
            GpuMaterial m{};
            f3(1.0f, 1.0f, 1.0f).store(m.base_color);
            m.roughness = 0.3f * rng_f32(st);
            m.ior = 1.5f;
            m.transmission = 1.0f;
            b.materials.push_back(m);
        } else {
            const float r = 0.05f + 0.95f * rng_f32(st);
            const float cr = 0.1f + 0.8f * rng_f32(st), cg = 0.1f + 0.8f * rng_f32(st), cb = 0.1f + 0.8f * rng_f32(st);
            b.materials.push_back(mat_glossy(f3(cr, cg, cb), r));
        }
        b.placements.push_back(pl);
    }
    // environment
    std::vector<float> rgba((size_t)env_w * env_h * 4);
    uint32_t es = pcg_hash((5u + seed) * 0x9e3779b9u + 11u);  // seed 0: the config-3 sky (seed 5)
    for (size_t i = 0; i < (size_t)env_w * env_h; ++i) {
        const float z = std::sqrt(-2.0f * std::log(std::max(rng_f32(es), 1e-12f))) * std::cos(TWO_PI * rng_f32(es));
        const float L = std::exp(1.5f * z);
        rgba[4 * i] = rgba[4 * i + 1] = rgba[4 * i + 2] = L;
        rgba[4 * i + 3] = 1.0f;
    }
    for (int k = 0; k < 3; ++k) {
        const float cu = rng_f32(es) * env_w, cv = (0.15f + 0.5f * rng_f32(es)) * env_h;
        const float sigma = 0.004f * env_w;
        const int rad = (int)std::ceil(4.0f * sigma);
        for (int dv = -rad; dv <= rad; ++dv)
            for (int du = -rad; du <= rad; ++du) {
                const int x = ((int)cu + du + (int)env_w) % (int)env_w, y = (int)cv + dv;
                if (y < 0 || y >= (int)env_h) continue;
                const float g = 1.0e4f * std::exp(-(float)(du * du + dv * dv) / (2.0f * sigma * sigma));
                for (int ch = 0; ch < 3; ++ch) rgba[4 * ((size_t)y * env_w + x) + ch] += g;
            }
    }
    env_build(b, env_w, env_h, rgba.data());
    camera_look_at(b, f3(0.5f * extent, 0.62f * extent, -0.85f * extent), f3(0.5f * extent, 0.45f * extent, 0.5f * extent),
                   50.0f);
    b.background = f3(0, 0, 0);
}

}  // namespace lucida

using namespace lucida;

extern "C" {

lucida_status lucida_blob_create(const char* name, const lucida_blob_params* params, lucida_blob** out) {
    if (!name || !out) return set_error(LUCIDA_INVALID_ARG, "lucida_blob_create: null argument");
    lucida_blob_params p{};
    if (params) p = *params;
    auto* b = new Blob();
    std::string n(name);
    if (n == "cornell-box") {
        cornell_box_ref(*b);
    } else if (n == "cornell-mesh") {
        cornell_mesh_ref(*b);
    } else if (n == "bvh-stress") {
        bvh_stress_ref(*b);
    } else if (n == "furnace-test") {
        furnace_test_ref(*b);
    } else if (n == "metal-sweep") {
        metal_sweep_ref(*b);
    } else if (n == "mis-scene") {
        mis_scene_ref(*b);
    } else if (n == "glass-box") {
        glass_box_ref(*b);
    } else if (n == "sunset") {
        sunset_ref(*b);
    } else if (n == "instance-forest") {
        instance_forest_ref(*b);
    } else if (n == "box-c1") {
        box_c1(*b);
    } else if (n == "sphere-c2") {
        uint32_t f = p.frequency ? p.frequency : 229u;
        float amp = p.displacement != 0.0f ? p.displacement : 0.05f;
        sphere_c2(*b, f, amp, p.seed);
    } else if (n == "instances-c3") {
        instance_field(*b, p.frequency ? p.frequency : 22u, 4096u, 100.0f, 2048u, 1024u, p.seed);
    } else if (n == "instances-small") {
        instance_field(*b, p.frequency ? p.frequency : 4u, 64u, 10.0f, 256u, 128u, p.seed);
    } else if (n.rfind("icosphere-", 0) == 0) {
        int lv = std::atoi(n.c_str() + 10);
        if (lv < 0 || lv > 9) {
            delete b;
            return set_error(LUCIDA_INVALID_ARG, "icosphere level out of range 0..9");
        }
        b->materials = {mat_diffuse(f3::splat(0.8f))};
        mesh_icosphere(f3(0, 0, 0), 1.0f, (uint32_t)lv).append_to(*b, 0);
        camera_look_at(*b, f3(0, 0, 4), f3(0, 0, 0), 40.0f);
    } else {
        delete b;
        return set_error(LUCIDA_INVALID_ARG, ("unknown scene '" + n + "'").c_str());
    }
    build_lights(*b);
    *out = reinterpret_cast<lucida_blob*>(b);
    return LUCIDA_OK;
}

void lucida_blob_destroy(lucida_blob* blob) { delete reinterpret_cast<Blob*>(blob); }

lucida_status lucida_blob_desc(const lucida_blob* blob, lucida_scene_desc* d) {
    if (!blob || !d) return set_error(LUCIDA_INVALID_ARG, "lucida_blob_desc: null argument");
    const Blob& b = *reinterpret_cast<const Blob*>(blob);
    std::memset(d, 0, sizeof *d);
    d->materials = b.materials.data();
    d->num_materials = b.materials.size();
    d->primitives = b.primitives.data();
    d->num_primitives = b.primitives.size();
    d->instances = b.instances.data();
    d->num_instances = b.instances.size();
    d->positions = b.positions.data();
    d->num_positions = b.positions.size() / 4;
    d->vertex_attrs = b.vertex_attrs.data();
    d->triangles = b.triangles.data();
    d->num_triangles = b.triangles.size();
    d->bvh_nodes = b.bvh_nodes.data();
    d->num_bvh_nodes = b.bvh_nodes.size();
    d->lights = b.lights.data();
    d->num_lights = b.lights.size();
    d->tlas_root = b.tlas_root;
    d->env_width = b.env_width;
    d->env_height = b.env_height;
    d->env_pixels = b.env_pixels.data();
    d->env_conditional = b.env_conditional.data();
    d->env_marginal = b.env_marginal.data();
    d->env_total_weight = b.env_total_weight;
    d->build_lbvh = 0;
    return LUCIDA_OK;
}

lucida_status lucida_blob_uniforms(const lucida_blob* blob, uint32_t width, uint32_t height, GpuUniforms* u) {
    if (!blob || !u || width == 0 || height == 0)
        return set_error(LUCIDA_INVALID_ARG, "lucida_blob_uniforms: bad argument");
    const Blob& b = *reinterpret_cast<const Blob*>(blob);
    std::memset(u, 0, sizeof *u);
    u->width = width;
    u->height = height;
    u->num_primitives = (uint32_t)b.primitives.size();
    u->num_triangles = (uint32_t)b.triangles.size();
    u->num_bvh_nodes = (uint32_t)b.bvh_nodes.size();
    u->num_lights = (uint32_t)b.lights.size();
    b.background.store(u->background);
    u->env_width = b.env_width;
    u->env_height = b.env_height;
    u->env_total_weight = b.env_total_weight;
    u->num_instances = (uint32_t)b.instances.size();
    u->tlas_root = b.tlas_root;
    u->max_depth = 8;
    u->frame_seed = 0x5eed1234u;
    u->sampling_mode = LUCIDA_SAMPLING_MIS;
    u->sampler_kind = LUCIDA_SAMPLER_SOBOL;
    camera_write_uniforms(b, *u, (float)width / (float)height);
    return LUCIDA_OK;
}

lucida_status lucida_blob_set_bvh(lucida_blob* blob, const GpuBvhNode* nodes, size_t num_nodes,
                                  const GpuTriangle* tris) {
    if (!blob || (num_nodes && !nodes)) return set_error(LUCIDA_INVALID_ARG, "lucida_blob_set_bvh: bad argument");
    Blob& b = *reinterpret_cast<Blob*>(blob);
    b.bvh_nodes.assign(nodes, nodes + num_nodes);
    if (tris) b.triangles.assign(tris, tris + b.triangles.size());
    return LUCIDA_OK;
}

lucida_status lucida_blob_instancing(const lucida_blob* blob, const lucida_mesh_range** meshes, size_t* num_meshes,
                                     const lucida_instance_placement** placements, size_t* num_placements) {
    if (!blob || !meshes || !num_meshes || !placements || !num_placements)
        return set_error(LUCIDA_INVALID_ARG, "lucida_blob_instancing: null argument");
    const Blob& b = *reinterpret_cast<const Blob*>(blob);
    *meshes = b.meshes.data();
    *num_meshes = b.meshes.size();
    *placements = b.placements.data();
    *num_placements = b.placements.size();
    return LUCIDA_OK;
}

lucida_status lucida_camera_write_uniforms(const lucida_camera* cam, float aspect, GpuUniforms* u) {
    if (!cam || !u || !(aspect > 0.0f)) return set_error(LUCIDA_INVALID_ARG, "lucida_camera_write_uniforms: bad argument");
    write_camera(f3(cam->eye[0], cam->eye[1], cam->eye[2]), f3(cam->look_at[0], cam->look_at[1], cam->look_at[2]),
                 f3(cam->up[0], cam->up[1], cam->up[2]), cam->vfov_deg, cam->aperture, cam->focus_distance, *u, aspect);
    return LUCIDA_OK;
}

lucida_status lucida_blob_camera(const lucida_blob* blob, lucida_camera* out) {
    if (!blob || !out) return set_error(LUCIDA_INVALID_ARG, "lucida_blob_camera: null argument");
    const Blob& b = *reinterpret_cast<const Blob*>(blob);
    b.cam_eye.store(out->eye);
    b.cam_look_at.store(out->look_at);
    b.cam_up.store(out->up);
    out->vfov_deg = b.cam_vfov_deg;
    out->aperture = b.cam_aperture;
    out->focus_distance = b.cam_focus_distance;
    return LUCIDA_OK;
}

lucida_status lucida_blob_set_env(lucida_blob* blob, uint32_t width, uint32_t height, const float* rgba) {
    if (!blob || !rgba || width == 0 || height == 0)
        return set_error(LUCIDA_INVALID_ARG, "lucida_blob_set_env: bad argument");
    env_build(*reinterpret_cast<Blob*>(blob), width, height, rgba);
    return LUCIDA_OK;
}

}  // extern "C"
