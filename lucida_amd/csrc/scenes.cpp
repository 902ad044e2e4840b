// Host-side scene generation: the inputs of the hot path, laid out per the
// reference's crates/core/src/gpu_layout.rs.
//
// Restated from the reference: GpuMaterial constructors (gpu_layout.rs:74-160),
// GpuPrimitive::quad/sphere (gpu_layout.rs:214-240), MeshBuilder::append_to,
// quad and icosphere (mesh.rs:20-249), build_lights (light.rs:62-120),
// Camera::look_at / basis / write_uniforms (camera.rs:28-80), EnvMap::new
// (envmap.rs:112-180), cornell_box / cornell_mesh (scenes.rs:56-214). The two
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

static void camera_write_uniforms(const Blob& b, GpuUniforms& u, float aspect) {
    f3 forward = normalize(b.cam_look_at - b.cam_eye);
    f3 right = normalize(cross(forward, b.cam_up));
    f3 up = cross(right, forward);
    f3 down = -up;
    const float DEG = 3.14159265358979323846f / 180.0f;
    float half_height = std::tan(b.cam_vfov_deg * DEG * 0.5f) * b.cam_focus_distance;
    float half_width = half_height * aspect;
    f3 horizontal = right * (2.0f * half_width);
    f3 vertical = down * (2.0f * half_height);
    f3 upper_left = b.cam_eye + forward * b.cam_focus_distance - right * half_width - down * half_height;
    b.cam_eye.store(u.cam_origin);
    upper_left.store(u.cam_upper_left);
    horizontal.store(u.cam_horizontal);
    vertical.store(u.cam_vertical);
    u.lens_radius = b.cam_aperture * 0.5f;
    u.focus_distance = b.cam_focus_distance;
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

// --- reference scenes (scenes.rs:56-214) ---------------------------------------

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
    uint32_t es = pcg_hash(5u * 0x9e3779b9u + 11u);
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

lucida_status lucida_blob_set_env(lucida_blob* blob, uint32_t width, uint32_t height, const float* rgba) {
    if (!blob || !rgba || width == 0 || height == 0)
        return set_error(LUCIDA_INVALID_ARG, "lucida_blob_set_env: bad argument");
    env_build(*reinterpret_cast<Blob*>(blob), width, height, rgba);
    return LUCIDA_OK;
}

}  // extern "C"
