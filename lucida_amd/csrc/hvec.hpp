// Host-side float3 for scene generation. Operations follow glam's scalar
// semantics (normalize = v * (1 / length), per-component scalar division) so
// generated geometry matches the reference's generators bit for bit where a
// test compares them.
#pragma once
#include <cmath>
#include <cstdint>

namespace lucida {

struct f3 {
    float x = 0, y = 0, z = 0;
    f3() = default;
    f3(float a, float b, float c) : x(a), y(b), z(c) {}
    static f3 splat(float v) { return f3(v, v, v); }
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    void store(float* p) const { p[0] = x; p[1] = y; p[2] = z; }
};
inline f3 operator+(f3 a, f3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline f3 operator-(f3 a, f3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline f3 operator-(f3 a) { return f3(-a.x, -a.y, -a.z); }
inline f3 operator*(f3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
inline f3 operator*(float s, f3 a) { return f3(s * a.x, s * a.y, s * a.z); }
inline f3 operator/(f3 a, float s) { return f3(a.x / s, a.y / s, a.z / s); }
inline float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline f3 cross(f3 a, f3 b) { return f3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y); }
inline float length(f3 a) { return std::sqrt(dot(a, a)); }
inline f3 normalize(f3 a) { return a * (1.0f / length(a)); }
inline f3 normalize_or(f3 a, f3 fb) {
    float r = 1.0f / length(a);
    return (std::isfinite(r) && r > 0.0f) ? a * r : fb;
}

}  // namespace lucida
