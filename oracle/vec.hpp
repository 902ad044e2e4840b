// Oracle vector type: a scalar restatement of the glam 0.33.7 Vec2/Vec3
// operations the reference's CPU twin uses (Cargo.lock:202-204). glam's Vec3 is
// scalar on x86_64: dot is x*x' + y*y' + z*z' evaluated left to right,
// normalize is v * (1 / length), divisions by a scalar are per component.
// Compiled with -ffp-contract=off so no product is fused into a sum, as in Rust.
//
// TEST INFRASTRUCTURE: part of the CPU oracle. Only tests, the smoke check and
// the bench's CPU-baseline leg use it; the product never links it.
#pragma once
#include <cmath>
#include <cstdint>

namespace oracle {

constexpr float PI = 3.14159265358979323846f;
constexpr float TAU = 6.28318530717958647692f;
constexpr float INV_PI = 0.318309886183790671538f;

struct Vec2 {
    float x, y;
};

struct Vec3 {
    float x, y, z;
    Vec3() : x(0), y(0), z(0) {}
    Vec3(float a, float b, float c) : x(a), y(b), z(c) {}
    static Vec3 splat(float v) { return Vec3(v, v, v); }
    static Vec3 from(const float* p) { return Vec3(p[0], p[1], p[2]); }
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    float& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    void store(float* p) const { p[0] = x; p[1] = y; p[2] = z; }
};

inline Vec3 operator+(Vec3 a, Vec3 b) { return Vec3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline Vec3 operator-(Vec3 a, Vec3 b) { return Vec3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline Vec3 operator*(Vec3 a, Vec3 b) { return Vec3(a.x * b.x, a.y * b.y, a.z * b.z); }
inline Vec3 operator/(Vec3 a, Vec3 b) { return Vec3(a.x / b.x, a.y / b.y, a.z / b.z); }
inline Vec3 operator*(Vec3 a, float s) { return Vec3(a.x * s, a.y * s, a.z * s); }
inline Vec3 operator*(float s, Vec3 a) { return Vec3(s * a.x, s * a.y, s * a.z); }
inline Vec3 operator/(Vec3 a, float s) { return Vec3(a.x / s, a.y / s, a.z / s); }
inline Vec3 operator-(Vec3 a) { return Vec3(-a.x, -a.y, -a.z); }
inline Vec3& operator+=(Vec3& a, Vec3 b) { a = a + b; return a; }
inline Vec3& operator*=(Vec3& a, Vec3 b) { a = a * b; return a; }
inline Vec3& operator*=(Vec3& a, float s) { a = a * s; return a; }
inline bool operator==(Vec3 a, Vec3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }

inline float dot(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline Vec3 cross(Vec3 a, Vec3 b) {
    return Vec3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
inline float length(Vec3 a) { return std::sqrt(dot(a, a)); }
inline float length_squared(Vec3 a) { return dot(a, a); }
inline Vec3 normalize(Vec3 a) { return a * (1.0f / length(a)); }
inline Vec3 normalize_or(Vec3 a, Vec3 fallback) {
    float r = 1.0f / length(a);
    if (std::isfinite(r) && r > 0.0f) return a * r;
    return fallback;
}
inline Vec3 normalize_or_zero(Vec3 a) { return normalize_or(a, Vec3()); }
inline Vec3 vmin(Vec3 a, Vec3 b) { return Vec3(std::fmin(a.x, b.x), std::fmin(a.y, b.y), std::fmin(a.z, b.z)); }
inline Vec3 vmax(Vec3 a, Vec3 b) { return Vec3(std::fmax(a.x, b.x), std::fmax(a.y, b.y), std::fmax(a.z, b.z)); }
inline float max_element(Vec3 a) { return std::fmax(std::fmax(a.x, a.y), a.z); }
inline float min_element(Vec3 a) { return std::fmin(std::fmin(a.x, a.y), a.z); }
inline Vec3 vabs(Vec3 a) { return Vec3(std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)); }
inline Vec3 clamp(Vec3 a, Vec3 lo, Vec3 hi) { return vmin(vmax(a, lo), hi); }
inline Vec3 lerp(Vec3 a, Vec3 b, float s) { return a * (1.0f - s) + b * s; }
inline Vec3 vpow(Vec3 a, float e) { return Vec3(std::pow(a.x, e), std::pow(a.y, e), std::pow(a.z, e)); }
inline bool is_finite(Vec3 a) { return std::isfinite(a.x) && std::isfinite(a.y) && std::isfinite(a.z); }

// Rust's f32::clamp: NaN passes through; used where the reference clamps a scalar.
inline float clampf(float v, float lo, float hi) {
    if (v < lo) v = lo;
    if (v > hi) v = hi;
    return v;
}
// Rust's `f as u32`: saturating, NaN to 0.
inline uint32_t f2u_sat(float v) {
    if (!(v > 0.0f)) return 0u;
    if (v >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)v;
}
inline float copysignf_(float mag, float sgn) { return std::copysign(mag, sgn); }

}  // namespace oracle
