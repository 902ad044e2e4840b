// The optional 4-wide BVH: collapsed on the device from the reference-layout
// binary BVH (GpuBvhNode), one wide node per binary internal node at even
// depth, whose children are its binary grandchildren (or its leaf children).
// 64 bytes per node: the node's box minimum and per-axis power-of-two scales,
// the four child words, and each child's box as 8-bit offsets from that
// minimum in units of the scale, rounded outward (the decoded box contains the
// child's exact box), so a visit is four 16-byte loads.
#pragma once
#include <cstdint>

namespace lucida {

struct alignas(64) WideNode {
    float origin[3];    // the node's box minimum
    int8_t exp[3];      // child box planes are origin + q * 2^exp (per axis)
    uint8_t _pad0;
    uint32_t child[4];  // WIDE_EMPTY, WIDE_LEAF | (first << 5) | count, or a wide node index
    uint8_t qlo_x[4], qlo_y[4], qlo_z[4];
    uint8_t qhi_x[4], qhi_y[4], qhi_z[4];
    uint32_t _pad1[2];
};
static_assert(sizeof(WideNode) == 64, "WideNode is 64 bytes");

constexpr uint32_t WIDE_LEAF = 0x80000000u;
constexpr uint32_t WIDE_EMPTY = 0xFFFFFFFFu;
// Largest scale exponent accepted: keeps 2^exp times a clamped reciprocal
// direction (at most 1e16) finite.
constexpr int WIDE_MAX_EXP = 64;

}  // namespace lucida
