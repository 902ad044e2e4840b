// The 4-wide BVH the tracer walks by default: collapsed on the device from the
// reference-layout binary BVH (GpuBvhNode), one wide node per binary internal
// node at even depth, whose children are its binary grandchildren (or its
// leaf children). 128 bytes per node: child boxes as structure-of-arrays, so a
// visit is seven 16-byte loads from one 128-byte line.
#pragma once
#include <cstdint>

namespace lucida {

struct alignas(128) WideNode {
    float lo_x[4], lo_y[4], lo_z[4];
    float hi_x[4], hi_y[4], hi_z[4];
    uint32_t child[4];  // WIDE_EMPTY, WIDE_LEAF | (first << 5) | count, or a wide node index
    uint32_t _pad[4];
};
static_assert(sizeof(WideNode) == 128, "WideNode is one 128-byte line");

constexpr uint32_t WIDE_LEAF = 0x80000000u;
constexpr uint32_t WIDE_EMPTY = 0xFFFFFFFFu;

}  // namespace lucida
