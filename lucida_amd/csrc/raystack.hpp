// A lane's BVH traversal stack (bvh.wgsl:126-133: 32 entries, a push onto a
// full stack is dropped). The newest entry is held in a register (a pop
// right after a push, the common case below a node whose near child is a
// leaf, needs no LDS round trip), the next LDS entries in an LDS ring (one
// column per lane of the block, RING_STRIDE words apart), older ones in a
// per-lane column of device memory, so the LDS held per wave stays small and
// eight waves fit on a SIMD; deep stacks are rare, so the spill column is
// rarely touched.
//
// The methods are host-callable too, so the bookkeeping is checked on the
// CPU against a plain array stack (tests/raystack_check.cpp).
#pragma once
#include <cstddef>
#include <cstdint>

#ifndef LUCIDA_HD
#define LUCIDA_HD __host__ __device__
#endif

namespace lucida {

template <int LDS, int RING_STRIDE>
struct RayStackT {
    static_assert(LDS > 0 && (LDS & (LDS - 1)) == 0, "the ring size is a power of two");
    uint32_t* ring;        // entry k at ring[slot(k)] = ring[(k % LDS) * RING_STRIDE]
    uint32_t* spill;       // entry k at spill[k * spill_stride]
    uint32_t spill_stride;
    int sp, lo;            // entries [0, lo) spilled, [lo, sp - held) in the ring
    int cap;               // entries at most (32 for the binary walks)
    uint32_t top;          // entry sp - 1 when held
    bool held;

    // Ring word of entry k (k >= 0; LDS a power of two, so this is a mask).
    LUCIDA_HD static inline uint32_t slot(int k) { return ((uint32_t)k % (uint32_t)LDS) * (uint32_t)RING_STRIDE; }
    // Stores entry k into the ring, spilling the ring's oldest entry if full.
    LUCIDA_HD inline void sink(int k, uint32_t v) {
        if (k - lo == LDS) {
            spill[(size_t)lo * spill_stride] = ring[slot(lo)];
            ++lo;
        }
        ring[slot(k)] = v;
    }
    LUCIDA_HD inline bool push(uint32_t v) {
        if (sp >= cap) return false;
        if (held) sink(sp - 1, top);
        top = v;
        held = true;
        ++sp;
        return true;
    }
    // Requires sp > 0.
    LUCIDA_HD inline uint32_t pop() {
        --sp;
        if (held) {
            held = false;
            return top;
        }
        if (sp < lo) {
            lo = sp;
            return spill[(size_t)sp * spill_stride];
        }
        return ring[slot(sp)];
    }
    // Drops the entries above `base` (a walk that returned early). The entry
    // below them, if any, was stored by the first push above it.
    LUCIDA_HD inline void truncate(int base) {
        if (sp > base) {
            sp = base;
            held = false;
            if (lo > base) lo = base;
        }
    }
};

}  // namespace lucida
