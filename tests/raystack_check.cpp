// CPU check of the traversal stack's bookkeeping (lucida_amd/csrc/raystack.hpp):
// random pushes, pops and truncations of RayStackT against a plain array
// stack with the same capacity, for several lanes sharing one ring and one
// spill buffer, as in a workgroup. Prints "ok" and exits 0, or the first
// disagreement and exits 1.
#define LUCIDA_HD
#include "raystack.hpp"

#include <cstdio>
#include <random>
#include <vector>

using lucida::RayStackT;

constexpr int LDS = 16, STRIDE = 8, LANES = 8, CAP = 32;
using Stack = RayStackT<LDS, STRIDE>;

int main() {
    std::vector<uint32_t> ring(LDS * STRIDE, 0xDEADBEEFu);
    std::vector<uint32_t> spill((size_t)CAP * LANES, 0xDEADBEEFu);
    std::mt19937 rng(7);
    long ops = 0;
    for (int round = 0; round < 200; ++round) {
        std::vector<Stack> st(LANES);
        std::vector<std::vector<uint32_t>> ref(LANES);
        for (int l = 0; l < LANES; ++l)
            st[l] = Stack{ring.data() + l, spill.data() + l, (uint32_t)LANES, 0, 0, CAP, 0u, false};
        for (int step = 0; step < 4000; ++step) {
            const int l = (int)(rng() % LANES);
            Stack& s = st[l];
            std::vector<uint32_t>& r = ref[l];
            // bias pushes for the first half of a round so stacks get deep
            const unsigned kind = rng() % 100;
            const unsigned push_share = step < 2000 ? 60 : 40;
            ++ops;
            if (kind < push_share) {
                const uint32_t v = rng();
                const bool ok = s.push(v);
                const bool want = r.size() < (size_t)CAP;
                if (ok != want) {
                    std::printf("push accepted=%d, expected %d (depth %zu)\n", ok, want, r.size());
                    return 1;
                }
                if (want) r.push_back(v);
            } else if (kind < 97) {
                if (r.empty()) continue;
                const uint32_t got = s.pop();
                if (got != r.back()) {
                    std::printf("pop %u, expected %u (depth %zu, round %d, step %d)\n", got, r.back(), r.size(),
                                round, step);
                    return 1;
                }
                r.pop_back();
            } else {
                const int base = r.empty() ? 0 : (int)(rng() % (r.size() + 1));
                s.truncate(base);
                if ((int)r.size() > base) r.resize(base);
            }
            if (s.sp != (int)r.size()) {
                std::printf("depth %d, expected %zu\n", s.sp, r.size());
                return 1;
            }
        }
        // drain every lane
        for (int l = 0; l < LANES; ++l)
            while (!ref[l].empty()) {
                if (st[l].pop() != ref[l].back()) {
                    std::printf("drain mismatch on lane %d\n", l);
                    return 1;
                }
                ref[l].pop_back();
            }
    }
    std::printf("ok %ld operations\n", ops);
    return 0;
}
