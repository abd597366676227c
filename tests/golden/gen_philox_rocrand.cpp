// Writes Philox4x32-10 outputs computed by rocRAND's header-only engine
// (/opt/rocm/include/rocrand/rocrand_philox4x32_10.h) for a set of
// (counter, key) pairs, as a golden fixture for the oracle's Philox.
//
// rocRAND's engine maps (seed, subsequence, offset) to key = seed,
// counter = (offset/4 low, offset/4 high, subsequence low, subsequence high);
// we pick offset = 4 * (c1 << 32 | c0) and subsequence = (c3 << 32 | c2),
// which requires c1 < 2^30 (the offset is a 64-bit count of words).
//
// Build and run (host code only, no GPU needed):
//   hipcc -O1 -o /tmp/gen_philox tests/golden/gen_philox_rocrand.cpp
//   /tmp/gen_philox > tests/golden/philox_rocrand.txt
#include <rocrand/rocrand_philox4x32_10.h>

#include <cstdint>
#include <cstdio>

int main() {
    const uint32_t cases[][6] = {
        // c0, c1, c2, c3, k0, k1
        {0u, 0u, 0u, 0u, 0u, 0u},
        {0xffffffffu, 0x3fffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu},
        {0x243f6a88u, 0x05a308d3u, 0x13198a2eu, 0x03707344u, 0xa4093822u, 0x299f31d0u},
        {1u, 0u, 0u, 0u, 1234u, 0u},
        {12345u, 678u, 0u, 0u, 1234u, 0u},
        {7u, 3u, 1u << 24, 0u, 99u, 5u},
        {65535u, 1u << 20, 2u << 24, 42u, 0xdeadbeefu, 0x12345678u},
        {999999u, 124999u, 0u, 0u, 1234u, 0u},
    };
    for (const auto& c : cases) {
        const unsigned long long seed = ((unsigned long long)c[5] << 32) | c[4];
        const unsigned long long sub = ((unsigned long long)c[3] << 32) | c[2];
        const unsigned long long off = 4ull * (((unsigned long long)c[1] << 32) | c[0]);
        rocrand_device::philox4x32_10_engine eng(seed, sub, off);
        const uint4 r = eng.next4();
        std::printf("%08x %08x %08x %08x %08x %08x  %08x %08x %08x %08x\n", c[0], c[1], c[2], c[3], c[4],
                    c[5], r.x, r.y, r.z, r.w);
    }
    return 0;
}
