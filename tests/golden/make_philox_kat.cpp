// Writes Philox4x32-10 known-answer vectors computed with rocRAND's
// host-callable philox4x32_10_engine (an independent implementation from
// /opt/rocm/include/rocrand).  Build and run:
//   hipcc -O1 -o /tmp/make_philox_kat tests/golden/make_philox_kat.cpp
//   /tmp/make_philox_kat > tests/golden/philox_kat_rocrand.json
// The engine is keyed by the 64-bit seed (key = lo, hi); its counter starts
// at {offset_lo, offset_hi, subseq_lo, subseq_hi} when seeded with
// subsequence = (c3<<32|c2) and offset = 4*(c1<<32|c0); c1 < 2^30 keeps
// that offset inside 64 bits.
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>
#include <cstdint>

int main() {
    struct Case { uint32_t c[4]; uint32_t k[2]; };
    const Case cases[] = {
        {{0u, 0u, 0u, 0u}, {0u, 0u}},
        {{0xffffffffu, 0x3fffffffu, 0xffffffffu, 0xffffffffu}, {0xffffffffu, 0xffffffffu}},
        {{0x243f6a88u, 0x05a308d3u, 0x13198a2eu, 0x03707344u}, {0xa4093822u, 0x299f31d0u}},
        {{0xfffffff0u, 0x00001234u, 0xdeadbeefu, 0x7fffffffu}, {0x01234567u, 0x89abcdefu}},
        {{1u, 2u, 3u, 4u}, {5u, 6u}},
        {{12345u, 4999u, 9u, (7u << 4) | 1u}, {0x9e3779b9u, 0x12345678u}},
        {{99999u, 0u, 0u, 3u}, {42u, 0u}},
        {{0u, 1u, 0u, 0u}, {20251203u, 0u}},
    };
    std::printf("[\n");
    const int n = sizeof(cases) / sizeof(cases[0]);
    for (int i = 0; i < n; ++i) {
        const Case& cs = cases[i];
        unsigned long long seed = ((unsigned long long)cs.k[1] << 32) | cs.k[0];
        unsigned long long subseq = ((unsigned long long)cs.c[3] << 32) | cs.c[2];
        unsigned long long off = 4ull * (((unsigned long long)cs.c[1] << 32) | cs.c[0]);
        rocrand_device::philox4x32_10_engine eng(seed, subseq, off);
        uint32_t o[4];
        for (int j = 0; j < 4; ++j) o[j] = eng();
        std::printf("  {\"ctr\": [%u, %u, %u, %u], \"key\": [%u, %u], \"out\": [%u, %u, %u, %u]}%s\n",
                    cs.c[0], cs.c[1], cs.c[2], cs.c[3], cs.k[0], cs.k[1], o[0], o[1], o[2], o[3],
                    i + 1 < n ? "," : "");
    }
    std::printf("]\n");
    return 0;
}
