// Writes the Philox4x32-10 known-answer fixture (philox_kat.json) from PyTorch's
// open-source at::philox_engine (ATen/core/PhiloxRNGEngine.h): the raw 4x32-bit
// outputs for a set of (counter, key) pairs.  philox_engine(seed, subsequence,
// offset) sets key = {seed_lo, seed_hi}, counter = {offset_lo, offset_hi,
// subsequence_lo, subsequence_hi}; operator() returns the four outputs in order.
// Build and run with tests/golden/make_philox_kat.sh.
#include <ATen/core/PhiloxRNGEngine.h>

#include <cstdint>
#include <cstdio>

int main()
{
    struct Case { uint64_t seed, subseq, offset; };
    const Case cases[] = {
        {0ull, 0ull, 0ull},
        {0xFFFFFFFFFFFFFFFFull, 0xFFFFFFFFFFFFFFFFull, 0xFFFFFFFFFFFFFFFFull},
        {0x299f31d0a4093822ull, 0x0370734413198a2eull, 0x85a308d3243f6a88ull},
        {0x5EED0000ull, 0ull, 7ull},
        {123456789ull, 42ull, 1000ull},
        {0xC4ull, (0x100ull << 32) | 884735ull, 17ull},
    };
    std::printf("[\n");
    const int n = sizeof(cases) / sizeof(cases[0]);
    for (int c = 0; c < n; ++c) {
        at::philox_engine eng(cases[c].seed, cases[c].subseq, cases[c].offset);
        uint32_t out[4];
        for (int k = 0; k < 4; ++k) out[k] = eng();
        const uint32_t ctr[4] = {static_cast<uint32_t>(cases[c].offset), static_cast<uint32_t>(cases[c].offset >> 32),
                                 static_cast<uint32_t>(cases[c].subseq), static_cast<uint32_t>(cases[c].subseq >> 32)};
        const uint32_t key[2] = {static_cast<uint32_t>(cases[c].seed), static_cast<uint32_t>(cases[c].seed >> 32)};
        std::printf("  {\"counter\": [%u, %u, %u, %u], \"key\": [%u, %u], \"out\": [%u, %u, %u, %u]}%s\n", ctr[0], ctr[1], ctr[2], ctr[3],
                    key[0], key[1], out[0], out[1], out[2], out[3], c + 1 < n ? "," : "");
    }
    std::printf("]\n");
    return 0;
}
