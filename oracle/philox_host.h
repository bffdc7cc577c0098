// philox_host.h -- host Philox4x32-10 and the CL noise draw, for the CPU oracle.
//
// TEST INFRASTRUCTURE (CPU oracle).  The reference draws its noise with cuRAND
// XORWOW seeded from the wall clock (common.h:582-590, 922-932), which no run can
// reproduce.  Both this oracle and the engine instead use the counter-based
// Philox4x32-10 generator of Salmon, Moraes, Dror and Shaw, "Parallel random
// numbers: as easy as 1, 2, 3" (SC'11): 10 rounds of the published round
// function with multipliers 0xD2511F53 / 0xCD9E8D57 and Weyl key increments
// 0x9E3779B9 / 0xBB67AE85.  Pinned by a known-answer fixture generated from
// PyTorch's open-source at::philox_engine (tests/golden/philox_kat.json).
//
// Noise draw for (seed, step t, field f, site i):
//   counter = {i, f, t_lo, t_hi}, key = {seed_lo, seed_hi}
//   r = Philox4x32-10(counter, key)
//   u1 = ((r0<<32 | r1) >> 11 + 0.5) * 2^-53,  u2 = ((r2<<32 | r3) >> 11 + 0.5) * 2^-53
//   n  = sqrt(-2 ln u1) * cos(2 pi u2)            (Box-Muller, one normal per site)
#pragma once

#include <cmath>
#include <cstdint>

namespace oracle {

inline void philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int round = 0; round < 10; ++round) {
        const uint64_t p0 = static_cast<uint64_t>(0xD2511F53u) * c0;
        const uint64_t p1 = static_cast<uint64_t>(0xCD9E8D57u) * c2;
        const uint32_t hi0 = static_cast<uint32_t>(p0 >> 32), lo0 = static_cast<uint32_t>(p0);
        const uint32_t hi1 = static_cast<uint32_t>(p1 >> 32), lo1 = static_cast<uint32_t>(p1);
        const uint32_t n0 = hi1 ^ c1 ^ k0;
        const uint32_t n1 = lo1;
        const uint32_t n2 = hi0 ^ c3 ^ k1;
        const uint32_t n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        if (round < 9) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

inline double cl_normal(uint64_t seed, uint64_t step, uint32_t field, uint32_t site)
{
    const uint32_t ctr[4] = {site, field, static_cast<uint32_t>(step), static_cast<uint32_t>(step >> 32)};
    const uint32_t key[2] = {static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32)};
    uint32_t r[4];
    philox4x32_10(ctr, key, r);
    const uint64_t a = ((static_cast<uint64_t>(r[0]) << 32) | r[1]) >> 11;
    const uint64_t b = ((static_cast<uint64_t>(r[2]) << 32) | r[3]) >> 11;
    const double u1 = (static_cast<double>(a) + 0.5) * 0x1.0p-53;
    const double u2 = (static_cast<double>(b) + 0.5) * 0x1.0p-53;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

// both Box-Muller normals of one draw: cl_normal's value (cos) and its sine partner
inline void cl_normal_pair(uint64_t seed, uint64_t step, uint32_t field, uint32_t site, double* out)
{
    const uint32_t ctr[4] = {site, field, static_cast<uint32_t>(step), static_cast<uint32_t>(step >> 32)};
    const uint32_t key[2] = {static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32)};
    uint32_t r[4];
    philox4x32_10(ctr, key, r);
    const uint64_t a = ((static_cast<uint64_t>(r[0]) << 32) | r[1]) >> 11;
    const uint64_t b = ((static_cast<uint64_t>(r[2]) << 32) | r[3]) >> 11;
    const double u1 = (static_cast<double>(a) + 0.5) * 0x1.0p-53;
    const double u2 = (static_cast<double>(b) + 0.5) * 0x1.0p-53;
    const double rad = std::sqrt(-2.0 * std::log(u1));
    out[0] = rad * std::cos(6.283185307179586 * u2);
    out[1] = rad * std::sin(6.283185307179586 * u2);
}

} // namespace oracle
