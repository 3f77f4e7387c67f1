// Exact division by the common factor of a 17-term group, 13^2 17^3 7^6
// 11^3 = 830297 * 156590819 (pi_bs_gpu.cu:151-152, 760-779): the constants
// and one Hensel step, as plain host/device code so that the CPU can check
// them against 128-bit arithmetic.
#pragma once

#include <stdint.h>

#ifndef GPUPI_HD
#if defined(__HIPCC__)
#define GPUPI_HD __host__ __device__ __forceinline__
#else
#define GPUPI_HD inline
#endif
#endif

namespace gpupi {

constexpr uint64_t kBlockDivisor = 130016887243243ull;  // 830297 * 156590819 (47 bits)
constexpr uint32_t kBlockDivisorInv = 0x33261ec3u;      // its inverse mod 2^32

// One step of the division from the low limb up: the quotient limb
// q = (a - c) d^-1 mod 2^32, and the carry becomes (q d + c - a) / 2^32.
// With d = d_lo + 2^32 d_hi and c = c_lo + 2^32 c_hi, q d_lo + c_lo - a is
// 0 mod 2^32 and above -2^32, so it equals 2^32 [a < c_lo], and
// c' = q d_hi + mulhi(q, d_lo) + c_hi + [a < c_lo] (< d, no 128-bit product).
GPUPI_HD uint32_t hensel_step(uint32_t a, uint64_t &c) {
    constexpr uint32_t kDlo = (uint32_t)kBlockDivisor, kDhi = (uint32_t)(kBlockDivisor >> 32);
    const uint32_t clo = (uint32_t)c, chi = (uint32_t)(c >> 32);
    const uint32_t q = (a - clo) * kBlockDivisorInv;
    const uint64_t s = (((uint64_t)q * kDlo) >> 32) + chi + (a < clo ? 1u : 0u);
    c = (uint64_t)q * kDhi + s;
    return q;
}

}  // namespace gpupi
