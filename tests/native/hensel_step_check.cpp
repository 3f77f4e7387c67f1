// CPU check of the 17-term group stage's exact division (gpupi_amd/csrc/
// exact_div.hpp, the header the kernel includes): the divisor and its
// inverse, the short carry step against the 128-bit step (q d + c - a) / 2^32
// on random limbs and carries (and on the edges c = 0, c = d, a = 0,
// a = 2^32 - 1), and whole divisions of random multiples of the divisor.
#include <cstdio>
#include <random>
#include <vector>

#include "exact_div.hpp"

using namespace gpupi;

static int fails = 0;
#define EXPECT(c, ...) do { if (!(c)) { if (fails++ < 20) { printf("FAIL: " __VA_ARGS__); printf("\n"); } } } while (0)

static void step_vs_wide(uint32_t a, uint64_t c) {
    const uint32_t q = (a - (uint32_t)c) * kBlockDivisorInv;
    const unsigned __int128 t = (unsigned __int128)q * kBlockDivisor + c - a;
    uint64_t c2 = c;
    const uint32_t q2 = hensel_step(a, c2);
    EXPECT(q2 == q && c2 == (uint64_t)(t >> 32) && c2 <= kBlockDivisor, "step a=%u c=%llu", a,
           (unsigned long long)c);
}

int main() {
    EXPECT(kBlockDivisor == 830297ull * 156590819ull, "divisor");
    EXPECT((uint32_t)(kBlockDivisor * kBlockDivisorInv) == 1u, "inverse");
    std::mt19937_64 rng(7);
    for (uint64_t c : {(uint64_t)0, (uint64_t)1, kBlockDivisor - 1, kBlockDivisor})
        for (uint32_t a : {0u, 1u, 0x7fffffffu, 0xffffffffu}) step_vs_wide(a, c);
    for (long it = 0; it < 5000000; ++it) step_vs_wide((uint32_t)rng(), rng() % (kBlockDivisor + 1));
    // whole divisions: A = Q d (n limbs) divides back to Q
    for (int trial = 0; trial < 2000; ++trial) {
        const int n = 1 + (int)(rng() % 90);
        std::vector<uint32_t> Q(n), A(n);
        for (int i = 0; i + 2 < n; ++i) Q[i] = (uint32_t)rng();  // Q d fits in n limbs
        unsigned __int128 carry = 0;
        for (int i = 0; i < n; ++i) {
            carry += (unsigned __int128)Q[i] * kBlockDivisor;
            A[i] = (uint32_t)carry;
            carry >>= 32;
        }
        EXPECT(carry == 0, "test setup");
        uint64_t c = 0;
        for (int i = 0; i < n; ++i) EXPECT(hensel_step(A[i], c) == Q[i], "quotient limb %d of %d", i, n);
        EXPECT(c == 0, "exact division leaves no carry (n=%d)", n);
    }
    if (fails) {
        printf("hensel check: %d failures\n", fails);
        return 1;
    }
    printf("hensel check: ok\n");
    return 0;
}
