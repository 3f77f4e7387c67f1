// CPU check of products run in pieces (gpupi_amd/csrc/ragged.hpp, the header
// the ragged NTT launches and the assembly kernel use): for random shapes of
// both kinds the pieces (piece_row) cover the longer operand exactly, every
// staged word the assembly reads (assembly_source) lies inside its piece's
// product, and summing those words with carries gives the schoolbook
// product of the two operands.  Exit status 0 = all checks passed.
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

#include "ragged.hpp"

using namespace gpupi;

static int fails = 0;
#define EXPECT(c, ...) do { if (!(c)) { if (fails++ < 20) { printf("FAIL: " __VA_ARGS__); printf("\n"); } } } while (0)

// schoolbook a * b (little-endian 32-bit limbs), |a| + |b| limbs
static std::vector<uint32_t> mul(const uint32_t *a, size_t la, const uint32_t *b, size_t lb) {
    std::vector<uint32_t> c(la + lb, 0);
    for (size_t i = 0; i < la; ++i) {
        uint64_t carry = 0;
        for (size_t j = 0; j < lb; ++j) {
            const uint64_t t = (uint64_t)a[i] * b[j] + c[i + j] + carry;
            c[i + j] = (uint32_t)t;
            carry = t >> 32;
        }
        for (size_t k = i + lb; carry != 0; ++k) {
            const uint64_t t = (uint64_t)c[k] + carry;
            c[k] = (uint32_t)t;
            carry = t >> 32;
        }
    }
    return c;
}

static uint32_t limb(std::mt19937_64 &rng, int mode) {
    if (mode == 1) return 0xffffffffu;
    if (mode == 2) return (rng() % 4 == 0) ? (uint32_t)rng() : 0u;
    return (uint32_t)rng();
}

int main() {
    std::mt19937_64 rng(12345);
    int cases = 0;
    for (int trial = 0; trial < 4000; ++trial) {
        const int mode = trial % 3;  // random, all ones, sparse
        const uint32_t S = 1 + (uint32_t)(rng() % 48);
        const uint32_t m = 2 + (uint32_t)(rng() % 9);
        // two pieces may be shorter than S (split at K/2); more are not
        const uint32_t s = m == 2 ? 1 + (uint32_t)(rng() % 80) : S + (uint32_t)(rng() % 64);
        const uint32_t last = 1 + (uint32_t)(rng() % s);
        const uint32_t LL = (m - 1) * s + last;
        std::vector<uint32_t> L(LL), Sv(S);
        for (auto &x : L) x = limb(rng, mode);
        for (auto &x : Sv) x = limb(rng, mode);
        // the pieces tile L and are staged as their products
        std::vector<std::vector<uint32_t>> staged(m);
        uint32_t covered = 0;
        for (uint32_t i = 0; i < m; ++i) {
            const RaggedRow r = piece_row(i, L.data(), LL, Sv.data(), S, m, s, 7);
            EXPECT(r.A == L.data() + covered && r.B == Sv.data() && r.C == nullptr && r.L_b == S && r.b == 7,
                   "row %u of m=%u s=%u S=%u", i, m, s, S);
            EXPECT(r.L_a >= 1 && r.L_a <= s, "piece length %u (s=%u)", r.L_a, s);
            covered += r.L_a;
            staged[i] = mul(r.A, r.L_a, Sv.data(), S);
        }
        EXPECT(covered == LL, "pieces cover %u of %u limbs", covered, LL);
        // the assembly: staged words summed with carries
        const RaggedAssembly a{nullptr, LL + S, s, S, m, 0, 0};
        std::vector<uint32_t> out(a.c);
        uint64_t carry = 0;
        for (uint32_t x = 0; x < a.c; ++x) {
            uint32_t i, off;
            bool below;
            assembly_source(a, x, &i, &off, &below);
            uint64_t t = carry;
            EXPECT(i < m && off < staged[i].size(), "x=%u reads piece %u word %u (m=%u s=%u S=%u)", x, i, off, m, s, S);
            if (i < m && off < staged[i].size()) t += staged[i][off];
            if (below) {
                EXPECT(off + s < staged[i - 1].size(), "x=%u reads piece %u word %u", x, i - 1, off + s);
                if (off + s < staged[i - 1].size()) t += staged[i - 1][off + s];
            }
            out[x] = (uint32_t)t;
            carry = t >> 32;
        }
        EXPECT(carry == 0, "carry out of the top limb (m=%u s=%u S=%u)", m, s, S);
        EXPECT(out == mul(L.data(), LL, Sv.data(), S), "assembled product differs (m=%u s=%u S=%u last=%u mode=%d)",
               m, s, S, last, mode);
        EXPECT(assembly_chunks(a.c) == (a.c + kAssembleChunk - 1) / kAssembleChunk, "chunks of %u", a.c);
        ++cases;
    }
    printf("%d cases, %d failures\n", cases, fails);
    if (fails == 0) printf("ragged assembly check: ok\n");
    return fails == 0 ? 0 : 1;
}
