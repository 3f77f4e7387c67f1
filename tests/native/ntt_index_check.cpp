// CPU check of the NTT's index math (gpupi_amd/csrc/ntt_plan.hpp): every
// tile of every level stays inside [0, K), the tiles of a level partition
// [0, K) exactly, every in-tile step partitions the 4096-element tile, LDS
// slots stay below the allocation, and the final level's runs/spill slots
// stay in range.  Exit status 0 = all checks passed.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "ntt_plan.hpp"

using namespace gpupi::ntt;

static int fails = 0;
#define EXPECT(c, ...) do { if (!(c)) { if (fails++ < 20) { printf("FAIL: " __VA_ARGS__); printf("\n"); } } } while (0)

static void check_tile_steps() {
    for (int m = 1; m <= kTileLog2; ++m) {
        for (int s = 0; s < step_count(m); ++s) {
            const int lr = step_log2radix(m, s);
            EXPECT(lr >= 1 && lr <= 4, "radix m=%d s=%d", m, s);
            const uint32_t R = 1u << lr;
            std::vector<int> seen(kTile, 0);
            for (uint32_t q = 0; q < kTile / R; ++q) {
                const SmallDft d = small_dft(m, s, q);
                for (uint32_t j = 0; j < R; ++j) {
                    const uint32_t idx = d.base + j * d.stride;
                    EXPECT(idx < kTile, "step idx m=%d s=%d q=%u", m, s, q);
                    if (idx < kTile) seen[idx]++;
                    EXPECT(lds_index(idx) < kLdsSlots, "lds slot");
                }
                EXPECT((d.n2 << lr) < (1u << d.log2Ms) || d.log2Ms == lr, "twiddle exponent range");
            }
            for (uint32_t i = 0; i < kTile; ++i) EXPECT(seen[i] == 1, "step coverage m=%d s=%d i=%u", m, s, i);
        }
    }
}

static void check_plan(int k, bool full) {
    const Plan p = make_plan(k);
    const uint64_t K = 1ull << k;
    int sum = 0;
    for (int l = 0; l < p.levels; ++l) sum += p.log2M[l];
    EXPECT(sum == k, "plan sum k=%d", k);
    if (p.levels == 1) return;
    for (int l = 0; l + 1 < p.levels; ++l) EXPECT(p.log2M[l] <= kColMaxLog2 && p.log2M[l] >= 3, "col size k=%d", k);
    EXPECT(p.log2M[p.levels - 1] <= kTileLog2, "row size k=%d", k);
    const uint32_t tiles = tiles_per_product(p);
    EXPECT((uint64_t)tiles * kTile == K, "tiles k=%d", k);
    for (int l = 0; l < p.levels; ++l) {
        std::vector<uint8_t> seen(full ? K : 0, 0);
        const bool col = l + 1 < p.levels;
        const int lm = p.log2M[l];
        const uint32_t C = 1u << (kTileLog2 - lm);
        const uint32_t step = full ? 1 : (tiles > 64 ? tiles / 61 : 1);
        for (uint32_t t = 0; t < tiles; t += step) {
            for (uint32_t i = 0; i < kTile; ++i) {
                uint64_t off;
                if (col) {
                    const uint32_t g = i & (C - 1), r = i >> (kTileLog2 - lm);
                    off = col_offset(p, l, col_tile(p, l, t), g, r);
                } else {
                    off = row_offset(p, t, i >> lm, i & ((1u << lm) - 1));
                }
                EXPECT(off < K, "offset k=%d l=%d t=%u i=%u", k, l, t, i);
                if (full && off < K) seen[off]++;
            }
        }
        if (full)
            for (uint64_t x = 0; x < K; ++x) EXPECT(seen[x] == 1, "coverage k=%d l=%d x=%llu", k, l, (unsigned long long)x);
        if (l == 0) {
            // final inverse level: runs of C coefficients and their spill slots
            const uint32_t lc = kTileLog2 - lm;
            EXPECT(final_run_log2(p) == lc, "run length k=%d", k);
            const int ls = slab_log2(p, 0) - lm;
            for (uint32_t t = 0; t < tiles; t += step) {
                const ColTile ct = col_tile(p, 0, t);
                EXPECT(ct.slab == 0, "slab 0 at level 0");
                for (uint32_t r = 0; r < (1u << lm); ++r) {
                    const uint64_t n0 = ((uint64_t)r << ls) + ct.col0;
                    EXPECT(n0 % C == 0 && n0 + C <= K, "run start k=%d", k);
                    EXPECT((n0 >> lc) < (K >> lc), "spill slot k=%d", k);
                }
            }
        }
    }
}

int main() {
    check_tile_steps();
    for (int k = kMinLog2; k <= kMaxLog2; ++k) check_plan(k, k <= 23);
    printf(fails ? "index check: %d failures\n" : "index check: ok\n", fails);
    return fails ? 1 : 0;
}
