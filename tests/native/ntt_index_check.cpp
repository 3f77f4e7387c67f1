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
    for (int col = 0; col < 2; ++col)
    for (int m = 1; m <= kTileLog2; ++m) {
        for (int s = 0; s < step_count(m); ++s) {
            const int lr = step_log2radix(m, s);
            EXPECT(lr >= 1 && lr <= 4, "radix m=%d s=%d", m, s);
            const uint32_t R = 1u << lr;
            std::vector<int> seen(kTile, 0);
            for (uint32_t q = 0; q < kTile / R; ++q) {
                const SmallDft d = small_dft(m, s, q, col != 0);
                for (uint32_t j = 0; j < R; ++j) {
                    const uint32_t idx = d.base + j * d.stride;
                    EXPECT(idx < kTile, "step idx m=%d s=%d q=%u", m, s, q);
                    if (idx < kTile) seen[idx]++;
                    EXPECT(lds_index(idx) < kLdsSlots, "lds slot");
                }
                EXPECT((d.n2 << lr) < (1u << d.log2Ms), "twiddle exponent range");
                const uint32_t lss = d.log2Ms - lr;
                if (lss > 0) {
                    // twiddled steps are radix 16; table slot j*Ss + n2 < Ms
                    EXPECT(lr == 4, "twiddled step radix m=%d s=%d", m, s);
                    EXPECT(15u * (1u << lss) + d.n2 < (1u << d.log2Ms), "step table slot");
                }
                for (uint32_t j = 0; j < R; ++j)
                    EXPECT(lds_index(d.base + j * d.stride) == lds_index(d.base) + lds_index(j * d.stride),
                           "constant LDS offsets m=%d s=%d q=%u j=%u col=%d", m, s, q, j, col);
                // byte offsets fit the ds_read immediate field
                EXPECT(lds_index((R - 1) * d.stride) * 8u < 65536u, "ds offset");
            }
            EXPECT(step_ms_log2(m, s) == small_dft(m, s, 0, col != 0).log2Ms, "ms");
            for (uint32_t i = 0; i < kTile; ++i) EXPECT(seen[i] == 1, "step coverage m=%d s=%d i=%u", m, s, i);
            // each small DFT's points are r = b*Ms + j*Ss + n2 of one DFT g
            for (uint32_t q = 0; q < kTile / R; q += 7) {
                const SmallDft d = small_dft(m, s, q, col != 0);
                const uint32_t lg = kTileLog2 - m;
                for (uint32_t j = 0; j < R; ++j) {
                    const uint32_t idx = d.base + j * d.stride;
                    const uint32_t g = col ? (idx & ((1u << lg) - 1u)) : (idx >> m);
                    const uint32_t r = col ? (idx >> lg) : (idx & ((1u << m) - 1u));
                    const uint32_t g0 = col ? (d.base & ((1u << lg) - 1u)) : (d.base >> m);
                    const uint32_t Ss = 1u << (d.log2Ms - lr);
                    EXPECT(g == g0, "same DFT m=%d s=%d", m, s);
                    EXPECT((r & ((1u << d.log2Ms) - 1u)) == j * Ss + d.n2, "point order m=%d s=%d", m, s);
                }
            }
        }
    }
}

static void check_bitrev_split() {
    // bitrev(r0 + (u << (lm - 4)), lm) == bitrev(r0, lm) + bitrev(u, 4)
    for (int lm = 5; lm <= kColMaxLog2; ++lm)
        for (uint32_t r0 = 0; r0 < (1u << (lm - 4)); ++r0)
            for (uint32_t u = 0; u < 16; ++u)
                EXPECT(bitrev(r0 + (u << (lm - 4)), lm) == bitrev(r0, lm) + bitrev(u, 4), "bitrev split lm=%d", lm);
}

static void check_plan(const Plan &p, bool full) {
    const int k = p.log2K;
    const uint64_t K = 1ull << k;
    int sum = 0;
    for (int l = 0; l < p.levels; ++l) sum += p.log2M[l];
    EXPECT(sum == k, "plan sum k=%d", k);
    if (p.levels == 1) return;
    for (int l = 0; l + 1 < p.levels; ++l) EXPECT(p.log2M[l] <= kColMaxLog2 && p.log2M[l] >= 3, "col size k=%d", k);
    EXPECT(p.log2M[p.levels - 1] <= kTileLog2, "row size k=%d", k);
    // kernel instantiations: column levels 5..9, row level 6..12
    for (int l = 0; l + 1 < p.levels; ++l) EXPECT(p.log2M[l] >= 5 && p.log2M[l] <= 9, "col inst k=%d", k);
    EXPECT(p.log2M[p.levels - 1] >= 6 && p.log2M[p.levels - 1] <= 12, "row inst k=%d", k);
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
                // padded plane position: inside the product's plane
                const uint64_t pos = plane_index(p, off);
                EXPECT(pos < plane_row_elems(p), "plane position k=%d l=%d", k, l);
                if (col) {
                    // thread tid = i mod 256, u = i / 256: offset and twiddle exponent
                    const ColTile ct = col_tile(p, l, t);
                    const uint32_t tid = i & 255u, u = i >> 8;
                    const uint64_t want = col_thread_offset(p, l, ct, tid) + ((uint64_t)u << col_row_step_log2(p, l));
                    EXPECT(off == want, "col thread offset k=%d l=%d", k, l);
                    // the kernels' plane position: base + u * step
                    const int ps = plane_pad_shift(p);
                    const int us = col_row_step_log2(p, l);
                    const uint64_t pstep = (1ull << us) + ((ps && us >= ps) ? ((uint64_t)kPlanePad << (us - ps)) : 0ull);
                    EXPECT(pos == plane_index(p, col_thread_offset(p, l, ct, tid)) + u * pstep,
                           "col plane step k=%d l=%d", k, l);
                    const int lk = slab_log2(p, l);
                    const uint32_t g = i & (C - 1), r = i >> (kTileLog2 - lm);
                    const uint64_t cg = (uint64_t)(ct.col0 + g) << (kMaxLog2 - lk);
                    EXPECT(cg < (1ull << kMaxLog2) && cg * bitrev(r, lm) < (1ull << kMaxLog2), "twiddle exp k=%d", k);
                    // last-step layout: value k = i / 256 of thread tid is
                    // the element the last in-tile step gives it, at a plane
                    // position linear in its row, with twiddle multiplier
                    // bitrev(row); the loaded step exponent stays below 2^30
                    const int lr = col_last_lr(lm);
                    const int kv = (int)u;
                    const SmallDft d = small_dft(lm, step_count(lm) - 1, tid + 256u * (u >> lr), true);
                    const uint32_t idx = d.base + (u & ((1u << lr) - 1u)) * d.stride;
                    const uint32_t rl = col_last_row0(lm, tid) + col_last_row_delta(lm, kv);
                    EXPECT(d.log2Ms == lr && idx == rl * C + (tid & (C - 1)), "last-step element k=%d l=%d", k, l);
                    const uint64_t p0 = plane_index(p, col_offset(p, l, ct, tid & (C - 1), col_last_row0(lm, tid)));
                    EXPECT(plane_index(p, col_offset(p, l, ct, tid & (C - 1), rl)) ==
                               p0 + (uint64_t)col_last_row_delta(lm, kv) * col_plane_row_step(p, l),
                           "last-step plane position k=%d l=%d", k, l);
                    EXPECT(col_last_brow(lm, tid, kv) == bitrev(rl, lm), "last-step frequency k=%d l=%d", k, l);
                    const uint64_t es = (16 >> lr) == 1 ? cg << (lm - lr) : cg;
                    EXPECT(es < (1ull << kMaxLog2) && cg * col_last_b0(lm, tid) < (1ull << kMaxLog2),
                           "last-step twiddle exponents k=%d", k);
                } else {
                    EXPECT(off == (uint64_t)t * kTile + i, "row offset contiguous k=%d", k);
                    // a row tile stays contiguous in the padded plane
                    EXPECT(pos == plane_index(p, (uint64_t)t * kTile) + i, "row tile padded contiguous k=%d", k);
                }
            }
        }
        if (full)
            for (uint64_t x = 0; x < K; ++x) EXPECT(seen[x] == 1, "coverage k=%d l=%d x=%llu", k, l, (unsigned long long)x);
        // padding: rows of the first level's view start kPlanePad apart from
        // a power-of-two distance, and positions stay one to one
        if (plane_pad_shift(p)) {
            const int ps = plane_pad_shift(p);
            EXPECT(plane_index(p, 1ull << ps) == (1ull << ps) + kPlanePad, "pad k=%d", k);
            EXPECT(plane_index(p, K - 1) == plane_row_elems(p) - 1 - kPlanePad, "pad end k=%d", k);
        }
        if (l == 0) {
            // final inverse level: thread tid's element u is coefficient
            // col_thread_offset(tid) + u 2^(k-4); the 2^rl lanes of a run hold
            // adjacent coefficients of one row, the run starts at a multiple
            // of 2^rl inside the product, and its spill slot exists
            const uint32_t rl = final_run_log2(p);
            const uint32_t run = 1u << rl;
            EXPECT(rl <= 4 && run <= C && 64 % run == 0, "run length k=%d", k);
            for (uint32_t t = 0; t < tiles; t += step) {
                const ColTile ct = col_tile(p, 0, t);
                EXPECT(ct.slab == 0, "slab 0 at level 0");
                for (uint32_t tid = 0; tid < 256; ++tid) {
                    const uint32_t jr = tid & (run - 1);
                    for (uint32_t u = 0; u < 16; ++u) {
                        const uint64_t n = col_thread_offset(p, 0, ct, tid) + ((uint64_t)u << (k - 4));
                        const uint64_t n0 = n - jr;
                        EXPECT(n0 % run == 0 && n0 + run <= K && (n0 >> rl) < (K >> rl), "run start k=%d", k);
                        for (uint32_t j = 0; j < run; ++j)
                            EXPECT(col_thread_offset(p, 0, ct, tid - jr + j) + ((uint64_t)u << (k - 4)) == n0 + j,
                                   "run lanes adjacent k=%d", k);
                    }
                }
            }
        }
    }
}

// Paired column tiles: the block's threads map one to one onto (tile,
// tile-local thread); inside each group of 2^(lc + pl) consecutive threads
// the columns of one row run consecutively across both tiles (tiles 2b and
// 2b + 1 of a slab are adjacent); a final-level run of 2^RL aligned lanes is
// one row of one tile.
static void check_col_pairs() {
    for (int lm = 5; lm <= kColMaxLog2; ++lm) {
        const int pl = col_pair_log2(lm), lc = kTileLog2 - lm;
        const uint32_t threads = 256u << pl;
        std::vector<int> seen(threads, 0);
        for (uint32_t t = 0; t < threads; ++t) {
            const uint32_t h = col_pair_half(t, lm), v = col_pair_vtid(t, lm);
            EXPECT(h < (1u << pl) && v < 256u, "pair range lm=%d t=%u", lm, t);
            if (h < (1u << pl) && v < 256u) seen[h * 256u + v]++;
            // absolute column of thread t within a slab (tiles of this block
            // start at column 0), and its row
            const uint32_t col = (h << lc) + (v & ((1u << lc) - 1u)), row = v >> lc;
            if ((t & ((1u << (lc + pl)) - 1u)) != 0) {
                const uint32_t hp = col_pair_half(t - 1, lm), vp = col_pair_vtid(t - 1, lm);
                const uint32_t colp = (hp << lc) + (vp & ((1u << lc) - 1u)), rowp = vp >> lc;
                EXPECT(row == rowp && col == colp + 1, "pair lanes adjacent lm=%d t=%u", lm, t);
            }
            const int rl = run_log2_for(lm);
            const uint32_t first = t & ~((1u << rl) - 1u);
            EXPECT(col_pair_half(first, lm) == h && (col_pair_vtid(first, lm) >> lc) == row,
                   "run inside one tile row lm=%d t=%u", lm, t);
        }
        for (uint32_t i = 0; i < threads; ++i) EXPECT(seen[i] == 1, "pair bijection lm=%d i=%u", lm, i);
    }
    // paired levels exist only where a product has an even number of tiles
    for (int k = kTileLog2 + 1; k <= kMaxLog2; ++k) EXPECT(tiles_per_product(make_plan(k)) % 2 == 0, "tiles even");
}

static void check_xcd_tile() {
    for (int c = 0; c <= kXcdChunkLog2Max; ++c)
        for (uint32_t n = 1; n < 5000; n += (n < 100 ? 1 : 37)) {
            const uint32_t chunk = 1u << c;
            std::vector<int> seen(n, 0);
            for (uint32_t b = 0; b < n; ++b) {
                const uint32_t t = xcd_tile(b, n, c);
                EXPECT(t < n, "xcd_tile range n=%u c=%d", n, c);
                if (t < n) seen[t]++;
            }
            for (uint32_t t = 0; t < n; ++t) EXPECT(seen[t] == 1, "xcd_tile bijection n=%u c=%d", n, c);
            // inside whole groups, each run of 2^c tiles runs on one XCD
            const uint32_t whole = (n / (kXcds * chunk)) * (kXcds * chunk);
            for (uint32_t b = 0; b < whole; ++b)
                EXPECT((xcd_tile(b, n, c) / chunk) % kXcds == b % kXcds, "xcd_tile chunk on one XCD n=%u c=%d", n, c);
            if (c == 0)
                for (uint32_t b = 0; b < n; ++b) EXPECT(xcd_tile(b, n, 0) == b, "chunk 1 is the plain order");
        }
}

int main() {
    check_tile_steps();
    check_col_pairs();
    check_xcd_tile();
    check_bitrev_split();
    // the built-in plans, then every split the run-time table accepts
    int splits = 0;
    for (int k = kMinLog2; k <= kMaxLog2; ++k) {
        const Plan d = make_plan(k);
        EXPECT(plan_ok(d), "default plan k=%d", k);
        check_plan(d, k <= 23);
        for (int m0 = 1; m0 <= 12; ++m0)
            for (int m1 = 0; m1 <= 12; ++m1) {
                const Plan p = make_plan_split(k, m0, m1);
                if (p.levels == 1 || (p.levels == 2 && m1 != 0) || !plan_ok(p)) continue;
                ++splits;
                check_plan(p, k <= 21);
            }
    }
    EXPECT(splits > 60, "split count %d", splits);
    printf("plans with explicit splits checked: %d\n", splits);
    printf(fails ? "index check: %d failures\n" : "index check: ok\n", fails);
    return fails ? 1 : 0;
}
