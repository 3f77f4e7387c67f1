// NTT multiply: the pass plan and every index computation, as plain
// host/device functions so the addressing can be checked on the CPU.
//
// Transform: length K = 2^k over two primes p0, p1 < 2^49 (2^30 | p - 1),
// residues held as binary64 in a centred range.  A length-K transform is a
// four-step decomposition into levels K = M_1 * M_2 * ... * M_L.  Level l
// works on contiguous slabs of K_l = M_l * ... * M_L elements; inside a slab
// element (r, c) sits at r * S_l + c with S_l = K_l / M_l, and the level does
// M_l-point DFTs down each column c.  A block holds a 4096-element tile: G
// DFTs of M points each (G * M = 4096).  Column levels tile C = 4096 / M_l
// adjacent columns (8..128 of them, runs of >= 64 bytes per prime plane);
// the last level is contiguous (S = 1) and tiles G whole slabs.
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
namespace ntt {

constexpr int kTileLog2 = 12;                 // 4096 elements per block tile
constexpr uint32_t kTile = 1u << kTileLog2;
constexpr int kColMaxLog2 = 9;                // column levels: M <= 512, C >= 8
constexpr int kMaxLevels = 3;
constexpr int kMaxLog2 = 30;
constexpr int kMinLog2 = 4;

// LDS slot of tile element i: one pad slot every 16 elements, so lanes
// reading 8-byte elements at a stride of 16 elements spread over all banks.
// Inside an in-tile step lds_index(base + j * stride) equals
// lds_index(base) + lds_index(j * stride), so the per-element LDS offsets of
// a small DFT are compile-time constants.
GPUPI_HD uint32_t lds_index(uint32_t i) { return i + (i >> 4); }
constexpr uint32_t kLdsSlots = kTile + (kTile >> 4);

struct Plan {
    int log2K;
    int levels;
    int log2M[kMaxLevels];   // per level, first level first
};

// k <= 12: one local level.  k <= 21: a column level and a contiguous
// level.  Otherwise two column levels and a contiguous level.  The column
// sizes are the fastest measured on MI355X for every k (batched products of
// 2^28 coefficients in all, every split the kernels cover, single products
// and the tree's shared-operand pairs agreeing): first levels of 2^7 or 2^8
// points (the sizes whose first forward level runs four waves per SIMD),
// and contiguous levels of 2^6..2^12.  k = 29, 30 (not measured) keep a
// 4096-point contiguous level.
GPUPI_HD Plan make_plan(int log2K) {
    //                          k: 13 14 15 16 17 18 19 20 21 22 23 24 25 26 27 28 29 30
    constexpr int8_t kM0[18] = {7, 7, 7, 8, 7, 7, 8, 8, 9, 7, 7, 8, 7, 7, 7, 8, 8, 9};
    constexpr int8_t kM1[18] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 7, 8, 8, 8, 8, 8, 8, 9, 9};
    Plan p{};
    p.log2K = log2K;
    if (log2K <= kTileLog2) {
        p.levels = 1;
        p.log2M[0] = log2K;
    } else if (log2K <= kTileLog2 + kColMaxLog2) {
        p.levels = 2;
        p.log2M[0] = kM0[log2K - 13];
        p.log2M[1] = log2K - p.log2M[0];
    } else {
        p.levels = 3;
        p.log2M[0] = kM0[log2K - 13];
        p.log2M[1] = kM1[log2K - 13];
        p.log2M[2] = log2K - p.log2M[0] - p.log2M[1];
    }
    return p;
}

// Level sizes the kernels are instantiated for: column levels 2^5..2^9,
// the contiguous level 2^6..2^12 (local products 2^4..2^12).
constexpr int kColMinLog2 = 5, kRowMinLog2 = 6;

// A plan with the same number of levels as make_plan(log2K) and explicit
// column sizes: log2M0 for the first level and, with three levels, log2M1
// for the second (the contiguous level takes the rest).  Zero keeps
// make_plan's choice.  Check the result with plan_ok.
GPUPI_HD Plan make_plan_split(int log2K, int log2M0, int log2M1) {
    Plan p = make_plan(log2K);
    if (p.levels == 1 || log2M0 == 0) return p;
    p.log2M[0] = log2M0;
    if (p.levels == 2) {
        p.log2M[1] = log2K - log2M0;
    } else {
        p.log2M[1] = log2M1 != 0 ? log2M1 : log2K - kTileLog2 - log2M0;
        p.log2M[2] = log2K - log2M0 - p.log2M[1];
    }
    return p;
}

// Whether every level of p has a kernel instantiation and the sizes add up.
GPUPI_HD bool plan_ok(const Plan &p) {
    if (p.log2K < kMinLog2 || p.log2K > kMaxLog2) return false;
    if (p.levels != make_plan(p.log2K).levels) return false;
    if (p.levels == 1) return p.log2M[0] == p.log2K;
    int sum = 0;
    for (int l = 0; l < p.levels; ++l) sum += p.log2M[l];
    if (sum != p.log2K) return false;
    for (int l = 0; l + 1 < p.levels; ++l)
        if (p.log2M[l] < kColMinLog2 || p.log2M[l] > kColMaxLog2) return false;
    const int last = p.log2M[p.levels - 1];
    return last >= kRowMinLog2 && last <= kTileLog2;
}

// log2 of the slab size K_l at level l
GPUPI_HD int slab_log2(const Plan &p, int level) {
    int s = 0;
    for (int l = level; l < p.levels; ++l) s += p.log2M[l];
    return s;
}

// Column level l: tile t (within one product's K elements) covers slab
// `slab`, columns [col0, col0 + C).  Element (g = column in tile, r) of the
// tile sits at offset(...) within the product.
struct ColTile {
    uint32_t slab;
    uint32_t col0;
};

GPUPI_HD ColTile col_tile(const Plan &p, int level, uint32_t t) {
    const int lm = p.log2M[level];
    const int lk = slab_log2(p, level);
    const int lc = kTileLog2 - lm;              // C = 2^lc columns per tile
    const int tiles_per_slab_log2 = (lk - lm) - lc;
    ColTile ct;
    ct.slab = t >> tiles_per_slab_log2;
    ct.col0 = (t & ((1u << tiles_per_slab_log2) - 1u)) << lc;
    return ct;
}

GPUPI_HD uint64_t col_offset(const Plan &p, int level, ColTile ct, uint32_t g, uint32_t r) {
    const int lm = p.log2M[level];
    const int lk = slab_log2(p, level);
    return ((uint64_t)ct.slab << lk) + ((uint64_t)r << (lk - lm)) + ct.col0 + g;
}

// Column tiles: thread `tid` of the 256 holds elements i = tid + 256 u
// (u < 16) of the tile, i.e. column g = tid mod C and rows
// r = (tid >> lc) + u 2^(lm-4); moving from u to u + 1 advances the element
// offset by 2^(lk-4) and adds bitrev(u, 4) to bitrev(r, lm).
GPUPI_HD uint64_t col_thread_offset(const Plan &p, int level, ColTile ct, uint32_t tid) {
    const int lc = kTileLog2 - p.log2M[level];
    return col_offset(p, level, ct, tid & ((1u << lc) - 1u), tid >> lc);
}
GPUPI_HD int col_row_step_log2(const Plan &p, int level) { return slab_log2(p, level) - 4; }

// Residue planes are padded: after every row of the first level's 2D view
// (2^(k - M0) elements, the rows a first-level column tile reads 2^M0 of)
// come kPlanePad unused elements, so those rows are not a power-of-two
// distance apart (at K = 2^28: 8 MB + 128 B instead of 8 MB).  Deeper levels
// and the contiguous level work inside one such row (their tiles and slabs
// divide it), so only the row's start moves.  Plans whose first-level rows
// are shorter than a tile (2^(k - M0) < 4096) are not padded.
constexpr uint32_t kPlanePad = 16;  // binary64 elements (128 bytes)
GPUPI_HD int plane_pad_shift(const Plan &p) {
    if (p.levels < 2) return 0;
    const int s = p.log2K - p.log2M[0];
    return s >= kTileLog2 ? s : 0;
}
// Physical index of logical element e of a product's plane.
GPUPI_HD uint64_t plane_index(const Plan &p, uint64_t e) {
    const int s = plane_pad_shift(p);
    return s ? e + (e >> s) * kPlanePad : e;
}
// Elements one product's plane takes (its row stride in a launch).
GPUPI_HD uint64_t plane_row_elems(const Plan &p) {
    const int s = plane_pad_shift(p);
    const uint64_t K = 1ull << p.log2K;
    return s ? K + (K >> s) * kPlanePad : K;
}

// Column levels of 2^7 or more points (32 or fewer adjacent columns per
// tile) run 2^col_pair_log2 neighbouring tiles per block of 256 *
// 2^col_pair_log2 threads.  Thread t works for tile col_pair_half(t) of the
// block as that tile's thread col_pair_vtid(t): within a group of 2^lc * 2^pl
// consecutive threads, the first 2^lc serve the first tile's row, the next
// 2^lc the second tile's (adjacent columns), and so on, so a wave's accesses
// cover 2^pl * C adjacent columns per row.
GPUPI_HD constexpr int col_pair_log2(int lm) { return lm >= 7 ? 1 : 0; }
GPUPI_HD constexpr uint32_t col_pair_half(uint32_t t, int lm) {
    return (t >> (kTileLog2 - lm)) & ((1u << col_pair_log2(lm)) - 1u);
}
GPUPI_HD constexpr uint32_t col_pair_vtid(uint32_t t, int lm) {
    return (t & ((1u << (kTileLog2 - lm)) - 1u)) | ((t >> (kTileLog2 - lm + col_pair_log2(lm))) << (kTileLog2 - lm));
}

GPUPI_HD uint32_t tiles_per_product(const Plan &p) { return 1u << (p.log2K - kTileLog2 > 0 ? p.log2K - kTileLog2 : 0); }

// Row (last) level: tile t holds G = 4096 / M consecutive slabs.
GPUPI_HD uint64_t row_offset(const Plan &p, uint32_t t, uint32_t g, uint32_t r) {
    const int lm = p.log2M[p.levels - 1];
    const int lg = kTileLog2 - lm;
    return ((uint64_t)((t << lg) + g) << lm) + r;
}

// Bit reversal of the low `bits` bits.
GPUPI_HD uint32_t bitrev(uint32_t x, int bits) {
    uint32_t r = 0;
    for (int i = 0; i < bits; ++i) {
        r = (r << 1) | (x & 1u);
        x >>= 1;
    }
    return r;
}

// Radix of step s for an M = 2^m point in-tile DFT: 16s, then the rest.
// Only the last step can have a smaller radix, and the last step's
// sub-problems have Ss = 1 (no twiddles), so every twiddled step is radix 16.
GPUPI_HD constexpr int step_count(int m) { return (m + 3) / 4; }
GPUPI_HD constexpr int step_log2radix(int m, int s) { return s < m / 4 ? 4 : (m - 4 * (m / 4)); }
// log2 of the sub-problem size Ms that step s works on (Ms = R * Ss)
GPUPI_HD constexpr int step_ms_log2(int m, int s) {
    int x = m;
    for (int i = 0; i < s; ++i) x -= step_log2radix(m, i);
    return x;
}

// Small DFT q (of 4096 / R per tile) at step s of G = 4096 / 2^m DFTs of
// M = 2^m points.  Element j of the small DFT is point r = b*Ms + j*Ss + n2
// of DFT g.  Two tile layouts: row tiles (contiguous levels) store point r of
// DFT g at tile index g*M + r and let n2 vary fastest with q; column tiles
// store it at r*G + g (the order of global memory) and let g vary fastest,
// so consecutive lanes touch consecutive tile indices in both layouts.
struct SmallDft {
    uint32_t base;   // tile index of element j = 0
    uint32_t stride; // tile-index step between elements j and j + 1
    uint32_t n2;     // position inside the current sub-problem (twiddle)
    int log2Ms;
};

GPUPI_HD SmallDft small_dft(int m, int s, uint32_t q, bool col) {
    const int log2Ms = step_ms_log2(m, s);
    const int lr = step_log2radix(m, s);
    const int log2Ss = log2Ms - lr;
    const int lg = kTileLog2 - m;  // log2 G
    SmallDft d;
    uint32_t g, b;
    if (col) {
        g = q & ((1u << lg) - 1u);
        const uint32_t rest = q >> lg;
        d.n2 = rest & ((1u << log2Ss) - 1u);
        b = rest >> log2Ss;
        d.base = (((b << log2Ms) + d.n2) << lg) + g;
        d.stride = 1u << (log2Ss + lg);
    } else {
        d.n2 = q & ((1u << log2Ss) - 1u);
        const uint32_t rest = q >> log2Ss;
        b = rest & ((1u << (m - log2Ms)) - 1u);
        g = rest >> (m - log2Ms);
        d.base = (g << m) + (b << log2Ms) + d.n2;
        d.stride = 1u << log2Ss;
    }
    d.log2Ms = log2Ms;
    return d;
}

// Last in-tile step of a column tile (DFT size 2^lm): radix R = 2^lr with
// lr = col_last_lr(lm), 16 / R small DFTs per thread.  Value k = u R + j of
// tile-local thread tid is point j of its small DFT u: row
// col_last_row0(lm, tid) + col_last_row_delta(lm, k) of column tid mod C.
// After the forward DFT that row holds frequency bitrev(row, lm), the
// multiplier of the inter-level twiddle, which splits as
// B0 + bitrev(u, 4 - lr) + bitrev(j, lr) 2^(lm - lr) with
// B0 = bitrev(tid >> lc, lm - 4) 2^(4 - lr).  Forward column levels store
// their results and inverse column levels load their inputs in this layout:
// for each k the lanes of a wave still cover the adjacent columns of a few
// rows, and no LDS exchange to or from the step-0 layout is needed.
GPUPI_HD constexpr int col_last_lr(int lm) { return step_log2radix(lm, step_count(lm) - 1); }
GPUPI_HD constexpr uint32_t col_last_row_delta(int lm, int k) {
    return ((uint32_t)(k >> col_last_lr(lm)) << (lm - 4 + col_last_lr(lm))) +
           ((uint32_t)k & ((1u << col_last_lr(lm)) - 1u));
}
GPUPI_HD uint32_t col_last_row0(int lm, uint32_t tid) { return (tid >> (kTileLog2 - lm)) << col_last_lr(lm); }
GPUPI_HD uint32_t col_last_b0(int lm, uint32_t tid) {
    return bitrev(tid >> (kTileLog2 - lm), lm - 4) << (4 - col_last_lr(lm));
}
GPUPI_HD uint32_t col_last_brow(int lm, uint32_t tid, int k) {
    const int lr = col_last_lr(lm);
    const uint32_t u = (uint32_t)k >> lr, j = (uint32_t)k & ((1u << lr) - 1u);
    return col_last_b0(lm, tid) + bitrev(u, 4 - lr) + (bitrev(j, lr) << (lm - lr));
}
// Plane positions between consecutive rows of a column tile at `level`
// (first-level rows carry the plane padding, deeper levels stay inside one
// padded row).
GPUPI_HD uint64_t col_plane_row_step(const Plan &p, int level) {
    const int rs = slab_log2(p, level) - p.log2M[level];
    const int s = plane_pad_shift(p);
    return (1ull << rs) + ((s && rs >= s) ? ((uint64_t)kPlanePad << (rs - s)) : 0ull);
}

// Workgroups are dealt to the 8 XCDs (each with its own L2) round-robin:
// block b runs on XCD b mod 8.  xcd_tile(b, n, c) renumbers blocks in
// groups of 8 * 2^c: within a group each XCD gets 2^c consecutive tiles, so
// neighbouring tiles (which share 128-byte lines of limbs) meet in one L2
// while the blocks resident at one time still cover one compact range of
// memory.  Blocks past the last whole group keep their number; c = 0 is the
// plain order.  A bijection on [0, n).  The chunk is a launch parameter
// (gpupi_ntt_set_launch); measured at 10^9 digits: one contiguous range per
// XCD 961.8 ms of kernels per step, chunks of 16 / 32 / 128 / 256 tiles
// 940.7 / 940.5 / 941.0 / 944.0 ms, the plain order 983.2 ms.
constexpr uint32_t kXcds = 8;
constexpr int kXcdChunkLog2 = 5;  // built-in chunk: 32 tiles
constexpr int kXcdChunkLog2Max = 8;
GPUPI_HD uint32_t xcd_tile(uint32_t b, uint32_t n, int chunk_log2) {
    const int gl = 3 + chunk_log2;  // log2 of the group: kXcds << chunk_log2
    if (b >= ((n >> gl) << gl)) return b;
    return ((b >> gl) << gl) + ((b & (kXcds - 1u)) << chunk_log2) + ((b >> 3) & ((1u << chunk_log2) - 1u));
}

// Run length of the final inverse column level (limbs per run): the C
// adjacent coefficients of one tile row, cut into runs of at most 16.  Thread
// tid of a first-level tile holds coefficients col_thread_offset(tid) +
// u 2^(k-4) (u < 16); the 2^run_log2 threads with equal tid >> run_log2 hold
// adjacent columns of one row, so each run lies in adjacent lanes of one wave.
GPUPI_HD constexpr int run_log2_for(int lm0) { return (kTileLog2 - lm0) < 4 ? (kTileLog2 - lm0) : 4; }
GPUPI_HD uint32_t final_run_log2(const Plan &p) {
    return p.levels == 1 ? 4u : (uint32_t)run_log2_for(p.log2M[0]);
}

}  // namespace ntt
}  // namespace gpupi
