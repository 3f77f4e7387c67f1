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
// adjacent columns (>= 8, i.e. >= 128-byte runs of double2); the last level
// is contiguous (S = 1) and tiles G whole slabs.
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

// LDS index of tile element i (one pad slot every 32 elements).
GPUPI_HD uint32_t lds_index(uint32_t i) { return i + (i >> 5); }
constexpr uint32_t kLdsSlots = kTile + (kTile >> 5);

struct Plan {
    int log2K;
    int levels;
    int log2M[kMaxLevels];   // per level, first level first
};

// k <= 12: one local level.  k <= 21: column level of ceil(k/2) (<= 9) and
// a contiguous level.  Otherwise two column levels and a 4096-point
// contiguous level.
GPUPI_HD Plan make_plan(int log2K) {
    Plan p{};
    p.log2K = log2K;
    if (log2K <= kTileLog2) {
        p.levels = 1;
        p.log2M[0] = log2K;
    } else if (log2K <= kTileLog2 + kColMaxLog2) {
        int c = (log2K + 1) / 2;
        if (c > kColMaxLog2) c = kColMaxLog2;
        if (log2K - c > kTileLog2) c = log2K - kTileLog2;
        p.levels = 2;
        p.log2M[0] = c;
        p.log2M[1] = log2K - c;
    } else {
        const int rest = log2K - kTileLog2;  // 10..18
        p.levels = 3;
        p.log2M[0] = (rest + 1) / 2;
        p.log2M[1] = rest - p.log2M[0];
        p.log2M[2] = kTileLog2;
    }
    return p;
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
GPUPI_HD int step_count(int m) { return (m + 3) / 4; }
GPUPI_HD int step_log2radix(int m, int s) {
    const int full = m / 4;
    return s < full ? 4 : (m - 4 * full);
}

// Small DFT q (of 4096 / R per tile) at step s: element j sits at tile index
// g*M + b*Ms + j*Ss + n2.  n2 varies fastest with q.
struct SmallDft {
    uint32_t base;   // tile index of element j = 0
    uint32_t stride; // Ss
    uint32_t n2;     // position inside the current sub-problem (twiddle)
    int log2Ms;
};

GPUPI_HD SmallDft small_dft(int m, int s, uint32_t q) {
    int log2Ms = m;
    for (int i = 0; i < s; ++i) log2Ms -= step_log2radix(m, i);
    const int lr = step_log2radix(m, s);
    const int log2Ss = log2Ms - lr;
    SmallDft d;
    d.n2 = q & ((1u << log2Ss) - 1u);
    const uint32_t rest = q >> log2Ss;
    const uint32_t b = rest & ((1u << (m - log2Ms)) - 1u);
    const uint32_t g = rest >> (m - log2Ms);
    d.base = (g << m) + (b << log2Ms) + d.n2;
    d.stride = 1u << log2Ss;
    d.log2Ms = log2Ms;
    return d;
}

// Run length of the final inverse column level (limbs per run).
GPUPI_HD uint32_t final_run_log2(const Plan &p) {
    return p.levels == 1 ? 4 : (uint32_t)(kTileLog2 - p.log2M[0]);
}

}  // namespace ntt
}  // namespace gpupi
