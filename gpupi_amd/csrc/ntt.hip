// Batched exact big-integer product by NTT (replaces the reference's
// batch_mul_ntt.cu, which works over 2^96 - 2^32 + 1 with 32-bit integer
// multiplies).
//
// MI355X design: CDNA4 issues FP64 FMA at full rate while 32-bit integer
// multiplies are quarter rate, so residues live in binary64 and each modular
// product is six FP64 operations (Dekker/FMA two-product + rounded quotient):
//     h = a*b; l = fma(a,b,-h); q = rint(h/p); r = fma(-q,p,h) + l.
// Two primes p0, p1 < 2^49 with 2^30 | p - 1 give exact coefficients up to
// p0 p1 ~ 2^98 (a coefficient of a product of 32-bit limbs is < 2^91 for
// lengths up to 2^27), so each limb is one coefficient and transforms reach
// 2^30 points.  Values are kept centred (|v| <= 0.75 p) so no compare/select
// is needed inside butterflies; sums are re-centred every other stage.
//
// Residues are stored planar (one array of binary64 per prime) and every
// kernel runs the two primes one after the other through a 4096-element LDS
// tile (34.8 KB).  Inside a tile the transform runs radix-16 steps
// (bit-reversed slot order); the DFT size of a level is a template parameter
// so every LDS address is a per-thread base plus a compile-time offset.
// Long transforms are split into column levels (<= 512 points, 8..128
// adjacent columns per tile) and a final contiguous level.  The last forward
// level, the pointwise product and the first inverse level are one kernel;
// the final inverse level applies 1/K, recombines the primes (Garner) and
// emits limbs with carries resolved inside runs of <= 16 adjacent
// coefficients, leaving one 64-bit spill word per run for the carry pass.
//
// Twiddles: in-tile steps read a per-step table laid out [j][n2] so that a
// wave's loads are contiguous; the inter-level twiddle of a column element
// is w^(c * bitrev(r)), generated per thread as base * step^v from two
// lookups in a two-level table of w_{2^30} powers.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdio>
#include <mutex>
#include <type_traits>
#include <vector>

#include "common.hpp"
#include "gpupi_pi.h"
#include "ntt_arith.hpp"
#include "ntt_plan.hpp"

namespace gpupi {
namespace {

using namespace ntt;

constexpr uint64_t kPrime0 = 0x1ffffc0000001ull;  // 562948879679489 = 524287 * 2^30 + 1
constexpr uint64_t kPrime1 = 0x1fffe00000001ull;  // 562941363486721 = 524280 * 2^30 + 1 (2^33 | p - 1)
constexpr uint64_t kGen0 = 3, kGen1 = 13;          // primitive roots

// Table layout (binary64) inside tables.roots, one block per prime.
//   step tables: for Ms = 2^x (x = 5..12), at (2^x - 32): [j * Ms/16 + n2] =
//                w_Ms^(n2 * bitrev(j, 4)), j < 16, n2 < Ms/16
//   lo/hi: w_{2^30}^e and w_{2^30}^(e 2^15), e < 2^15 (and the inverses)
constexpr size_t kOffStep = 0;
constexpr size_t kOffStepInv = 8192;
constexpr size_t kOffLo = 16384;
constexpr size_t kOffHi = kOffLo + 32768;
constexpr size_t kOffLoInv = kOffHi + 32768;
constexpr size_t kOffHiInv = kOffLoInv + 32768;
constexpr size_t kPrimeTable = kOffHiInv + 32768;
constexpr size_t kTableLen = 2 * kPrimeTable;

struct NttConsts {
    double w16[2][8];
    double w16i[2][8];
    double kinv[2][kMaxLog2 + 1];
    double garner;  // p0^-1 mod p1, centred
};
__constant__ NttConsts c_ntt;

constexpr uint32_t brev4(uint32_t x) { return ((x & 1u) << 3) | ((x & 2u) << 1) | ((x & 4u) >> 1) | ((x & 8u) >> 3); }

// In-tile steps.  Step S of the 2^M-point DFTs of a tile is 4096 / R small
// DFTs of radix R; thread q = threadIdx.x handles 16 / R of them (q, q + 256,
// ...), 16 values in registers v[u * R + j].  The first and last steps of a
// transform run on values already in registers (the step-0 elements of a
// thread are the ones it reads from HBM), the others go through LDS.
template <int M, int S, bool COL>
struct StepMap {
    static constexpr int LR = step_log2radix(M, S);
    static constexpr int R = 1 << LR;
    static constexpr int PER = 16 / R;
    static constexpr int LMS = step_ms_log2(M, S);
    static constexpr int LSS = LMS - LR;
    static constexpr uint32_t kStride = COL ? 1u << (LSS + kTileLog2 - M) : 1u << LSS;
    uint32_t base[PER];  // tile index of element j = 0 of each small DFT
    uint32_t n2[PER];
    // tid: the thread's index within its tile's 256 (threadIdx.x, or its
    // virtual index in a block of paired column tiles)
    __device__ __forceinline__ explicit StepMap(uint32_t tid) {
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const SmallDft d = small_dft(M, S, tid + 256u * u, COL);
            base[u] = d.base;
            n2[u] = d.n2;
        }
    }
    // tile index of value v[k]
    __device__ __forceinline__ uint32_t index(int k) const { return base[k / R] + (k % R) * kStride; }
};

// Butterfly kind of forward step S (FINAL: the kind of the last step) and
// of inverse step S.
template <int M, int S, int FINAL>
constexpr int fwd_kind() { return S < step_count(M) - 1 ? kDifMid : FINAL; }
template <int M, int S>
constexpr int inv_kind() {
    return S == step_count(M) - 1 ? (S == 0 ? kDitOnly : kDitFirst) : (S == 0 ? kDitLast : kDitMid);
}

// Butterflies (and the step's twiddles) of step S on registers; KIND picks
// the re-centring plan.
template <int M, int S, bool INV, bool COL, int KIND>
__device__ __forceinline__ void step_regs(double *v, const StepMap<M, S, COL> &sm, const double *tab, int P, Mod m) {
    using SM = StepMap<M, S, COL>;
    constexpr int R = SM::R;
    const double *w16 = INV ? c_ntt.w16i[P] : c_ntt.w16[P];
#pragma unroll
    for (int u = 0; u < SM::PER; ++u) {
        if (!INV) dif<KIND, R>(&v[u * R], w16, m);
        if constexpr (SM::LSS > 0) {
            static_assert(R == 16, "twiddled steps are radix 16");
            const double *tw = tab + (INV ? kOffStepInv : kOffStep) + ((1u << SM::LMS) - 32u) + sm.n2[u];
#pragma unroll
            for (int j = 1; j < 16; ++j) v[u * 16 + j] = mulp(v[u * 16 + j], tw[j << SM::LSS], m);
        }
        if (INV) dit<KIND, R>(&v[u * R], w16, m);
    }
}

template <int M, int S, bool COL>
__device__ __forceinline__ void lds_load(double *v, const double *lds, const StepMap<M, S, COL> &sm) {
    using SM = StepMap<M, S, COL>;
#pragma unroll
    for (int u = 0; u < SM::PER; ++u) {
        const uint32_t b = lds_index(sm.base[u]);
#pragma unroll
        for (int j = 0; j < SM::R; ++j) v[u * SM::R + j] = lds[b + lds_index(j * SM::kStride)];
    }
}

template <int M, int S, bool COL>
__device__ __forceinline__ void lds_store(double *lds, const double *v, const StepMap<M, S, COL> &sm) {
    using SM = StepMap<M, S, COL>;
#pragma unroll
    for (int u = 0; u < SM::PER; ++u) {
        const uint32_t b = lds_index(sm.base[u]);
#pragma unroll
        for (int j = 0; j < SM::R; ++j) lds[b + lds_index(j * SM::kStride)] = v[u * SM::R + j];
    }
}

// Steps S..LAST (forward, increasing) or S..LAST (inverse, decreasing) as
// LDS -> registers -> LDS passes, each followed by a barrier.
template <int M, bool INV, bool COL, int S, int LAST>
__device__ __forceinline__ void lds_steps(double *lds, const double *tab, int P, Mod m, uint32_t tid) {
    if constexpr (INV ? S >= LAST : S <= LAST) {
        const StepMap<M, S, COL> sm(tid);
        double v[16];
        lds_load(v, lds, sm);
        step_regs<M, S, INV, COL, INV ? inv_kind<M, S>() : fwd_kind<M, S, kDifMid>()>(v, sm, tab, P, m);
        lds_store(lds, v, sm);
        __syncthreads();
        lds_steps<M, INV, COL, INV ? S - 1 : S + 1, LAST>(lds, tab, P, m, tid);
    }
}

// Forward transform of the tile.  On entry v holds this thread's step-0
// elements; on exit its last-step elements (bit-reversed DFT order).  LDS is
// used when M > 4: the caller must make sure no thread still reads `lds`.
template <int M, bool COL, int FINAL>
__device__ __forceinline__ void fwd_dft(double *v, double *lds, const double *tab, int P, Mod m,
                                        uint32_t tid = threadIdx.x) {
    constexpr int kLast = step_count(M) - 1;
    step_regs<M, 0, false, COL, fwd_kind<M, 0, FINAL>()>(v, StepMap<M, 0, COL>(tid), tab, P, m);
    if constexpr (kLast > 0) {
        lds_store(lds, v, StepMap<M, 0, COL>(tid));
        __syncthreads();
        lds_steps<M, false, COL, 1, kLast - 1>(lds, tab, P, m, tid);
        const StepMap<M, kLast, COL> sl(tid);
        lds_load(v, lds, sl);
        step_regs<M, kLast, false, COL, FINAL>(v, sl, tab, P, m);
    }
}

// Inverse transform of the tile.  On entry v holds this thread's last-step
// elements (as left by fwd_dft, so the first LDS store touches only slots
// this thread read); on exit its step-0 elements.
template <int M, bool COL, int FINAL = kDitLast>
__device__ __forceinline__ void inv_dft(double *v, double *lds, const double *tab, int P, Mod m,
                                        uint32_t tid = threadIdx.x) {
    constexpr int kLast = step_count(M) - 1;
    const StepMap<M, kLast, COL> sl(tid);
    step_regs<M, kLast, true, COL, inv_kind<M, kLast>()>(v, sl, tab, P, m);
    if constexpr (kLast > 0) {
        lds_store(lds, v, sl);
        __syncthreads();
        lds_steps<M, true, COL, kLast - 1, 1>(lds, tab, P, m, tid);
        const StepMap<M, 0, COL> s0(tid);
        lds_load(v, lds, s0);
        step_regs<M, 0, true, COL, FINAL>(v, s0, tab, P, m);
    }
}

constexpr uint32_t brev_c(uint32_t x, int bits) {
    uint32_t r = 0;
    for (int i = 0; i < bits; ++i) r |= ((x >> i) & 1u) << (bits - 1 - i);
    return r;
}

// Column twiddles are applied in the last in-tile step's layout: value
// k = u R + j (R = 2^LR the last step's radix) takes w^(cg E) with
// E = B0 + bitrev(u, 4 - LR) + bitrev(j, LR) 2^(M - LR) (cg: the column's
// exponent, B0: the thread's row base).  Table entries, as lo/hi pairs of
// the two-level table (w_{2^30}^E = lo[E mod 2^15] * hi[E >> 15],
// E < 2^30), are loaded early (they depend only on the tile geometry) so
// their latency overlaps the data loads: w^(cg B0) and the step
// w^(cg 2^(M - LR)) between j's when a thread holds one small DFT (M = 8),
// else w^cg, the step between u's.
struct ColTw {
    double lb, hb, ls, hs;
};
template <int M>
__device__ __forceinline__ ColTw col_tw_load(const double *tab, uint32_t cg, uint32_t B0, bool inv) {
    constexpr int LR = col_last_lr(M);
    const double *lo = tab + (inv ? kOffLoInv : kOffLo), *hi = tab + (inv ? kOffHiInv : kOffHi);
    const uint32_t eb = cg * B0;
    const uint32_t es = (16 >> LR) == 1 ? cg << (M - LR) : cg;
    return ColTw{lo[eb & 32767u], hi[eb >> 15], lo[es & 32767u], hi[es >> 15]};
}

// Multiplies value k of x by its column twiddle (layout above): chains
// base * su^a * sj^b over a = bitrev(u), b = bitrev(j); two interleaved
// chains when the thread holds one small DFT of 16.
template <int M>
__device__ __forceinline__ void column_twiddles(double *x, const ColTw &tw, Mod m) {
    constexpr int LR = col_last_lr(M), R = 1 << LR, PER = 16 / R;
    const double base = mulp(tw.lb, tw.hb, m);
    const double step = mulp(tw.ls, tw.hs, m);
    if constexpr (PER == 1) {
        const double step2 = mulp(step, step, m);
        double c0 = base, c1 = mulp(base, step, m);
#pragma unroll
        for (int b = 0; b < 16; b += 2) {
            x[brev4(b)] = mulp(x[brev4(b)], c0, m);
            x[brev4(b + 1)] = mulp(x[brev4(b + 1)], c1, m);
            if (b + 2 < 16) {
                c0 = mulp(c0, step2, m);
                c1 = mulp(c1, step2, m);
            }
        }
    } else {
        double sj = step;  // w^(cg 2^(M - LR))
#pragma unroll
        for (int i = 0; i < M - LR; ++i) sj = mulp(sj, sj, m);
        double cu = base;
#pragma unroll
        for (int a = 0; a < PER; ++a) {
            const int u = (int)brev_c((uint32_t)a, 4 - LR);
            double c = cu;
#pragma unroll
            for (int b = 0; b < R; ++b) {
                const int k = u * R + (int)brev_c((uint32_t)b, LR);
                x[k] = mulp(x[k], c, m);
                if (b + 1 < R) c = mulp(c, sj, m);
            }
            if (a + 1 < PER) cu = mulp(cu, step, m);
        }
    }
}

// An integral double in [0, 2^52) as an integer: adding 2^52 puts its value
// in the mantissa bits (exact), which one integer subtraction reads out.
__device__ __forceinline__ uint64_t small_double_to_u64(double x) {
    return (uint64_t)__double_as_longlong(x + 0x1p52) - (uint64_t)__double_as_longlong(0x1p52);
}

// Garner recombination of centred residues -> exact coefficient (< 2^98):
// r0 + p0 t with t = (r1 - r0) / p0 mod p1; p0 t is formed from shifts, since
// p0 = 2^49 - 2^30 + 1.
__device__ __forceinline__ unsigned __int128 recombine(double v0, double v1) {
    static_assert(kPrime0 == (1ull << 49) - (1ull << 30) + 1ull, "p0 = 2^49 - 2^30 + 1");
    const Mod m1 = mod_of(1);
    const double r0 = v0 < 0.0 ? v0 + kP0 : v0;
    const double r1 = v1 < 0.0 ? v1 + kP1 : v1;
    double t = mulp(r1 - r0, c_ntt.garner, m1);
    t = t < 0.0 ? t + kP1 : t;
    const unsigned __int128 T = small_double_to_u64(t);
    return (T << 49) - (T << 30) + T + small_double_to_u64(r0);
}

// Forward operand o of a multi-level launch: limbs of row i at X + i s (the
// first level reads them), residues of prime P of row i at
// T + base + P pstride + i K.
struct NttOp {
    const uint32_t *X;
    uint64_t base, pstride;
    uint32_t L, s;
};

// Product j of a multi-level launch: operand a times operand b (row 0 of b
// for every row when b_bcast); the product overwrites a's residues, then
// goes to C + i sC (Lc limbs) with its spill words from spill + spill_base.
struct NttProd {
    uint32_t *C;
    uint64_t spill_base;
    uint32_t Lc, sC;
    uint32_t a, b, b_bcast;
};

struct PassArgs {
    // one-tile products: A, B, C with strides, N products
    const uint32_t *A;
    const uint32_t *B;
    uint32_t La, Lb, sA, sB;
    uint32_t *C;
    uint32_t Lc, sC;
    uint32_t N;
    // multi-level products: residue planes, spill words, operands and
    // products; forward launches cover operand o's rows from block
    // op_start[o], the others product j's rows from block prod_start[j]
    double *T;
    uint64_t *spill;
    NttOp ops[3];
    NttProd prods[2];
    uint32_t op_start[3];
    uint32_t prod_start[2];
    uint32_t n_ops, n_prods;
    Plan plan;
    const double *tables;
    // ragged launches: per-product rows and the set of B operands they use
    // (device arrays), or null when the operands above describe every product
    const RaggedRow *rows;
    const RaggedB *bset;
    // log2 of the tiles each XCD takes in turn (xcd_tile)
    int xcd_log2;
};

// Global memory instructions between memory_first_begin() and
// memory_first_end() issue at raised wave priority: a wave about to wait on
// memory gets its requests out ahead of the waves that have arithmetic
// ready, so that the latency overlaps their work (the accesses are issued,
// not waited for, in between).  Used where it measured faster: the row
// levels' and the first forward level's loads (at 10^9 digits about 1 %
// each), and the first forward level's stores (1.4 % more); around the
// second column levels' and the first inverse level's loads it was slower
// (first inverse 123 -> 137 ms), around the row levels' stores neutral.
__device__ __forceinline__ void memory_first_begin() { __builtin_amdgcn_s_setprio(1); }
__device__ __forceinline__ void memory_first_end() { __builtin_amdgcn_s_setprio(0); }

// The (segment, row, tile) a block works on: segments are operands (forward
// launches) or products (the others), each `rows x tiles` blocks in turn.
struct BlockPos {
    uint32_t seg, row, t;
};
__device__ __forceinline__ BlockPos block_pos_of(uint32_t bid, const uint32_t *start, uint32_t nseg, uint32_t tiles) {
    uint32_t seg = 0, first = 0;
    for (uint32_t i = 1; i < nseg; ++i)
        if (bid >= start[i]) {
            seg = i;
            first = start[i];
        }
    const uint32_t local = bid - first;
    return BlockPos{seg, local / tiles, local % tiles};
}
__device__ __forceinline__ BlockPos block_pos(const uint32_t *start, uint32_t nseg, uint32_t tiles, int xcd_log2) {
    return block_pos_of(xcd_tile(blockIdx.x, gridDim.x, xcd_log2), start, nseg, tiles);
}

// Column kernels of DFT size 2^M >= 2^7 (tiles of 32 or fewer adjacent
// columns) can run two neighbouring tiles per block of 512 threads, so that
// a wave's global accesses cover the 2C adjacent columns of both tiles (runs
// of 2C elements instead of C): thread t works for tile half(t) as virtual
// thread col_vtid(t), and each half has its own LDS tile.  On the first
// level's access pattern at K = 2^28 a bare copy kernel moves limbs -> two
// binary64 planes at 3.1 -> 4.1 TB/s and binary64 -> binary64 at 3.9 -> 5.2
// TB/s with runs doubled, but in the NTT kernels (per 10^9 step) only the
// first forward level gains (195 -> 191 ms) and the second inverse level a
// little (64.4 -> 63.0 ms); the second forward level loses 2 ms and the
// first inverse level 13 ms (barriers over eight waves instead of four), so
// by default only those two pair their tiles; with the last-step layout
// (after the measurements above) the choice still holds: first inverse
// level paired +11 ms per 10^9 step, the others within 2 ms.  The choice is
// a launch parameter per level kind (gpupi_ntt_set_launch, the tuner's
// --launch): bit 0 first forward level, bit 1 second forward level, bit 2
// second inverse level, bit 3 first inverse level.
constexpr uint32_t kPairMaskDefault = 5;
constexpr int kPairFwdLimbs = 0, kPairFwd = 1, kPairInv = 2, kPairInvLimbs = 3;

template <int M, bool PAIR>
struct ColPair {
    static constexpr int kLog2 = PAIR ? col_pair_log2(M) : 0;
    static constexpr int kTiles = 1 << kLog2;
    static constexpr int kThreads = 256 * kTiles;
    static constexpr int LC = kTileLog2 - M;
    // the tiles of a block start 16 slots (32 four-byte banks) off a
    // multiple of the bank count, so that a wave's two tile halves, which
    // address the same slots of their tiles, fall on different banks (the
    // paired second inverse level: 25 % -> 0 % of LDS cycles in bank
    // conflicts, its time unchanged)
    static constexpr uint32_t kTileStride = kLdsSlots + 16;
    static constexpr uint32_t kLdsWords = kTiles * kTileStride - 16;
    __device__ __forceinline__ static uint32_t half() { return kLog2 ? col_pair_half(threadIdx.x, M) : 0u; }
    __device__ __forceinline__ static uint32_t vtid() { return kLog2 ? col_pair_vtid(threadIdx.x, M) : threadIdx.x; }
};

// Per-thread geometry of a column tile (level `level`, DFT size 2^M): the
// thread (tile-local index tid) holds elements tid + 256 u (u < 16), its
// step-0 elements, at the start of a forward level and the end of an
// inverse one, and its last-step values k (rows col_last_row0 +
// col_last_row_delta(k) of its column) at the end of a forward level and
// the start of an inverse one.
template <int M>
struct ColThread {
    uint64_t off0;     // element (coefficient) index of u = 0 within the product
    uint32_t ustep;    // log2 of the index step between u and u + 1
    uint64_t poff0;    // plane position of u = 0 (padded planes)
    uint64_t pstep;    // plane position step between u and u + 1
    uint64_t plast0;   // plane position of last-step value k = 0
    uint64_t prow;     // plane position step between rows of the tile
    uint32_t cg;       // (col0 + g) << (30 - lk): twiddle exponent per bitrev(r)
    uint32_t B0;       // bitrev of the last step's row base (col_last_b0)
    uint32_t lds0;     // LDS slot of u = 0
    uint32_t tid;      // index within the tile's 256 threads

    __device__ __forceinline__ ColThread(const Plan &p, int level, uint32_t t, uint32_t tid_) {
        constexpr int LC = kTileLog2 - M;
        tid = tid_;
        const ColTile ct = col_tile(p, level, t);
        const uint32_t g = tid & ((1u << LC) - 1u);
        off0 = col_thread_offset(p, level, ct, tid);
        ustep = (uint32_t)col_row_step_log2(p, level);
        poff0 = plane_index(p, off0);
        // a step of 2^ustep elements crosses 2^(ustep - s) padded rows when
        // it is at least one row long, and none otherwise
        const int s = plane_pad_shift(p);
        pstep = (1ull << ustep) + ((s && (int)ustep >= s) ? ((uint64_t)kPlanePad << (ustep - s)) : 0ull);
        plast0 = plane_index(p, col_offset(p, level, ct, g, col_last_row0(M, tid)));
        prow = col_plane_row_step(p, level);
        cg = (ct.col0 + g) << (kMaxLog2 - slab_log2(p, level));
        B0 = (__brev(tid >> LC) >> (32 - (M - 4))) << (4 - col_last_lr(M));
        lds0 = lds_index(tid);
    }
    // coefficient index of element u (limbs)
    __device__ __forceinline__ uint64_t off(int u) const { return off0 + ((uint64_t)u << ustep); }
    // plane position of element u (residues)
    __device__ __forceinline__ uint64_t poff(int u) const { return poff0 + (uint64_t)u * pstep; }
    // plane position of last-step value k
    __device__ __forceinline__ uint64_t plast(int k) const {
        return plast0 + (uint64_t)col_last_row_delta(M, k) * prow;
    }
};

// One prime of a forward column level on the thread's step-0 elements v:
// DFT down the columns, then the inter-level twiddle; v is left holding the
// thread's last-step values (stored at ct.plast(k)).  With BARRIER it ends
// with a barrier, so the next prime may reuse `lds`.
template <int M, int P, bool BARRIER>
__device__ __forceinline__ void col_fwd_prime(double *v, double *lds, const ColThread<M> &ct, const double *tables,
                                              const ColTw &tw) {
    const Mod m = mod_of(P);
    const double *tab = tables + P * kPrimeTable;
    fwd_dft<M, true, kDifToTwiddle>(v, lds, tab, P, m, ct.tid);
    column_twiddles<M>(v, tw, m);
    if (BARRIER) __syncthreads();
}

// col_fwd_prime with the column twiddles loaded after the transform (fewer
// registers held through it), then a barrier.
template <int M, int P>
__device__ __forceinline__ void col_fwd_prime_late(double *v, double *lds, const ColThread<M> &ct,
                                                   const double *tables) {
    const Mod m = mod_of(P);
    const double *tab = tables + P * kPrimeTable;
    fwd_dft<M, true, kDifToTwiddle>(v, lds, tab, P, m, ct.tid);
    column_twiddles<M>(v, col_tw_load<M>(tab, ct.cg, ct.B0, false), m);
    __syncthreads();
}

// Forward column level above the first: blocks cover (product, operand,
// tile); both primes pass through the tile in turn; the inter-level twiddle
// is applied on store.  Every global load (both primes' residues and column
// twiddles) is issued before the first store and both primes' results are
// stored at the end: loads and stores share one in-order counter, so a
// load issued after a store also waits for the store.
template <int M, bool PAIR>
__global__ void __launch_bounds__((ColPair<M, PAIR>::kThreads)) ntt_col_fwd_kernel(PassArgs a, int level) {
    using CP = ColPair<M, PAIR>;
    __shared__ double lds_tiles[CP::kLdsWords];
    double *const lds = lds_tiles + CP::half() * CP::kTileStride;
    const BlockPos bp = block_pos(a.op_start, a.n_ops, tiles_per_product(a.plan) >> CP::kLog2, a.xcd_log2);
    const NttOp o = a.ops[bp.seg];
    const ColThread<M> ct(a.plan, level, (bp.t << CP::kLog2) + CP::half(), CP::vtid());
    double *T0 = a.T + o.base + (size_t)bp.row * plane_row_elems(a.plan);
    double *T1 = T0 + o.pstride;
    double v0[16], v1[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) v0[u] = T0[ct.poff(u)];
#pragma unroll
    for (int u = 0; u < 16; ++u) v1[u] = T1[ct.poff(u)];
    const ColTw tw0 = col_tw_load<M>(a.tables, ct.cg, ct.B0, false);
    const ColTw tw1 = col_tw_load<M>(a.tables + kPrimeTable, ct.cg, ct.B0, false);
    col_fwd_prime<M, 0, true>(v0, lds, ct, a.tables, tw0);
    col_fwd_prime<M, 1, false>(v1, lds, ct, a.tables, tw1);
#pragma unroll
    for (int k = 0; k < 16; ++k) T0[ct.plast(k)] = v0[k];
#pragma unroll
    for (int k = 0; k < 16; ++k) T1[ct.plast(k)] = v1[k];
}

// Limbs of one tile block for the first forward level: element u of the
// thread's column (zero past the operand's length).
template <int M>
__device__ __forceinline__ void load_limbs(uint32_t *x, const PassArgs &a, const BlockPos &bp, const ColThread<M> &ct) {
    const NttOp o = a.ops[bp.seg];
    const uint32_t *X = o.X + (size_t)bp.row * o.s;
    uint32_t LX = o.L;
    if (a.rows) {
        if (bp.seg) {
            const RaggedB rb = a.bset[bp.row];
            X = rb.B;
            LX = (uint32_t)rb.L;
        } else {
            const RaggedRow rr = a.rows[bp.row];
            X = rr.A;
            LX = rr.L_a;
        }
    }
    memory_first_begin();
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const uint64_t off = ct.off(u);
        x[u] = off < LX ? X[off] : 0u;
    }
    memory_first_end();
}

// First forward column level, persistent: block b works through the tile
// blocks w = b, b + gridDim.x, ... (numbered and XCD-ordered as one block per
// tile block would be; gridDim.x is a multiple of the XCD count).  The limbs
// stay in registers as 32-bit words and each prime converts them; once the
// second prime has, the next tile block's limbs are loaded into the same
// registers, so that their latency (and the completion of the first
// prime's stores before them) overlaps the second prime's transform and
// stores.  Each prime's column twiddles are loaded after its transform (not
// held through it), which brings the 2^6..2^8-point levels to 121-125 VGPRs:
// four waves per SIMD without spills (at 10^9 digits 179 -> 165 ms per
// step, against three waves).  The 2^5 and 2^9-point levels would spill
// at four and keep three.
constexpr int limbs_waves(int m) { return m >= 6 && m <= 8 ? 4 : 3; }
template <int M, bool PAIR>
__global__ void __launch_bounds__((ColPair<M, PAIR>::kThreads)) __attribute__((amdgpu_waves_per_eu(limbs_waves(M))))
ntt_col_fwd_limbs_kernel(PassArgs a, uint32_t total) {
    using CP = ColPair<M, PAIR>;
    __shared__ double lds_tiles[CP::kLdsWords];
    double *const lds = lds_tiles + CP::half() * CP::kTileStride;
    const uint32_t tiles = tiles_per_product(a.plan) >> CP::kLog2;
    uint32_t w = blockIdx.x;
    if (w >= total) return;  // uniform over the block
    BlockPos bp = block_pos_of(xcd_tile(w, total, a.xcd_log2), a.op_start, a.n_ops, tiles);
    uint32_t x[16];
    load_limbs<M>(x, a, bp, ColThread<M>(a.plan, 0, (bp.t << CP::kLog2) + CP::half(), CP::vtid()));
    for (;;) {
        const ColThread<M> ct(a.plan, 0, (bp.t << CP::kLog2) + CP::half(), CP::vtid());
        const NttOp o = a.ops[bp.seg];
        double *T0 = a.T + o.base + (size_t)bp.row * plane_row_elems(a.plan);
        double *T1 = T0 + o.pstride;
        double v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) v[u] = (double)x[u];
        col_fwd_prime_late<M, 0>(v, lds, ct, a.tables);
        memory_first_begin();
#pragma unroll
        for (int k = 0; k < 16; ++k) T0[ct.plast(k)] = v[k];
        memory_first_end();
#pragma unroll
        for (int u = 0; u < 16; ++u) v[u] = (double)x[u];
        const uint32_t wn = w + gridDim.x;
        const bool more = wn < total;  // uniform over the block
        if (more) {
            bp = block_pos_of(xcd_tile(wn, total, a.xcd_log2), a.op_start, a.n_ops, tiles);
            load_limbs<M>(x, a, bp, ColThread<M>(a.plan, 0, (bp.t << CP::kLog2) + CP::half(), CP::vtid()));
        }
        // the barrier at its end lets the next tile block reuse `lds`
        col_fwd_prime_late<M, 1>(v, lds, ct, a.tables);
        memory_first_begin();
#pragma unroll
        for (int k = 0; k < 16; ++k) T1[ct.plast(k)] = v[k];
        memory_first_end();
        if (!more) break;
        w = wn;
    }
}

// Step-0 elements of a row tile (DFT size 2^M): v[j] is tile index
// row0 + j 2^(M-4).
template <int M>
__device__ __forceinline__ uint32_t row_step0_base() {
    return ((threadIdx.x >> (M - 4)) << M) + (threadIdx.x & ((1u << (M - 4)) - 1u));
}

// Last level: per prime, forward on A's tile (kept in registers), forward on
// B's tile, pointwise product, inverse; written back over A's plane.  A
// square (B is A) multiplies A's transformed tile by itself.  The next
// prime's A tile is loaded once A's values are multiplied in (its registers
// are free then), so that the load runs during the inverse.
template <int M>
__global__ void __launch_bounds__(256) ntt_row_fused_kernel(PassArgs a) {
    __shared__ double lds[kLdsSlots];
    constexpr uint32_t kJ = 1u << (M - 4);
    const BlockPos bp = block_pos(a.prod_start, a.n_prods, tiles_per_product(a.plan), a.xcd_log2);
    const NttProd pr = a.prods[bp.seg];
    const NttOp oa = a.ops[pr.a], ob = a.ops[pr.b];
    const uint64_t Kp = plane_row_elems(a.plan);
    const size_t off = plane_index(a.plan, (size_t)bp.t * kTile) + row_step0_base<M>();
    double *TA0 = a.T + oa.base + (size_t)bp.row * Kp + off;
    const uint32_t brow = a.rows ? a.rows[bp.row].b : pr.b_bcast ? 0u : bp.row;
    const double *TB0 = a.T + ob.base + (size_t)brow * Kp + off;
    const bool square = pr.a == pr.b;  // uniform over the launch
    double ra[16], v[16];
    memory_first_begin();
#pragma unroll
    for (int j = 0; j < 16; ++j) ra[j] = TA0[j * kJ];
    memory_first_end();
#pragma clang loop unroll(disable)
    for (int P = 0; P < 2; ++P) {
        const Mod m = mod_of(P);
        const double *tab = a.tables + P * kPrimeTable;
        double *TA = TA0 + P * oa.pstride;
        const double *TB = TB0 + P * ob.pstride;
        fwd_dft<M, false, kDifToPointA>(ra, lds, tab, P, m);
        __syncthreads();
        if (square) {
#pragma unroll
            for (int j = 0; j < 16; ++j) v[j] = mulp(ra[j], ra[j], m);
        } else {
            // an opaque copy of the table pointer keeps B's twiddle loads
            // after A's transform (fewer registers live)
            const double *tabb = tab;
            __asm__ volatile("" : "+s"(tabb));
            memory_first_begin();
#pragma unroll
            for (int j = 0; j < 16; ++j) v[j] = TB[j * kJ];
            memory_first_end();
            fwd_dft<M, false, kDifToPointB>(v, lds, tabb, P, m);
#pragma unroll
            for (int j = 0; j < 16; ++j) v[j] = mulp(v[j], ra[j], m);
        }
        if (P == 0) {
            memory_first_begin();
#pragma unroll
            for (int j = 0; j < 16; ++j) ra[j] = TA[oa.pstride + j * kJ];
            memory_first_end();
        }
        inv_dft<M, false>(v, lds, tab, P, m);
#pragma unroll
        for (int j = 0; j < 16; ++j) TA[j * kJ] = v[j];
        __syncthreads();
    }
}

// Largest last level the paired kernel takes over from the single-product
// one.
constexpr int kPairMaxM = 12;

// Last level of two products sharing B (C_q = A_q B, q = 0, 1; product q
// has rows0 / rows1 rows, B row r serves row r of both): per prime, B's tile
// is loaded and transformed once and kept in registers; each product's A
// tile is then transformed, multiplied, inverted and written back over A.
// Block b covers row b / tiles, tile b % tiles of both products.
template <int M>
__global__ void __launch_bounds__(256) ntt_row_pair_kernel(PassArgs a, uint32_t rows0, uint32_t rows1) {
    __shared__ double lds[kLdsSlots];
    constexpr uint32_t kJ = 1u << (M - 4);
    const uint32_t tiles = tiles_per_product(a.plan);
    const uint32_t bid = xcd_tile(blockIdx.x, gridDim.x, a.xcd_log2);
    const uint32_t row = bid / tiles, t = bid % tiles;
    const NttOp ob = a.ops[a.prods[0].b];
    const uint64_t Kp = plane_row_elems(a.plan);
    const size_t off = plane_index(a.plan, (size_t)t * kTile) + row_step0_base<M>();
    const double *TB0 = a.T + ob.base + (size_t)(a.prods[0].b_bcast ? 0u : row) * Kp + off;
#pragma clang loop unroll(disable)
    for (int P = 0; P < 2; ++P) {
        const Mod m = mod_of(P);
        const double *tab = a.tables + P * kPrimeTable;
        const double *TB = TB0 + P * ob.pstride;
        double rb[16], v[16];
        memory_first_begin();
#pragma unroll
        for (int j = 0; j < 16; ++j) rb[j] = TB[j * kJ];
        memory_first_end();
        fwd_dft<M, false, kDifToPointB>(rb, lds, tab, P, m);
#pragma clang loop unroll(disable)
        for (int q = 0; q < 2; ++q) {
            if (row >= (q ? rows1 : rows0)) continue;  // uniform over the block
            const NttOp oa = a.ops[a.prods[q].a];
            double *TA = a.T + oa.base + (size_t)row * Kp + off + P * oa.pstride;
            // the table pointer is opaque per product, so that the twiddle
            // loads stay inside the loop instead of occupying registers
            // across it
            const double *tabq = tab;
            __asm__ volatile("" : "+s"(tabq));
            __syncthreads();
            memory_first_begin();
#pragma unroll
            for (int j = 0; j < 16; ++j) v[j] = TA[j * kJ];
            memory_first_end();
            fwd_dft<M, false, kDifToPointA>(v, lds, tabq, P, m);
#pragma unroll
            for (int j = 0; j < 16; ++j) v[j] = mulp(rb[j], v[j], m);
            inv_dft<M, false>(v, lds, tabq, P, m);
#pragma unroll
            for (int j = 0; j < 16; ++j) TA[j * kJ] = v[j];
        }
        __syncthreads();
    }
}

// Inverse column level of prime P on the thread's last-step values v (as
// loaded from ct.plast(k)): inverse twiddle, inverse DFT; leaves the
// thread's elements tid + 256 u in v.  The caller must make sure no thread
// still reads `lds`.
template <int M, int P, int FINAL = kDitLast>
__device__ __forceinline__ void col_inv_prime(double *v, double *lds, const ColThread<M> &ct, const double *tables,
                                              const ColTw &tw) {
    const Mod m = mod_of(P);
    const double *tab = tables + P * kPrimeTable;
    column_twiddles<M>(v, tw, m);
    inv_dft<M, true, FINAL>(v, lds, tab, P, m, ct.tid);
}

// Loads both primes' values of a column tile and their inverse column
// twiddle table entries (every load before any store).
template <int M>
__device__ __forceinline__ void col_inv_load(double *v0, double *v1, ColTw &tw0, ColTw &tw1, const PassArgs &a,
                                             const ColThread<M> &ct, const double *T0, const double *T1) {
#pragma unroll
    for (int k = 0; k < 16; ++k) v0[k] = T0[ct.plast(k)];
#pragma unroll
    for (int k = 0; k < 16; ++k) v1[k] = T1[ct.plast(k)];
    tw0 = col_tw_load<M>(a.tables, ct.cg, ct.B0, true);
    tw1 = col_tw_load<M>(a.tables + kPrimeTable, ct.cg, ct.B0, true);
}

// Inverse column level above the first: both primes through one tile,
// stored together at the end.
template <int M, bool PAIR>
__global__ void __launch_bounds__((ColPair<M, PAIR>::kThreads)) ntt_col_inv_kernel(PassArgs a, int level) {
    using CP = ColPair<M, PAIR>;
    __shared__ double lds_tiles[CP::kLdsWords];
    double *const lds = lds_tiles + CP::half() * CP::kTileStride;
    const BlockPos bp = block_pos(a.prod_start, a.n_prods, tiles_per_product(a.plan) >> CP::kLog2, a.xcd_log2);
    const NttOp oa = a.ops[a.prods[bp.seg].a];
    const ColThread<M> ct(a.plan, level, (bp.t << CP::kLog2) + CP::half(), CP::vtid());
    double *T0 = a.T + oa.base + (size_t)bp.row * plane_row_elems(a.plan);
    double *T1 = T0 + oa.pstride;
    double v0[16], v1[16];
    ColTw tw0, tw1;
    col_inv_load<M>(v0, v1, tw0, tw1, a, ct, T0, T1);
    col_inv_prime<M, 0>(v0, lds, ct, a.tables, tw0);
    __syncthreads();
    col_inv_prime<M, 1>(v1, lds, ct, a.tables, tw1);
#pragma unroll
    for (int u = 0; u < 16; ++u) T0[ct.poff(u)] = v0[u];
#pragma unroll
    for (int u = 0; u < 16; ++u) T1[ct.poff(u)] = v1[u];
}

// x of lane - D inside the lane's row of 16 lanes (DPP row_shr; 0 where that
// lane lies before the row): a VALU operand modifier, where a general
// shuffle would go through the LDS crossbar.
template <int D>
__device__ __forceinline__ uint32_t row_shr(uint32_t x) {
    static_assert(D >= 1 && D <= 15, "row_shr:1..15");
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x110 + D, 0xf, 0xf, true);
}

// First inverse column level: both primes through one tile, 1/K, Garner,
// and the limbs of each run of adjacent coefficients with the run's spill
// word.  A thread's element u is (row r0 + u 2^(M-4), column g) and lanes
// hold adjacent columns, so a run of 2^RL adjacent coefficients sits in
// 2^RL adjacent lanes: coefficient c = w0 + 2^32 w1 + 2^64 w2 adds to the
// limb column sums s_j = w0_j + w1_(j-1) + w2_(j-2), which become limbs with
// one neighbour shift and one ballot carry resolution per run; what passes
// the run's top goes to the spill word (the carry engine adds it).
template <int M, bool PAIR>
__global__ void __launch_bounds__((ColPair<M, PAIR>::kThreads)) ntt_col_inv_limbs_kernel(PassArgs a) {
    using CP = ColPair<M, PAIR>;
    __shared__ double lds_tiles[CP::kLdsWords];
    double *const lds = lds_tiles + CP::half() * CP::kTileStride;
    constexpr int LC = kTileLog2 - M;
    constexpr int RL = run_log2_for(M);
    constexpr uint32_t kRun = 1u << RL;
    // a run (2^RL lanes aligned to 2^RL) lies inside one group of 2^LC
    // adjacent lanes, which hold one row of one tile also in paired blocks,
    // and inside one DPP row of 16 lanes (the neighbour shifts)
    static_assert(kRun <= (1u << LC) && kRun <= 16u, "a run is one tile row inside one DPP row");
    const BlockPos bp = block_pos(a.prod_start, a.n_prods, tiles_per_product(a.plan) >> CP::kLog2, a.xcd_log2);
    const uint32_t prod = bp.row;
    const NttProd pr = a.prods[bp.seg];
    const NttOp oa = a.ops[pr.a];
    const int log2K = a.plan.log2K;
    const ColThread<M> ct(a.plan, 0, (bp.t << CP::kLog2) + CP::half(), CP::vtid());
    const double *T0 = a.T + oa.base + (size_t)prod * plane_row_elems(a.plan);
    double v0[16], v1[16];
    ColTw tw0, tw1;
    col_inv_load<M>(v0, v1, tw0, tw1, a, ct, T0, T0 + oa.pstride);
    // 1/K rides on the column twiddles (the inverse transform is linear):
    // one product per prime instead of one per value; the last inverse step
    // re-centres its outputs (|v| < p/2 + p/1024) for the recombination
    tw0.lb = mulp(tw0.lb, c_ntt.kinv[0][log2K], mod_of(0));
    tw1.lb = mulp(tw1.lb, c_ntt.kinv[1][log2K], mod_of(1));
    col_inv_prime<M, 0, kDitToGarner>(v0, lds, ct, a.tables, tw0);
    __syncthreads();
    col_inv_prime<M, 1, kDitToGarner>(v1, lds, ct, a.tables, tw1);
    // prime 1's elements wait in the thread's own LDS slots while the limbs
    // are formed (registers for the carry resolution)
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 16; ++u) lds[ct.lds0 + 272u * u] = v1[u];

    const uint32_t lane = threadIdx.x & (kWave - 1u);
    const uint32_t jr = lane & (kRun - 1u);  // position inside the run
    uint64_t last_mask = 0;                   // last lanes of runs pass no carry on
#pragma unroll
    for (uint32_t l = kRun - 1; l < kWave; l += kRun) last_mask |= 1ull << l;
    uint32_t *out = pr.C + (size_t)prod * pr.sC;
    const uint32_t Lc = pr.Lc;
    uint64_t *spill = a.spill + pr.spill_base + (size_t)prod * ((1ull << log2K) >> RL);
    const uint64_t n_u0 = ct.off0;  // coefficient index of element u = 0
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const unsigned __int128 c = recombine(v0[u], lds[ct.lds0 + 272u * u]);
        const uint32_t w0 = (uint32_t)c, w1 = (uint32_t)(c >> 32), w2 = (uint32_t)(c >> 64);
        const uint32_t w1a = row_shr<1>(w1), w2a = row_shr<1>(w2), w2b = row_shr<2>(w2);
        const uint64_t s = (uint64_t)w0 + (jr >= 1 ? w1a : 0u) + (jr >= 2 ? w2b : 0u);
        const uint32_t lo = (uint32_t)s, hi = (uint32_t)(s >> 32);
        const uint32_t hia = row_shr<1>(hi);
        const uint64_t t = (uint64_t)lo + (jr >= 1 ? hia : 0u);
        const uint32_t lo2 = (uint32_t)t;
        const bool gen = (t >> 32) != 0, prop = lo2 == 0xffffffffu;
        const uint64_t G = ballot64(gen) & ~last_mask, P = ballot64(prop) & ~last_mask;
        uint32_t unused;
        const uint64_t cm = resolve_carries(G, P, 0u, &unused);
        const uint32_t ci = (uint32_t)(cm >> lane) & 1u;
        const uint64_t n = n_u0 + ((uint64_t)u << ct.ustep);
        if (n < Lc) out[n] = lo2 + ci;
        if (jr == kRun - 1) {
            // V >> 32 kRun: the carry out of the top limb, the top column's
            // high part and the words of the last two coefficients past the run
            const uint64_t cout = (gen || (prop && ci)) ? 1u : 0u;
            spill[(n - jr) >> RL] = cout + hi + (uint64_t)w1 + (uint64_t)w2a + ((uint64_t)w2 << 32);
        }
    }
}

// K = 2^M <= 4096, one prime: transforms of A (kept in registers) and B,
// pointwise, inverse without 1/K; res[j] is K times the result, re-centred
// (|res| < p/2 + p/1024), at tile element row_step0_base<M>() + j 2^(M-4),
// the element the last inverse step read from `lds` (so it may be written
// back there without a barrier).  The coefficients are below 2^76, so K
// times them stays below 2^88 < p0 p1: the recombination gives K c exactly
// and a shift by M bits divides it out.
template <int M>
__device__ __forceinline__ void local_prime(const PassArgs &a, double *lds, uint32_t prod0, int P, double *res) {
    constexpr uint32_t kJ = 1u << (M - 4);
    const Mod m = mod_of(P);
    const double *tab = a.tables + P * kPrimeTable;
    const uint32_t i0 = row_step0_base<M>();
    double ra[16], v[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const uint32_t i = i0 + j * kJ;
        const uint32_t prod = prod0 + (i >> M), r = i & ((1u << M) - 1u);
        ra[j] = 0.0;
        v[j] = 0.0;
        if (prod < a.N) {
            if (a.rows) {
                const RaggedRow it = a.rows[prod];
                if (r < it.L_a) ra[j] = (double)it.A[r];
                if (r < it.L_b) v[j] = (double)it.B[r];
            } else {
                if (r < a.La) ra[j] = (double)a.A[(size_t)prod * a.sA + r];
                if (r < a.Lb) v[j] = (double)a.B[(size_t)prod * a.sB + r];
            }
        }
    }
    fwd_dft<M, false, kDifToPointA>(ra, lds, tab, P, m);
    if constexpr (M > 4) __syncthreads();
    fwd_dft<M, false, kDifToPointB>(v, lds, tab, P, m);
#pragma unroll
    for (int j = 0; j < 16; ++j) v[j] = mulp(v[j], ra[j], m);
    inv_dft<M, false, kDitToGarner>(v, lds, tab, P, m);
#pragma unroll
    for (int j = 0; j < 16; ++j) res[j] = step_count(M) == 1 ? redp(v[j], m) : v[j];
}

// K = 2^M <= 4096: whole products inside one tile (4096 / K per block).
// One tile buffer: both primes transform in `work` in turn, prime 0's
// results wait half in registers and half in `hi_w` (16 KB), and the
// recombined coefficients (below 2^75:
// at most 2^11 products of two limbs each) replace the thread's own
// elements as a low 64-bit word in `work` and a high word in `hi_w`.
// 52.2 KB of LDS, three blocks per CU.
template <int M>
__global__ void __launch_bounds__(256) ntt_local_kernel(PassArgs a) {
    __shared__ double work[kLdsSlots];
    __shared__ uint64_t hi_w[kLdsSlots / 2];  // 32-bit high words, then the spill words
    __shared__ uint32_t wave_flags[4];
    constexpr uint32_t kJ = 1u << (M - 4);
    uint64_t *const lo_s = reinterpret_cast<uint64_t *>(work);
    uint32_t *const hi_s = reinterpret_cast<uint32_t *>(hi_w);
    const uint32_t prod0 = xcd_tile(blockIdx.x, gridDim.x, a.xcd_log2) * (kTile >> M);
    // prime 0's results: the first half stays in registers, the second
    // half waits in hi_w (8 x 256 doubles, lane-consecutive)
    double res0[16], res1[16];
    local_prime<M>(a, work, prod0, 0, res0);
    double *const park = reinterpret_cast<double *>(hi_w);
#pragma unroll
    for (int j = 8; j < 16; ++j) park[(j - 8) * 256 + threadIdx.x] = res0[j];
    __syncthreads();
    local_prime<M>(a, work, prod0, 1, res1);
#pragma unroll
    for (int j = 8; j < 16; ++j) res0[j] = park[(j - 8) * 256 + threadIdx.x];
    // every parked word is read before the high words take their place
    __syncthreads();
    const uint32_t is0 = row_step0_base<M>();
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const uint32_t e = lds_index(is0 + j * kJ);
        static_assert((unsigned __int128)kPrime0 * kPrime1 > ((unsigned __int128)1 << (75 + 12)),
                      "K times a one-tile coefficient stays below p0 p1");
        const unsigned __int128 c = recombine(res0[j], res1[j]) >> M;
        lo_s[e] = (uint64_t)c;
        hi_s[e] = (uint32_t)(c >> 64);
    }
    __syncthreads();

    // each thread: 16 consecutive coefficients of one product
    const uint32_t i0 = threadIdx.x * 16u;
    const uint32_t prod = prod0 + (i0 >> M);
    const uint32_t r0 = i0 & ((1u << M) - 1u);
    const uint32_t e0 = lds_index(i0);
    uint32_t limb[16];
    unsigned __int128 pending = 0;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        pending += ((unsigned __int128)hi_s[e0 + j] << 64) + lo_s[e0 + j];
        limb[j] = (uint32_t)pending;
        pending >>= 32;
    }
    // every high word is read before the spill words take their place
    __syncthreads();
    uint64_t *const spill_s = hi_w;
    spill_s[threadIdx.x] = (uint64_t)pending;
    __syncthreads();
    const uint64_t sin = (r0 == 0) ? 0ull : spill_s[threadIdx.x - 1];
    uint64_t acc = (uint64_t)limb[0] + (uint32_t)sin;
    limb[0] = (uint32_t)acc;
    acc = (uint64_t)limb[1] + (uint32_t)(sin >> 32) + (acc >> 32);
    limb[1] = (uint32_t)acc;
    uint32_t c = (uint32_t)(acc >> 32);
    bool all_ones = limb[0] == 0xffffffffu && limb[1] == 0xffffffffu;
#pragma unroll
    for (int j = 2; j < 16; ++j) {
        const uint64_t x = (uint64_t)limb[j] + c;
        limb[j] = (uint32_t)x;
        c = (uint32_t)(x >> 32);
        all_ones = all_ones && limb[j] == 0xffffffffu;
    }
    // the last thread of a product neither passes nor forwards a carry
    const bool last_of_product = (r0 + 16u) == (1u << M);
    const bool gen = !last_of_product && c != 0;
    const bool prop = !last_of_product && c == 0 && all_ones;
    const uint64_t Gm = ballot64(gen), Pm = ballot64(prop);
    const uint32_t wave = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
    uint32_t wout;
    (void)resolve_carries(Gm, Pm, 0u, &wout);
    if (lane == 0) wave_flags[wave] = wout | (Pm == ~0ull ? 2u : 0u);
    __syncthreads();
    uint32_t cin = 0;
    for (uint32_t w = 0; w < wave; ++w) {
        const uint32_t f = wave_flags[w];
        cin = (f & 1u) | ((f >> 1) & cin);
    }
    uint32_t dummy;
    const uint64_t cm = resolve_carries(Gm, Pm, cin, &dummy);
    uint32_t ci = (uint32_t)(cm >> lane) & 1u;
    if (prod < a.N) {
        uint32_t *out = a.C + (size_t)prod * a.sC;
        uint32_t Lc = a.Lc;
        if (a.rows) {
            const RaggedRow it = a.rows[prod];
            out = it.C;
            Lc = it.L_a + it.L_b;
        }
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const uint64_t x = (uint64_t)limb[j] + ci;
            ci = (uint32_t)(x >> 32);
            if (r0 + j < Lc) out[r0 + j] = (uint32_t)x;
        }
    }
}

// Calls f(std::integral_constant<int, M>) for the M in [LO, HI] equal to m.
template <int LO, int HI, typename F>
void dispatch_m(int m, F &&f) {
    if constexpr (LO <= HI) {
        if (m == LO) f(std::integral_constant<int, LO>{});
        else dispatch_m<LO + 1, HI>(m, f);
    }
}
constexpr int kColLo = 5, kColHi = 9, kRowLo = 6, kRowHi = 12, kLocalLo = kMinLog2, kLocalHi = kTileLog2;

// ---------------------------------------------------------------- host side

uint64_t mulmod_u64(uint64_t a, uint64_t b, uint64_t m) { return (uint64_t)((unsigned __int128)a * b % m); }
uint64_t powmod_u64(uint64_t b, uint64_t e, uint64_t m) {
    uint64_t r = 1 % m;
    b %= m;
    while (e) {
        if (e & 1) r = mulmod_u64(r, b, m);
        b = mulmod_u64(b, b, m);
        e >>= 1;
    }
    return r;
}
double centred(uint64_t v, uint64_t p) { return v > p / 2 ? -(double)(p - v) : (double)v; }

constexpr size_t kChunkElems = size_t(1) << 28;  // products * K per launch chunk

int log2_ceil(uint64_t x) {
    int l = 0;
    while ((1ull << l) < x) ++l;
    return l;
}

int transform_log2(uint32_t L_a, uint32_t L_b) {
    int k = log2_ceil((uint64_t)L_a + L_b);
    return k < kMinLog2 ? kMinLog2 : k;
}

size_t chunk_products(uint32_t N, int log2K) {
    size_t n = kChunkElems >> log2K;
    if (n < 1) n = 1;
    if (n > N) n = N;
    return n;
}

// A ragged launch runs as one chunk (its B set stays resident throughout).
uint32_t ragged_max_rows(int log2K) { return (uint32_t)std::max<size_t>(1, kChunkElems >> log2K); }

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Optional per-launch timing of the NTT kernels (HIP events on the launch
// stream) with each launch's algorithmic bytes, for live roofline numbers.
enum KernelKind { kColFwdLimbs, kColFwd, kRowFused, kRowPair, kColInv, kColInvLimbs, kLocal, kKindCount };
const char *const kKindNames[kKindCount] = {"ntt_col_fwd_kernel<limbs>", "ntt_col_fwd_kernel",
                                            "ntt_row_fused_kernel", "ntt_row_pair_kernel",
                                            "ntt_col_inv_kernel", "ntt_col_inv_limbs_kernel",
                                            "ntt_local_kernel"};
struct StatRecord {
    int kind, log2K;
    hipEvent_t start, stop;
    double bytes;
};
// Process-wide; several threads (the serving loop's workers) may launch at
// once, so the records and the event pool are guarded by `mu`.
struct Stats {
    std::atomic<bool> enabled{false};
    std::mutex mu;
    std::vector<StatRecord> records;
    std::vector<hipEvent_t> pool;
} g_stats;

// Called with g_stats.mu held.
hipEvent_t stats_event() {
    if (!g_stats.pool.empty()) {
        hipEvent_t e = g_stats.pool.back();
        g_stats.pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Set on a thread while it runs a launch sequence that must not create
// events (the allocation-free multiply).
thread_local bool t_stats_off = false;

// Records `launch` between two events when stats are enabled.
template <typename F>
hipError_t timed(int kind, int log2K, double bytes, hipStream_t stream, F &&launch) {
    if (t_stats_off || !g_stats.enabled.load(std::memory_order_relaxed)) {
        launch();
        return hipGetLastError();
    }
    StatRecord r{kind, log2K, nullptr, nullptr, bytes};
    {
        std::lock_guard<std::mutex> lk(g_stats.mu);
        r.start = stats_event();
        r.stop = stats_event();
    }
    if (r.start == nullptr || r.stop == nullptr) return hipErrorOutOfMemory;
    GPUPI_RETURN_IF(hipEventRecord(r.start, stream));
    launch();
    GPUPI_RETURN_IF(hipGetLastError());
    GPUPI_RETURN_IF(hipEventRecord(r.stop, stream));
    std::lock_guard<std::mutex> lk(g_stats.mu);
    g_stats.records.push_back(r);
    return hipSuccess;
}

// Level sizes chosen at run time per transform size (gpupi_ntt_set_split):
// m0 | m1 << 8, zero for make_plan's choice.  Read once per product launch
// sequence, so a change takes effect from the next multiply on.
std::atomic<uint32_t> g_split[kMaxLog2 + 1];

Plan active_plan(int k) {
    const uint32_t s = g_split[k].load(std::memory_order_relaxed);
    return make_plan_split(k, (int)(s & 0xffu), (int)(s >> 8));
}

// Launch shape (gpupi_ntt_set_launch): which column level kinds pair their
// tiles, and the XCD chunk of the block order.  Read once per product
// launch sequence.
std::atomic<uint32_t> g_pair_mask{kPairMaskDefault};
std::atomic<int> g_xcd_log2{kXcdChunkLog2};

// Blocks of the persistent first-level kernel: as many as fit on the device
// at once (a multiple of the XCD count), or one per tile block when there
// are fewer.
uint32_t device_cus() {
    static std::atomic<int> cus{0};
    int n = cus.load(std::memory_order_relaxed);
    if (n == 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n < 1)
            n = 1;
        cus.store(n, std::memory_order_relaxed);
    }
    return (uint32_t)n;
}
template <int M, bool PAIR>
__global__ void ntt_col_fwd_limbs_kernel(PassArgs a, uint32_t total);
template <int M, bool PAIR>
uint32_t persistent_grid(uint32_t total) {
    static std::atomic<int> per_cu{0};
    int b = per_cu.load(std::memory_order_relaxed);
    if (b == 0) {
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, ntt_col_fwd_limbs_kernel<M, PAIR>,
                                                         ColPair<M, PAIR>::kThreads, 0) != hipSuccess || b < 1)
            b = 1;
        per_cu.store(b, std::memory_order_relaxed);
    }
    const uint64_t fit = ((uint64_t)b * device_cus()) / kXcds * kXcds;
    return (uint32_t)std::min<uint64_t>(total, fit < kXcds ? kXcds : fit);
}

// Calls f(std::integral_constant<bool, PAIR>) with PAIR = pair where a level
// of 2^M points can pair its tiles, false otherwise.
template <int M, typename F>
void with_pair(bool pair, F &&f) {
    if constexpr (col_pair_log2(M) > 0) {
        if (pair) {
            f(std::true_type{});
            return;
        }
    }
    f(std::false_type{});
}

// Spill words per product for any plan of size 2^k: runs are never shorter
// than run_log2_for(kColMaxLog2) limbs.
size_t spill_runs_bound(int k) { return (size_t(1) << k) >> run_log2_for(kColMaxLog2); }

}  // namespace

void ntt_stats_suspend(bool suspend) { t_stats_off = suspend; }

bool ntt_supports(uint32_t L_a, uint32_t L_b) {
    return L_a > 0 && L_b > 0 && transform_log2(L_a, L_b) <= kMaxLog2;
}

hipError_t ntt_tables_init(NTTPrecomputedTables *tables) {
    const uint64_t primes[2] = {kPrime0, kPrime1};
    const uint64_t gens[2] = {kGen0, kGen1};
    std::vector<double> host(kTableLen);
    NttConsts cst{};
    for (int q = 0; q < 2; ++q) {
        const uint64_t P = primes[q];
        double *tab = host.data() + q * kPrimeTable;
        const uint64_t w30 = powmod_u64(gens[q], (P - 1) >> kMaxLog2, P);
        const uint64_t w30i = powmod_u64(w30, P - 2, P);
        // in-tile step tables, [j][n2] for Ms = 2^x
        for (int x = 5; x <= kTileLog2; ++x) {
            const uint64_t wx = powmod_u64(w30, 1ull << (kMaxLog2 - x), P);
            const uint64_t wxi = powmod_u64(wx, P - 2, P);
            const uint32_t ss = 1u << (x - 4);
            for (uint32_t j = 0; j < 16; ++j) {
                const uint32_t bj = bitrev(j, 4);
                for (uint32_t n2 = 0; n2 < ss; ++n2) {
                    const size_t slot = ((size_t)1 << x) - 32 + (size_t)j * ss + n2;
                    tab[kOffStep + slot] = centred(powmod_u64(wx, (uint64_t)n2 * bj, P), P);
                    tab[kOffStepInv + slot] = centred(powmod_u64(wxi, (uint64_t)n2 * bj, P), P);
                }
            }
        }
        const uint64_t step_hi = powmod_u64(w30, 32768, P), step_hi_i = powmod_u64(w30i, 32768, P);
        uint64_t lo = 1, loi = 1, hi = 1, hii = 1;
        for (size_t e = 0; e < 32768; ++e) {
            tab[kOffLo + e] = centred(lo, P);
            tab[kOffLoInv + e] = centred(loi, P);
            tab[kOffHi + e] = centred(hi, P);
            tab[kOffHiInv + e] = centred(hii, P);
            lo = mulmod_u64(lo, w30, P);
            loi = mulmod_u64(loi, w30i, P);
            hi = mulmod_u64(hi, step_hi, P);
            hii = mulmod_u64(hii, step_hi_i, P);
        }
        const uint64_t w16 = powmod_u64(w30, 1ull << (kMaxLog2 - 4), P);
        const uint64_t w16i = powmod_u64(w16, P - 2, P);
        for (int e = 0; e < 8; ++e) {
            cst.w16[q][e] = centred(powmod_u64(w16, e, P), P);
            cst.w16i[q][e] = centred(powmod_u64(w16i, e, P), P);
        }
        for (int k = 0; k <= kMaxLog2; ++k)
            cst.kinv[q][k] = centred(powmod_u64(powmod_u64(2, k, P), P - 2, P), P);
    }
    cst.garner = centred(powmod_u64(kPrime0 % kPrime1, kPrime1 - 2, kPrime1), kPrime1);

    double *d = nullptr;
    GPUPI_RETURN_IF(hipMalloc(&d, kTableLen * sizeof(double)));
    hipError_t err = hipMemcpy(d, host.data(), kTableLen * sizeof(double), hipMemcpyHostToDevice);
    if (err == hipSuccess) err = hipMemcpyToSymbol(HIP_SYMBOL(c_ntt), &cst, sizeof(cst));
    if (err != hipSuccess) {
        (void)hipFree(d);
        return err;
    }
    tables->roots = d;
    tables->max_log2 = kMaxLog2;
    return hipSuccess;
}

void ntt_tables_release(NTTPrecomputedTables *tables) {
    if (tables->roots) (void)hipFree(tables->roots);
    tables->roots = nullptr;
    tables->max_log2 = 0;
}

namespace {

// Copies limbs [0, L_a + L_b) of each staged product row (stride K) to its
// row's C (rows without C stay staged for their assembly).  Blocks: (row,
// 1024-limb slice).
__global__ void __launch_bounds__(256) ragged_copy_kernel(const uint32_t *staging, uint64_t K,
                                                          const RaggedRow *rows, uint32_t slices) {
    const uint32_t row = blockIdx.x / slices, slice = blockIdx.x % slices;
    const RaggedRow it = rows[row];
    if (it.C == nullptr) return;
    const uint32_t Lc = it.L_a + it.L_b;
    const uint32_t *src = staging + (size_t)row * K;
    for (uint32_t i = slice * 1024u + threadIdx.x; i < Lc && i < (slice + 1) * 1024u; i += 256u) it.C[i] = src[i];
}

// One NTT launch sequence over products of one transform size 2^k: up to
// three forward operands and two products with rows at strides (a product's
// A operand receives the product and is used by nothing else; a B operand
// may be shared by both products, or be one row at stride 0 that every row
// multiplies), or ragged rows (operand 0: each row's A, operand 1: the
// launch's B set, one product; one launch chunk).
struct NttJob {
    struct Op {
        const uint32_t *X = nullptr;
        uint32_t L = 0, s = 0;
    };
    struct Prod {
        uint32_t a = 0, b = 1;
        uint32_t *C = nullptr;
        uint32_t sC = 0, N = 0;
    };
    Op ops[3];
    Prod prods[2];
    uint32_t n_ops = 0, n_prods = 0;
    bool ragged = false;
    const RaggedRow *rows = nullptr;  // ragged: device rows, B set and assemblies, with host copies
    const RaggedRow *host_rows = nullptr;
    const RaggedB *bset = nullptr;
    const RaggedB *host_bset = nullptr;
    uint32_t nb = 0;
    const RaggedAssembly *asms = nullptr;
    const RaggedAssembly *host_asms = nullptr;
    uint32_t n_asm = 0, asm_chunks = 0;
    int k = 0;
    size_t max_chunk = 0;  // rows per launch chunk at most (0: the default)
};

// Rows each operand needs (the most of any product using it; one for a
// broadcast row), the rows of the whole job and each product's length.
struct JobShape {
    uint32_t op_rows[3] = {0, 0, 0};
    bool op_bcast[3] = {false, false, false};
    uint32_t rows = 0;
    uint32_t Lc[2] = {0, 0};
};

JobShape shape_of(const NttJob &j) {
    JobShape sh;
    const size_t K = size_t(1) << j.k;
    bool used_as_a[3] = {false, false, false};
    for (uint32_t q = 0; q < j.n_prods; ++q) {
        const NttJob::Prod &pr = j.prods[q];
        sh.rows = std::max(sh.rows, pr.N);
        sh.op_rows[pr.a] = std::max(sh.op_rows[pr.a], pr.N);
        sh.op_rows[pr.b] = std::max(sh.op_rows[pr.b], pr.N);
        used_as_a[pr.a] = true;
        sh.Lc[q] = j.ragged ? (uint32_t)K : j.ops[pr.a].L + j.ops[pr.b].L;
    }
    for (uint32_t o = 0; o < j.n_ops; ++o) {
        sh.op_bcast[o] = !j.ragged && !used_as_a[o] && j.ops[o].s == 0 && sh.op_rows[o] > 1;
        if (sh.op_bcast[o]) sh.op_rows[o] = 1;
    }
    if (j.ragged) sh.op_rows[1] = j.nb;
    return sh;
}

// Whether the operands and products describe a job run_ntt can do.
bool job_ok(const NttJob &j) {
    if (j.n_ops < 1 || j.n_ops > 3 || j.n_prods < 1 || j.n_prods > 2) return false;
    if (j.ragged && (j.rows == nullptr || j.host_rows == nullptr || j.bset == nullptr || j.host_bset == nullptr ||
                     j.nb == 0 || j.nb > j.prods[0].N || j.n_ops != 2 || j.n_prods != 1 ||
                     (j.n_asm > 0 && (j.asms == nullptr || j.host_asms == nullptr))))
        return false;
    for (uint32_t q = 0; q < j.n_prods; ++q) {
        const NttJob::Prod &pr = j.prods[q];
        if (pr.a >= j.n_ops || pr.b >= j.n_ops) return false;
        if (pr.a == pr.b && (j.ragged || j.n_prods != 1)) return false;  // squares: one product
        for (uint32_t r = 0; r < j.n_prods; ++r)
            if (r != q && (j.prods[r].a == pr.a || j.prods[r].b == pr.a)) return false;
        if (!j.ragged && transform_log2(j.ops[pr.a].L, j.ops[pr.b].L) != j.k) return false;
    }
    return true;
}

struct JobLayout {
    size_t plane_elems[3] = {0, 0, 0};  // per prime, per operand (elements)
    size_t op_base[3] = {0, 0, 0};      // element offset of operand o's prime-0 plane
    size_t spill_elems = 0;             // per product
    size_t spill_off = 0, flags_off = 0, staging_off = 0, asm_flags_off = 0, total = 0;  // bytes
    size_t nc = 0;
};

JobLayout layout_of(const NttJob &j, const JobShape &sh) {
    JobLayout l;
    const size_t K = size_t(1) << j.k;
    // a ragged launch is one chunk (its B set and staged pieces stay resident)
    l.nc = j.ragged ? sh.rows : chunk_products(sh.rows, j.k);
    if (j.max_chunk != 0 && l.nc > j.max_chunk) l.nc = j.max_chunk;
    size_t off = 0;
    for (uint32_t o = 0; o < j.n_ops; ++o) {
        const size_t cap = sh.op_bcast[o] ? 1 : std::min<size_t>(l.nc, sh.op_rows[o]);
        l.plane_elems[o] = align256(cap * plane_row_elems(active_plan(j.k)) * sizeof(double)) / sizeof(double);
        l.op_base[o] = off / sizeof(double);
        off += 2 * l.plane_elems[o] * sizeof(double);
    }
    l.spill_off = off;
    l.spill_elems = l.nc * spill_runs_bound(j.k);
    off += j.n_prods * align256(l.spill_elems * sizeof(uint64_t));
    l.flags_off = off;
    size_t fl = 0;
    for (uint32_t q = 0; q < j.n_prods; ++q) fl = std::max(fl, carry_flags_bytes((uint32_t)l.nc, sh.Lc[q]));
    off += align256(fl);
    l.staging_off = off;
    if (j.ragged) off += align256(l.nc * K * sizeof(uint32_t));
    l.asm_flags_off = off;
    if (j.ragged) off += align256(j.asm_chunks);
    l.total = off;
    return l;
}

size_t job_workspace_bytes(const NttJob &j) {
    if (make_plan(j.k).levels == 1) return 0;
    const JobShape sh = shape_of(j);
    if (sh.rows == 0) return 0;
    return layout_of(j, sh).total;
}

// One-tile products (K <= 4096): every product in one launch of the local
// kernel.
hipError_t run_local(const NttJob &j, const NTTPrecomputedTables &tables, hipStream_t stream) {
    const int k = j.k;
    for (uint32_t q = 0; q < j.n_prods; ++q) {
        const NttJob::Prod &pr = j.prods[q];
        if (pr.N == 0) continue;
        PassArgs a{};
        a.plan = make_plan(k);
        a.tables = tables.roots;
        a.A = j.ops[pr.a].X;
        a.B = j.ops[pr.b].X;
        a.La = j.ops[pr.a].L;
        a.Lb = j.ops[pr.b].L;
        a.sA = j.ops[pr.a].s;
        a.sB = j.ops[pr.b].s;
        a.C = pr.C;
        a.Lc = a.La + a.Lb;
        a.sC = pr.sC;
        a.N = pr.N;
        a.rows = j.rows;
        a.xcd_log2 = g_xcd_log2.load(std::memory_order_relaxed);
        const uint32_t G = kTile >> k;
        const uint64_t blocks = ceil_div64(pr.N, G);
        if (blocks > 0x7fffffffull) return hipErrorInvalidValue;
        double limbs = 0;
        if (j.ragged) {
            for (uint32_t i = 0; i < pr.N; ++i) limbs += (double)j.host_rows[i].L_a + j.host_rows[i].L_b;
        } else {
            limbs = (double)pr.N * ((double)a.La + a.Lb);
        }
        GPUPI_RETURN_IF(timed(kLocal, k, 8.0 * limbs, stream, [&] {
            dispatch_m<kLocalLo, kLocalHi>(k, [&](auto mc) {
                hipLaunchKernelGGL((ntt_local_kernel<decltype(mc)::value>), dim3((uint32_t)blocks), dim3(256), 0,
                                   stream, a);
            });
        }));
    }
    return hipSuccess;
}

hipError_t run_ntt(const NttJob &j, void *workspace, size_t workspace_bytes, const NTTPrecomputedTables &tables,
                   hipStream_t stream) {
    if (!job_ok(j)) return hipErrorInvalidValue;
    const JobShape sh = shape_of(j);
    if (sh.rows == 0) return hipSuccess;
    const int k = j.k;
    if (make_plan(k).levels == 1) return run_local(j, tables, stream);
    const Plan p = active_plan(k);
    const size_t K = size_t(1) << k;
    const JobLayout l = layout_of(j, sh);
    if (workspace == nullptr || workspace_bytes < l.total) return hipErrorInvalidValue;
    const size_t runs = K >> final_run_log2(p);
    char *ws = static_cast<char *>(workspace);
    uint8_t *flags = reinterpret_cast<uint8_t *>(ws + l.flags_off);
    uint32_t *staging = reinterpret_cast<uint32_t *>(ws + l.staging_off);
    const uint32_t tiles = tiles_per_product(p);

    PassArgs a{};
    a.plan = p;
    a.tables = tables.roots;
    a.xcd_log2 = g_xcd_log2.load(std::memory_order_relaxed);
    const uint32_t pair_mask = g_pair_mask.load(std::memory_order_relaxed);
    a.T = reinterpret_cast<double *>(ws);
    a.spill = reinterpret_cast<uint64_t *>(ws + l.spill_off);
    a.n_ops = j.n_ops;
    a.n_prods = j.n_prods;
    const size_t spill_stride = align256(l.spill_elems * sizeof(uint64_t)) / sizeof(uint64_t);
    for (size_t r0 = 0; r0 < sh.rows; r0 += l.nc) {
        // operands: rows of this chunk (a broadcast row only in the first)
        uint64_t fwd_blocks = 0, limb_reads = 0;
        uint32_t op_rows[3] = {0, 0, 0};
        for (uint32_t o = 0; o < j.n_ops; ++o) {
            const NttJob::Op &op = j.ops[o];
            if (sh.op_bcast[o]) op_rows[o] = r0 == 0 ? 1u : 0u;
            else op_rows[o] = (uint32_t)std::min<size_t>(l.nc, sh.op_rows[o] > r0 ? sh.op_rows[o] - r0 : 0);
            a.ops[o].X = op.X + (sh.op_bcast[o] ? 0 : r0 * op.s);
            a.ops[o].L = op.L;
            a.ops[o].s = op.s;
            a.ops[o].base = l.op_base[o];
            a.ops[o].pstride = l.plane_elems[o];
            if (o < 3) a.op_start[o] = (uint32_t)fwd_blocks;
            fwd_blocks += (uint64_t)op_rows[o] * tiles;
            if (!j.ragged) limb_reads += (uint64_t)op_rows[o] * op.L;
        }
        uint64_t prod_blocks = 0, prod_rows = 0, limb_writes = 0;
        uint32_t rows[2] = {0, 0};
        for (uint32_t q = 0; q < j.n_prods; ++q) {
            const NttJob::Prod &pr = j.prods[q];
            rows[q] = (uint32_t)std::min<size_t>(l.nc, pr.N > r0 ? pr.N - r0 : 0);
            NttProd &d = a.prods[q];
            d.a = pr.a;
            d.b = pr.b;
            d.b_bcast = sh.op_bcast[pr.b] ? 1u : 0u;
            d.C = j.ragged ? staging : pr.C + r0 * pr.sC;
            d.sC = j.ragged ? (uint32_t)K : pr.sC;
            d.Lc = sh.Lc[q];
            d.spill_base = q * spill_stride;
            a.prod_start[q] = (uint32_t)prod_blocks;
            prod_blocks += (uint64_t)rows[q] * tiles;
            prod_rows += rows[q];
            limb_writes += (uint64_t)rows[q] * d.Lc;
        }
        if (fwd_blocks > 0x7fffffffull || prod_blocks > 0x7fffffffull) return hipErrorInvalidValue;
        if (j.ragged) {
            a.rows = j.rows;
            a.bset = j.bset;
            for (uint32_t i = 0; i < rows[0]; ++i) limb_reads += j.host_rows[i].L_a;
            for (uint32_t b = 0; b < j.nb; ++b) limb_reads += j.host_bset[b].L;
        }
        // algorithmic bytes: 16 B per residue pair (one binary64 per prime),
        // 4 B per limb
        const double fwd_tb = 16.0 * (double)(fwd_blocks / tiles) * (double)K;
        const double prod_tb = 16.0 * (double)prod_rows * (double)K;
        // a column level of 2^lm points runs 2^col_pair_log2(lm) tiles per
        // block: block starts and grid in units of such blocks
        const auto pair_log2_for = [&](int kind, int lm) {
            return ((pair_mask >> kind) & 1u) ? col_pair_log2(lm) : 0;
        };
        const auto col_args = [&](int kind, int lm) {
            PassArgs c = a;
            const int pl = pair_log2_for(kind, lm);
            for (int o = 0; o < 3; ++o) c.op_start[o] >>= pl;
            for (int q = 0; q < 2; ++q) c.prod_start[q] >>= pl;
            return c;
        };
        if (fwd_blocks > 0) {
            GPUPI_RETURN_IF(timed(kColFwdLimbs, k, 4.0 * (double)limb_reads + fwd_tb, stream, [&] {
                dispatch_m<kColLo, kColHi>(p.log2M[0], [&](auto mc) {
                    constexpr int M = decltype(mc)::value;
                    with_pair<M>((pair_mask >> kPairFwdLimbs) & 1u, [&](auto pc) {
                        constexpr int pl = ColPair<M, decltype(pc)::value>::kLog2;
                        const uint32_t total = (uint32_t)(fwd_blocks >> pl);
                        hipLaunchKernelGGL((ntt_col_fwd_limbs_kernel<M, decltype(pc)::value>),
                                           dim3(persistent_grid<M, decltype(pc)::value>(total)), dim3(256 << pl), 0,
                                           stream, col_args(kPairFwdLimbs, M), total);
                    });
                });
            }));
            if (p.levels == 3) {
                GPUPI_RETURN_IF(timed(kColFwd, k, 2 * fwd_tb, stream, [&] {
                    dispatch_m<kColLo, kColHi>(p.log2M[1], [&](auto mc) {
                        constexpr int M = decltype(mc)::value;
                        with_pair<M>((pair_mask >> kPairFwd) & 1u, [&](auto pc) {
                            constexpr int pl = ColPair<M, decltype(pc)::value>::kLog2;
                            hipLaunchKernelGGL((ntt_col_fwd_kernel<M, decltype(pc)::value>),
                                               dim3((uint32_t)(fwd_blocks >> pl)), dim3(256 << pl), 0, stream,
                                               col_args(kPairFwd, M), 1);
                        });
                    });
                }));
            }
        }
        if (prod_blocks == 0) continue;
        if (j.n_prods == 2 && j.prods[0].b == j.prods[1].b && p.log2M[p.levels - 1] <= kPairMaxM) {
            // shared B: its last-level transform once per tile for both products
            const uint32_t pair_rows = std::max(rows[0], rows[1]);
            const double pair_tb = 16.0 * (double)pair_rows * (double)K + 2 * prod_tb;
            GPUPI_RETURN_IF(timed(kRowPair, k, pair_tb, stream, [&] {
                dispatch_m<kRowLo, kRowHi>(p.log2M[p.levels - 1], [&](auto mc) {
                    hipLaunchKernelGGL((ntt_row_pair_kernel<decltype(mc)::value>),
                                       dim3((uint32_t)((uint64_t)pair_rows * tiles)), dim3(256), 0, stream, a,
                                       rows[0], rows[1]);
                });
            }));
        } else {
            // a square reads its one operand's planes once
            const double row_tb = (j.prods[0].a == j.prods[0].b ? 2 : 3) * prod_tb;
            GPUPI_RETURN_IF(timed(kRowFused, k, row_tb, stream, [&] {
                dispatch_m<kRowLo, kRowHi>(p.log2M[p.levels - 1], [&](auto mc) {
                    hipLaunchKernelGGL((ntt_row_fused_kernel<decltype(mc)::value>), dim3((uint32_t)prod_blocks),
                                       dim3(256), 0, stream, a);
                });
            }));
        }
        if (p.levels == 3) {
            GPUPI_RETURN_IF(timed(kColInv, k, 2 * prod_tb, stream, [&] {
                dispatch_m<kColLo, kColHi>(p.log2M[1], [&](auto mc) {
                    constexpr int M = decltype(mc)::value;
                    with_pair<M>((pair_mask >> kPairInv) & 1u, [&](auto pc) {
                        constexpr int pl = ColPair<M, decltype(pc)::value>::kLog2;
                        hipLaunchKernelGGL((ntt_col_inv_kernel<M, decltype(pc)::value>),
                                           dim3((uint32_t)(prod_blocks >> pl)), dim3(256 << pl), 0, stream,
                                           col_args(kPairInv, M), 1);
                    });
                });
            }));
        }
        GPUPI_RETURN_IF(timed(kColInvLimbs, k, prod_tb + 4.0 * (double)limb_writes + 8.0 * prod_rows * runs, stream,
                              [&] {
                                  dispatch_m<kColLo, kColHi>(p.log2M[0], [&](auto mc) {
                                      constexpr int M = decltype(mc)::value;
                                      with_pair<M>((pair_mask >> kPairInvLimbs) & 1u, [&](auto pc) {
                                          constexpr int pl = ColPair<M, decltype(pc)::value>::kLog2;
                                          hipLaunchKernelGGL((ntt_col_inv_limbs_kernel<M, decltype(pc)::value>),
                                                             dim3((uint32_t)(prod_blocks >> pl)), dim3(256 << pl),
                                                             0, stream, col_args(kPairInvLimbs, M));
                                      });
                                  });
                              }));
        for (uint32_t q = 0; q < j.n_prods; ++q) {
            if (rows[q] == 0) continue;
            const NttProd &d = a.prods[q];
            GPUPI_RETURN_IF(launch_add_spills(d.C, a.spill + d.spill_base, rows[q], d.Lc, d.sC,
                                              1u << final_run_log2(p), (uint32_t)runs, flags, stream));
        }
        if (j.ragged) {
            bool copies = false;
            for (uint32_t i = 0; i < rows[0] && !copies; ++i) copies = j.host_rows[i].C != nullptr;
            const uint32_t slices = (uint32_t)ceil_div64(K, 1024);
            if ((uint64_t)rows[0] * slices > 0x7fffffffull) return hipErrorInvalidValue;
            if (copies) {
                hipLaunchKernelGGL(ragged_copy_kernel, dim3(rows[0] * slices), dim3(256), 0, stream, staging,
                                   (uint64_t)K, a.rows, slices);
                GPUPI_RETURN_IF(hipGetLastError());
            }
            GPUPI_RETURN_IF(launch_ragged_assemble(staging, K, j.asms, j.n_asm, j.asm_chunks,
                                                   reinterpret_cast<uint8_t *>(ws + l.asm_flags_off), stream));
        }
    }
    return hipSuccess;
}

// A uniform product C = A B of N rows; a broadcast A (stride 0) is swapped
// into the B position so that its transform is computed once, and a square
// (B is A: same rows) transforms its one operand once.
NttJob uniform_job(const uint32_t *A, const uint32_t *B, uint32_t *C, uint32_t N, uint32_t L_a, uint32_t L_b,
                   uint32_t stride_A, uint32_t stride_B, uint32_t stride_C) {
    NttJob j;
    if (A == B && A != nullptr && L_a == L_b && stride_A == stride_B) {
        j.n_ops = 1;
        j.n_prods = 1;
        j.ops[0] = NttJob::Op{A, L_a, stride_A};
        j.prods[0].a = 0;
        j.prods[0].b = 0;
        j.prods[0].C = C;
        j.prods[0].sC = stride_C;
        j.prods[0].N = N;
        j.k = transform_log2(L_a, L_b);
        return j;
    }
    if (stride_A == 0 && stride_B != 0) {
        std::swap(A, B);
        std::swap(L_a, L_b);
        std::swap(stride_A, stride_B);
    }
    j.n_ops = 2;
    j.n_prods = 1;
    j.ops[0] = NttJob::Op{A, L_a, stride_A};
    j.ops[1] = NttJob::Op{B, L_b, stride_B};
    j.prods[0].a = 0;
    j.prods[0].b = 1;
    j.prods[0].C = C;
    j.prods[0].sC = stride_C;
    j.prods[0].N = N;
    j.k = transform_log2(L_a, L_b);
    return j;
}

// Two products sharing B: C1 = A1 B (N1 rows), C2 = A2 B (N2 rows).
NttJob shared_job(const uint32_t *A1, uint32_t L1, uint32_t s1, uint32_t N1, uint32_t *C1, uint32_t sC1,
                  const uint32_t *A2, uint32_t L2, uint32_t s2, uint32_t N2, uint32_t *C2, uint32_t sC2,
                  const uint32_t *B, uint32_t Lb, uint32_t sB) {
    NttJob j;
    j.n_ops = 3;
    j.n_prods = 2;
    j.ops[0] = NttJob::Op{A1, L1, s1};
    j.ops[1] = NttJob::Op{B, Lb, sB};
    j.ops[2] = NttJob::Op{A2, L2, s2};
    j.prods[0].a = 0;
    j.prods[0].b = 1;
    j.prods[0].C = C1;
    j.prods[0].sC = sC1;
    j.prods[0].N = N1;
    j.prods[1].a = 2;
    j.prods[1].b = 1;
    j.prods[1].C = C2;
    j.prods[1].sC = sC2;
    j.prods[1].N = N2;
    j.k = transform_log2(L1, Lb);
    return j;
}

}  // namespace

size_t ntt_workspace_bytes(uint32_t N, uint32_t L_a, uint32_t L_b) {
    if (!ntt_supports(L_a, L_b) || N == 0) return 0;
    // rows at a stride: the most any call of this shape needs
    return job_workspace_bytes(uniform_job(nullptr, nullptr, nullptr, N, L_a, L_b, L_a, L_b, L_a + L_b));
}

hipError_t launch_mul_ntt(const uint32_t *A, const uint32_t *B, uint32_t *C, void *workspace,
                          size_t workspace_bytes, const NTTPrecomputedTables &tables, uint32_t N,
                          uint32_t L_a, uint32_t L_b, uint32_t stride_A, uint32_t stride_B,
                          uint32_t stride_C, hipStream_t stream) {
    if (N == 0) return hipSuccess;
    if (!ntt_supports(L_a, L_b) || tables.roots == nullptr) return hipErrorInvalidValue;
    return run_ntt(uniform_job(A, B, C, N, L_a, L_b, stride_A, stride_B, stride_C), workspace, workspace_bytes,
                   tables, stream);
}

hipError_t launch_mul_ntt_bounded(const uint32_t *A, const uint32_t *B, uint32_t *C, void *workspace,
                                  size_t workspace_bytes, const NTTPrecomputedTables &tables, uint32_t N,
                                  uint32_t L_a, uint32_t L_b, uint32_t stride_A, uint32_t stride_B,
                                  uint32_t stride_C, hipStream_t stream) {
    if (N == 0) return hipSuccess;
    if (!ntt_supports(L_a, L_b) || tables.roots == nullptr) return hipErrorInvalidValue;
    NttJob j = uniform_job(A, B, C, N, L_a, L_b, stride_A, stride_B, stride_C);
    if (make_plan(j.k).levels > 1) {
        // the most rows per chunk whose planes, spills and flags fit
        const JobShape sh = shape_of(j);
        size_t nc = chunk_products(sh.rows, j.k);
        for (;;) {
            j.max_chunk = nc;
            if (layout_of(j, sh).total <= workspace_bytes) break;
            if (nc == 1) return hipErrorInvalidValue;  // one row does not fit
            nc = (nc + 1) / 2;
        }
    }
    const bool prev = t_stats_off;
    t_stats_off = true;
    const hipError_t err = run_ntt(j, workspace, workspace_bytes, tables, stream);
    t_stats_off = prev;
    return err;
}

bool ntt_shared_supports(uint32_t L1, uint32_t L2, uint32_t Lb) {
    if (!ntt_supports(L1, Lb) || !ntt_supports(L2, Lb)) return false;
    const int k = transform_log2(L1, Lb);
    return k == transform_log2(L2, Lb) && make_plan(k).levels > 1;
}

size_t ntt_shared_workspace_bytes(uint32_t L1, uint32_t N1, uint32_t L2, uint32_t N2, uint32_t Lb, uint32_t sB) {
    if (!ntt_shared_supports(L1, L2, Lb)) return 0;
    return job_workspace_bytes(shared_job(nullptr, L1, L1, N1, nullptr, L1 + Lb, nullptr, L2, L2, N2, nullptr,
                                          L2 + Lb, nullptr, Lb, sB));
}

hipError_t launch_mul_ntt_shared(const uint32_t *A1, uint32_t L1, uint32_t s1, uint32_t N1, uint32_t *C1,
                                 uint32_t sC1, const uint32_t *A2, uint32_t L2, uint32_t s2, uint32_t N2,
                                 uint32_t *C2, uint32_t sC2, const uint32_t *B, uint32_t Lb, uint32_t sB,
                                 void *workspace, size_t workspace_bytes, const NTTPrecomputedTables &tables,
                                 hipStream_t stream) {
    if (!ntt_shared_supports(L1, L2, Lb) || tables.roots == nullptr) return hipErrorInvalidValue;
    return run_ntt(shared_job(A1, L1, s1, N1, C1, sC1, A2, L2, s2, N2, C2, sC2, B, Lb, sB), workspace,
                   workspace_bytes, tables, stream);
}

int ntt_transform_log2(uint32_t L_a, uint32_t L_b) { return transform_log2(L_a, L_b); }

namespace {
NttJob ragged_job(const RaggedRow *rows, const RaggedRow *host_rows, uint32_t N, const RaggedB *bset,
                  const RaggedB *host_bset, uint32_t nb, const RaggedAssembly *asms,
                  const RaggedAssembly *host_asms, uint32_t n_asm, uint32_t asm_chunks, int log2K) {
    NttJob j;
    j.n_ops = 2;
    j.n_prods = 1;
    j.prods[0].N = N;
    j.ragged = true;
    j.rows = rows;
    j.host_rows = host_rows;
    j.bset = bset;
    j.host_bset = host_bset;
    j.nb = nb;
    j.asms = asms;
    j.host_asms = host_asms;
    j.n_asm = n_asm;
    j.asm_chunks = asm_chunks;
    j.k = log2K;
    return j;
}

// A launch's rows, B set and assemblies agree: every row inside its
// transform and B set, assembled rows only above one tile, each assembly's
// pieces rows without C in its range, chunks numbered in order.
bool ragged_ok(const RaggedRow *rows, uint32_t N, const RaggedB *bset, uint32_t nb, const RaggedAssembly *asms,
               uint32_t n_asm, uint32_t chunks, int log2K) {
    const uint64_t K = 1ull << log2K;
    for (uint32_t i = 0; i < N; ++i) {
        const RaggedRow &r = rows[i];
        if (r.b >= nb || (uint64_t)r.L_a + r.L_b > K || bset[r.b].L != r.L_b) return false;
        if (r.C == nullptr && log2K <= kTileLog2) return false;
    }
    uint32_t next = 0;
    for (uint32_t q = 0; q < n_asm; ++q) {
        const RaggedAssembly &x = asms[q];
        if (x.C == nullptr || x.m < 2 || x.s == 0 || x.S == 0 || x.chunk0 != next || x.row0 > N || x.m > N - x.row0)
            return false;
        if (x.m > 2 && x.s < x.S) return false;  // pieces of one parity would overlap
        const uint64_t last = (uint64_t)x.c - x.S - (uint64_t)(x.m - 1) * x.s;
        if ((uint64_t)x.c < (uint64_t)x.S + (uint64_t)(x.m - 1) * x.s + 1 || last > x.s) return false;
        for (uint32_t i = 0; i < x.m; ++i) {
            const RaggedRow &r = rows[x.row0 + i];
            if (r.C != nullptr || r.L_b != x.S || r.L_a != (i + 1 < x.m ? x.s : (uint32_t)last)) return false;
        }
        next += assembly_chunks(x.c);
    }
    return next == chunks;
}
}  // namespace

uint32_t ntt_ragged_max_rows(int log2K) { return ragged_max_rows(log2K); }

namespace {
// The rows of a product in pieces (piece_row), its one B entry and its
// assembly, written on the device (no host upload).
__global__ void __launch_bounds__(256) piece_rows_kernel(RaggedRow *rows, RaggedB *bset, RaggedAssembly *as,
                                                         const uint32_t *L, uint32_t LL, const uint32_t *S,
                                                         uint32_t LS, uint32_t *C, uint32_t m, uint32_t s) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i == 0) {
        bset[0] = RaggedB{S, LS};
        as[0] = RaggedAssembly{C, LL + LS, s, LS, m, 0, 0};
    }
    if (i < m) rows[i] = piece_row(i, L, LL, S, LS, m, s, 0);
}

struct PiecesLayout {
    size_t rows_bytes, bset_bytes, asm_bytes, ntt_bytes;
    uint32_t chunks;
    int k;
};
PiecesLayout pieces_layout(uint32_t LL, uint32_t LS, uint32_t m, uint32_t s) {
    PiecesLayout l;
    l.k = transform_log2(s, LS);
    l.chunks = assembly_chunks(LL + LS);
    l.rows_bytes = align256((size_t)m * sizeof(RaggedRow));
    l.bset_bytes = align256(sizeof(RaggedB));
    l.asm_bytes = align256(sizeof(RaggedAssembly));
    l.ntt_bytes =
        job_workspace_bytes(ragged_job(nullptr, nullptr, m, nullptr, nullptr, 1, nullptr, nullptr, 1, l.chunks, l.k));
    return l;
}
}  // namespace

size_t ntt_pieces_workspace_bytes(uint32_t LL, uint32_t LS, uint32_t m, uint32_t s) {
    const PiecesLayout l = pieces_layout(LL, LS, m, s);
    return l.rows_bytes + l.bset_bytes + l.asm_bytes + l.ntt_bytes;
}

hipError_t launch_mul_ntt_pieces(const uint32_t *L, uint32_t LL, const uint32_t *S, uint32_t LS, uint32_t *C,
                                 uint32_t m, uint32_t s, void *workspace, size_t workspace_bytes,
                                 const NTTPrecomputedTables &tables, hipStream_t stream) {
    if (L == nullptr || S == nullptr || C == nullptr || m < 2 || s == 0 || LS == 0 || (m > 2 && s < LS) ||
        (uint64_t)(m - 1) * s >= LL || (uint64_t)LL - (uint64_t)(m - 1) * s > s ||
        (uint64_t)LL + LS > 0xffffffffull || tables.roots == nullptr)
        return hipErrorInvalidValue;
    const PiecesLayout l = pieces_layout(LL, LS, m, s);
    if (l.k <= kTileLog2 || l.k > kMaxLog2 || workspace == nullptr ||
        workspace_bytes < l.rows_bytes + l.bset_bytes + l.asm_bytes + l.ntt_bytes)
        return hipErrorInvalidValue;
    char *ws = static_cast<char *>(workspace);
    auto *rows = reinterpret_cast<RaggedRow *>(ws);
    auto *bset = reinterpret_cast<RaggedB *>(ws + l.rows_bytes);
    auto *as = reinterpret_cast<RaggedAssembly *>(ws + l.rows_bytes + l.bset_bytes);
    hipLaunchKernelGGL(piece_rows_kernel, dim3((m + 255) / 256), dim3(256), 0, stream, rows, bset, as, L, LL, S, LS,
                       C, m, s);
    GPUPI_RETURN_IF(hipGetLastError());
    // the same descriptors on the host, for the launch's bookkeeping
    std::vector<RaggedRow> host_rows(m);
    for (uint32_t i = 0; i < m; ++i) host_rows[i] = piece_row(i, L, LL, S, LS, m, s, 0);
    const RaggedB host_b{S, LS};
    const RaggedAssembly host_as{C, LL + LS, s, LS, m, 0, 0};
    return run_ntt(ragged_job(rows, host_rows.data(), m, bset, &host_b, 1, as, &host_as, 1, l.chunks, l.k),
                   ws + l.rows_bytes + l.bset_bytes + l.asm_bytes, l.ntt_bytes, tables, stream);
}

size_t ntt_ragged_workspace_bytes(uint32_t N, uint32_t nb, uint32_t chunks, int log2K) {
    return job_workspace_bytes(
        ragged_job(nullptr, nullptr, N, nullptr, nullptr, nb, nullptr, nullptr, chunks ? 1u : 0u, chunks, log2K));
}

hipError_t launch_mul_ntt_ragged(const RaggedRow *rows, const RaggedRow *host_rows, uint32_t N, const RaggedB *bset,
                                 const RaggedB *host_bset, uint32_t nb, const RaggedAssembly *asms,
                                 const RaggedAssembly *host_asms, uint32_t n_asm, uint32_t chunks, int log2K,
                                 void *workspace, size_t workspace_bytes, const NTTPrecomputedTables &tables,
                                 hipStream_t stream) {
    if (N == 0) return hipSuccess;
    if (tables.roots == nullptr || log2K < kMinLog2 || log2K > kMaxLog2 || rows == nullptr || host_rows == nullptr ||
        bset == nullptr || host_bset == nullptr || (n_asm > 0 && (asms == nullptr || host_asms == nullptr)) ||
        !ragged_ok(host_rows, N, host_bset, nb, host_asms, n_asm, chunks, log2K))
        return hipErrorInvalidValue;
    return run_ntt(ragged_job(rows, host_rows, N, bset, host_bset, nb, asms, host_asms, n_asm, chunks, log2K),
                   workspace, workspace_bytes, tables, stream);
}

}  // namespace gpupi

// ---------------------------------------------------------------------------
// C entry points (reference batch_mul_ntt.h)
// ---------------------------------------------------------------------------
extern "C" {

void init_ntt_precomputed_tables(NTTPrecomputedTables *tables) {
    if (tables == nullptr) return;
    if (gpupi::ntt_tables_init(tables) != hipSuccess) {
        tables->roots = nullptr;
        tables->max_log2 = 0;
    }
}

void release_ntt_precomputed_tables(NTTPrecomputedTables *tables) {
    if (tables != nullptr) gpupi::ntt_tables_release(tables);
}

void batch_mul_ntt(uint32_t *A, uint32_t *B, uint32_t *ret, uint32_t *workspace, NTTPrecomputedTables tables,
                   uint32_t N, uint32_t L_a, uint32_t L_b, uint32_t stride_A, uint32_t stride_B,
                   uint32_t stride_ret) {
    (void)gpupi::launch_mul_ntt(A, B, ret, workspace, gpupi::ntt_workspace_bytes(N, L_a, L_b), tables, N, L_a,
                                L_b, stride_A, stride_B, stride_ret, 0);
}

size_t batch_mul_ntt_workspace_size(uint32_t N, uint32_t L_a, uint32_t L_b) {
    return gpupi::ntt_workspace_bytes(N, L_a, L_b);
}

void gpupi_kernel_stats_reset(int enable) {
    using namespace gpupi;
    std::lock_guard<std::mutex> lk(g_stats.mu);
    for (auto &r : g_stats.records) {
        g_stats.pool.push_back(r.start);
        g_stats.pool.push_back(r.stop);
    }
    g_stats.records.clear();
    g_stats.enabled = enable != 0;
}

namespace {
// Aggregates the records per kernel kind, or per (kind, log2 K).
int stats_query(GpupiKernelStat *out, int max_entries, bool by_size) {
    using namespace gpupi;
    constexpr int kSizes = ntt::kMaxLog2 + 1;
    std::vector<GpupiKernelStat> agg((size_t)kKindCount * kSizes);
    std::lock_guard<std::mutex> lk(g_stats.mu);
    for (auto &r : g_stats.records) {
        if (hipEventSynchronize(r.stop) != hipSuccess) return -1;
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, r.start, r.stop) != hipSuccess) return -1;
        GpupiKernelStat &a = agg[(size_t)r.kind * kSizes + (by_size ? r.log2K : 0)];
        a.launches += 1;
        a.total_ms += ms;
        a.total_bytes += r.bytes;
    }
    int n = 0;
    for (int k = 0; k < kKindCount; ++k)
        for (int s = 0; s < kSizes; ++s) {
            GpupiKernelStat a = agg[(size_t)k * kSizes + s];
            if (a.launches == 0) continue;
            if (n == max_entries) return n;
            if (by_size) std::snprintf(a.name, sizeof(a.name), "%s@2^%d", kKindNames[k], s);
            else std::snprintf(a.name, sizeof(a.name), "%s", kKindNames[k]);
            out[n++] = a;
        }
    return n;
}
}  // namespace

int gpupi_kernel_stats_query(GpupiKernelStat *out, int max_entries) { return stats_query(out, max_entries, false); }

int gpupi_kernel_stats_query_sizes(GpupiKernelStat *out, int max_entries) {
    return stats_query(out, max_entries, true);
}

hipError_t gpupi_ntt_set_split(int log2_k, int log2_m0, int log2_m1) {
    using namespace gpupi;
    using namespace gpupi::ntt;
    if (log2_k < kMinLog2 || log2_k > kMaxLog2 || log2_m0 < 0 || log2_m1 < 0 || log2_m0 > 255 || log2_m1 > 255)
        return hipErrorInvalidValue;
    if (log2_m0 == 0) {
        if (log2_m1 != 0) return hipErrorInvalidValue;
        g_split[log2_k].store(0);
        return hipSuccess;
    }
    const Plan p = make_plan_split(log2_k, log2_m0, log2_m1);
    if (p.levels == 1 || !plan_ok(p) || (p.levels == 2 && log2_m1 != 0)) return hipErrorInvalidValue;
    g_split[log2_k].store((uint32_t)log2_m0 | (uint32_t)log2_m1 << 8);
    return hipSuccess;
}

hipError_t gpupi_ntt_set_launch(uint32_t pair_mask, int xcd_chunk_log2) {
    using namespace gpupi::ntt;
    if (pair_mask > 15u || xcd_chunk_log2 < 0 || xcd_chunk_log2 > kXcdChunkLog2Max) return hipErrorInvalidValue;
    gpupi::g_pair_mask.store(pair_mask);
    gpupi::g_xcd_log2.store(xcd_chunk_log2);
    return hipSuccess;
}

void gpupi_ntt_get_launch(uint32_t *pair_mask, int *xcd_chunk_log2) {
    if (pair_mask) *pair_mask = gpupi::g_pair_mask.load();
    if (xcd_chunk_log2) *xcd_chunk_log2 = gpupi::g_xcd_log2.load();
}

int gpupi_ntt_get_split(int log2_k, int *log2_m) {
    using namespace gpupi;
    using namespace gpupi::ntt;
    if (log2_k < kMinLog2 || log2_k > kMaxLog2 || log2_m == nullptr) return 0;
    const Plan p = active_plan(log2_k);
    for (int l = 0; l < kMaxLevels; ++l) log2_m[l] = l < p.levels ? p.log2M[l] : 0;
    return p.levels;
}

}  // extern "C"
