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
// is needed inside butterflies.
//
// Each kernel works on a 4096-element tile in LDS (radix-16 steps, bit-
// reversed slot order inside the tile); long transforms are split into
// column levels (<= 512 points, >= 8 adjacent columns per tile so HBM
// accesses stay in >= 128-byte runs) and a final contiguous level.  The last
// forward level, the pointwise product and the first inverse level are fused
// in one kernel; the final inverse level applies 1/K, recombines the primes
// (Garner) and emits limbs with carries resolved inside runs of adjacent
// coefficients, leaving one 64-bit spill word per run for the carry pass.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <vector>

#include "common.hpp"
#include "gpupi_pi.h"
#include "ntt_plan.hpp"

namespace gpupi {
namespace {

using namespace ntt;

constexpr uint64_t kPrime0 = 0x1ffffc0000001ull;  // 562948879679489 = 524287 * 2^30 + 1
constexpr uint64_t kPrime1 = 0x1fffe00000001ull;  // 562941363486721 = 524280 * 2^30 + 1 (2^33 | p - 1)
constexpr uint64_t kGen0 = 3, kGen1 = 13;          // primitive roots
constexpr double kP0 = 562948879679489.0;
constexpr double kP1 = 562941363486721.0;
constexpr double kPinv0 = 1.0 / kP0;
constexpr double kPinv1 = 1.0 / kP1;
constexpr double kRoundMagic = 6755399441055744.0;  // 1.5 * 2^52

// Table layout (double2 = (prime0, prime1) pairs) inside tables.roots.
constexpr size_t kOffTw = 0;                // w_4096^e, e < 4096
constexpr size_t kOffTwInv = 4096;          // w_4096^-e
constexpr size_t kOffLo = 8192;             // w_2^30^e, e < 2^15
constexpr size_t kOffHi = kOffLo + 32768;   // w_2^30^(e 2^15)
constexpr size_t kOffLoInv = kOffHi + 32768;
constexpr size_t kOffHiInv = kOffLoInv + 32768;
constexpr size_t kTableLen = kOffHiInv + 32768;

struct NttConsts {
    double2 w16[8];
    double2 w16i[8];
    double2 kinv[kMaxLog2 + 1];
    double garner;  // p0^-1 mod p1, centred
};
__constant__ NttConsts c_ntt;

__device__ __forceinline__ double qround(double x, double pinv) {
    return __builtin_fma(x, pinv, kRoundMagic) - kRoundMagic;
}
__device__ __forceinline__ double redp(double x, double p, double pinv) {
    return __builtin_fma(-qround(x, pinv), p, x);
}
__device__ __forceinline__ double mulp(double a, double b, double p, double pinv) {
    const double h = a * b;
    const double l = __builtin_fma(a, b, -h);
    return __builtin_fma(-qround(h, pinv), p, h) + l;
}
__device__ __forceinline__ double2 add2(double2 u, double2 v) {
    return make_double2(redp(u.x + v.x, kP0, kPinv0), redp(u.y + v.y, kP1, kPinv1));
}
__device__ __forceinline__ double2 sub2(double2 u, double2 v) {
    return make_double2(redp(u.x - v.x, kP0, kPinv0), redp(u.y - v.y, kP1, kPinv1));
}
__device__ __forceinline__ double2 mul2(double2 u, double2 w) {
    return make_double2(mulp(u.x, w.x, kP0, kPinv0), mulp(u.y, w.y, kP1, kPinv1));
}
__device__ __forceinline__ double2 submul2(double2 u, double2 v, double2 w) {
    return make_double2(mulp(u.x - v.x, w.x, kP0, kPinv0), mulp(u.y - v.y, w.y, kP1, kPinv1));
}

__device__ __forceinline__ uint32_t brev(uint32_t x, int bits) {
    return bits == 0 ? 0u : (__brev(x) >> (32 - bits));
}

// In-register radix-R DIF (natural in, bit-reversed out) and its inverse DIT.
template <int R>
__device__ __forceinline__ void dif(double2 *v) {
#pragma unroll
    for (int h = R / 2; h >= 1; h >>= 1) {
#pragma unroll
        for (int j = 0; j < R; ++j) {
            if (j & h) continue;
            const int e = (j & (h - 1)) * (8 / h);
            const double2 a = v[j], b = v[j + h];
            v[j] = add2(a, b);
            v[j + h] = (e == 0) ? sub2(a, b) : submul2(a, b, c_ntt.w16[e]);
        }
    }
}

template <int R>
__device__ __forceinline__ void dit(double2 *v) {
#pragma unroll
    for (int h = 1; h < R; h <<= 1) {
#pragma unroll
        for (int j = 0; j < R; ++j) {
            if (j & h) continue;
            const int e = (j & (h - 1)) * (8 / h);
            const double2 t = (e == 0) ? v[j + h] : mul2(v[j + h], c_ntt.w16i[e]);
            const double2 a = v[j];
            v[j] = add2(a, t);
            v[j + h] = sub2(a, t);
        }
    }
}

// One radix-R step over the tile (4096 / R small DFTs, 16 / R per thread).
template <int R, bool INV>
__device__ __forceinline__ void tile_step(double2 *lds, int m, int s, const double2 *tw) {
    constexpr int PER = 16 / R;
    constexpr int LR = (R == 16) ? 4 : (R == 8) ? 3 : (R == 4) ? 2 : 1;
    double2 v[16];
    SmallDft d[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        d[u] = small_dft(m, s, threadIdx.x + 256u * u);
#pragma unroll
        for (int j = 0; j < R; ++j) v[u * R + j] = lds[lds_index(d[u].base + j * d[u].stride)];
    }
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int shift = kTileLog2 - d[u].log2Ms;
        if (!INV) {
            dif<R>(&v[u * R]);
#pragma unroll
            for (int j = 1; j < R; ++j) {
                const uint32_t e = d[u].n2 * brev((uint32_t)j, LR);
                v[u * R + j] = mul2(v[u * R + j], tw[e << shift]);
            }
        } else {
#pragma unroll
            for (int j = 1; j < R; ++j) {
                const uint32_t e = d[u].n2 * brev((uint32_t)j, LR);
                v[u * R + j] = mul2(v[u * R + j], tw[e << shift]);
            }
            dit<R>(&v[u * R]);
        }
    }
#pragma unroll
    for (int u = 0; u < PER; ++u) {
#pragma unroll
        for (int j = 0; j < R; ++j) lds[lds_index(d[u].base + j * d[u].stride)] = v[u * R + j];
    }
}

// All steps of G = 4096 / 2^m DFTs of 2^m points held in the tile.
template <bool INV>
__device__ void tile_dft(double2 *lds, int m, const double2 *tw) {
    const int steps = step_count(m);
    for (int i = 0; i < steps; ++i) {
        const int s = INV ? steps - 1 - i : i;
        switch (step_log2radix(m, s)) {
            case 4: tile_step<16, INV>(lds, m, s, tw); break;
            case 3: tile_step<8, INV>(lds, m, s, tw); break;
            case 2: tile_step<4, INV>(lds, m, s, tw); break;
            default: tile_step<2, INV>(lds, m, s, tw); break;
        }
        __syncthreads();
    }
}

// w_{2^30}^E (or its inverse) from the two-level table.
__device__ __forceinline__ double2 big_twiddle(const double2 *tables, uint64_t E, bool inv) {
    const double2 lo = tables[(inv ? kOffLoInv : kOffLo) + (E & 32767u)];
    const double2 hi = tables[(inv ? kOffHiInv : kOffHi) + (E >> 15)];
    return mul2(lo, hi);
}

// Garner recombination of centred residues -> exact coefficient (< 2^98).
__device__ __forceinline__ unsigned __int128 recombine(double2 v) {
    const double r0 = v.x < 0.0 ? v.x + kP0 : v.x;
    const double r1 = v.y < 0.0 ? v.y + kP1 : v.y;
    double t = mulp(r1 - r0, c_ntt.garner, kP1, kPinv1);
    t = t < 0.0 ? t + kP1 : t;
    return (unsigned __int128)kPrime0 * (uint64_t)t + (uint64_t)r0;
}

struct PassArgs {
    const uint32_t *A;
    const uint32_t *B;
    uint32_t La, Lb, sA, sB;
    double2 *TA;
    double2 *TB;
    uint32_t *C;
    uint32_t Lc, sC;
    uint64_t *spill;
    uint32_t N;  // products in this launch
    Plan plan;
    const double2 *tables;
};

// Forward column level `level`.  Level 0 reads limbs; blocks cover
// (product, operand, tile).  The inter-level twiddle is applied on store.
template <bool FROM_LIMBS>
__global__ void __launch_bounds__(256) ntt_col_fwd_kernel(PassArgs a, int level) {
    __shared__ double2 lds[kLdsSlots];
    const Plan &p = a.plan;
    const uint32_t tiles = tiles_per_product(p);
    const uint32_t t = blockIdx.x % tiles;
    const uint32_t op = (blockIdx.x / tiles) & 1u;
    const uint32_t prod = blockIdx.x / (2u * tiles);
    const int lm = p.log2M[level];
    const int lc = kTileLog2 - lm;
    const int lk = slab_log2(p, level);
    const uint64_t K = 1ull << p.log2K;
    const ColTile ct = col_tile(p, level, t);
    double2 *T = (op ? a.TB : a.TA) + (size_t)prod * K;
    const uint32_t *X = op ? a.B : a.A;
    const uint32_t LX = op ? a.Lb : a.La;
    const uint32_t sX = op ? a.sB : a.sA;
#pragma unroll 4
    for (int u = 0; u < 16; ++u) {
        const uint32_t i = threadIdx.x + 256u * u;
        const uint32_t g = i & ((1u << lc) - 1u), r = i >> lc;
        const uint64_t off = col_offset(p, level, ct, g, r);
        double2 v;
        if (FROM_LIMBS) {
            const double x = off < LX ? (double)X[(size_t)prod * sX + off] : 0.0;
            v = make_double2(x, x);
        } else {
            v = T[off];
        }
        lds[lds_index((g << lm) + r)] = v;
    }
    __syncthreads();
    tile_dft<false>(lds, lm, a.tables + kOffTw);
#pragma unroll 4
    for (int u = 0; u < 16; ++u) {
        const uint32_t i = threadIdx.x + 256u * u;
        const uint32_t g = i & ((1u << lc) - 1u), r = i >> lc;
        const uint64_t off = col_offset(p, level, ct, g, r);
        const uint64_t e = (uint64_t)(ct.col0 + g) * brev(r, lm);
        const double2 w = big_twiddle(a.tables, e << (kMaxLog2 - lk), false);
        T[off] = mul2(lds[lds_index((g << lm) + r)], w);
    }
}

// Last level: forward on A and B tiles, pointwise product, inverse.
__global__ void __launch_bounds__(256) ntt_row_fused_kernel(PassArgs a) {
    extern __shared__ double2 dyn[];
    double2 *la = dyn;
    double2 *lb = dyn + kLdsSlots;
    const Plan &p = a.plan;
    const uint32_t tiles = tiles_per_product(p);
    const uint32_t t = blockIdx.x % tiles;
    const uint32_t prod = blockIdx.x / tiles;
    const int lm = p.log2M[p.levels - 1];
    const uint64_t K = 1ull << p.log2K;
    double2 *TA = a.TA + (size_t)prod * K;
    const double2 *TB = a.TB + (size_t)prod * K;
#pragma unroll 4
    for (int u = 0; u < 16; ++u) {
        const uint32_t i = threadIdx.x + 256u * u;
        const uint64_t off = row_offset(p, t, i >> lm, i & ((1u << lm) - 1u));
        la[lds_index(i)] = TA[off];
        lb[lds_index(i)] = TB[off];
    }
    __syncthreads();
    tile_dft<false>(la, lm, a.tables + kOffTw);
    tile_dft<false>(lb, lm, a.tables + kOffTw);
#pragma unroll 4
    for (int u = 0; u < 16; ++u) {
        const uint32_t i = lds_index(threadIdx.x + 256u * u);
        la[i] = mul2(la[i], lb[i]);
    }
    __syncthreads();
    tile_dft<true>(la, lm, a.tables + kOffTwInv);
#pragma unroll 4
    for (int u = 0; u < 16; ++u) {
        const uint32_t i = threadIdx.x + 256u * u;
        const uint64_t off = row_offset(p, t, i >> lm, i & ((1u << lm) - 1u));
        TA[off] = la[lds_index(i)];
    }
}

// Inverse column level.  Level 0 finishes the product: 1/K, Garner, limbs
// of each run of C adjacent coefficients, and the run's spill word.
template <bool TO_LIMBS>
__global__ void __launch_bounds__(256) ntt_col_inv_kernel(PassArgs a, int level) {
    __shared__ double2 lds[kLdsSlots];
    const Plan &p = a.plan;
    const uint32_t tiles = tiles_per_product(p);
    const uint32_t t = blockIdx.x % tiles;
    const uint32_t prod = blockIdx.x / tiles;
    const int lm = p.log2M[level];
    const int lc = kTileLog2 - lm;
    const int lk = slab_log2(p, level);
    const uint64_t K = 1ull << p.log2K;
    const ColTile ct = col_tile(p, level, t);
    double2 *T = a.TA + (size_t)prod * K;
#pragma unroll 4
    for (int u = 0; u < 16; ++u) {
        const uint32_t i = threadIdx.x + 256u * u;
        const uint32_t g = i & ((1u << lc) - 1u), r = i >> lc;
        const uint64_t off = col_offset(p, level, ct, g, r);
        const uint64_t e = (uint64_t)(ct.col0 + g) * brev(r, lm);
        const double2 w = big_twiddle(a.tables, e << (kMaxLog2 - lk), true);
        lds[lds_index((g << lm) + r)] = mul2(T[off], w);
    }
    __syncthreads();
    tile_dft<true>(lds, lm, a.tables + kOffTwInv);
    if (!TO_LIMBS) {
#pragma unroll 4
        for (int u = 0; u < 16; ++u) {
            const uint32_t i = threadIdx.x + 256u * u;
            const uint32_t g = i & ((1u << lc) - 1u), r = i >> lc;
            T[col_offset(p, level, ct, g, r)] = lds[lds_index((g << lm) + r)];
        }
        return;
    }
    // level 0: slab 0 covers the product; run r = coefficients r*S + col0 + [0, C)
    const double2 kinv = c_ntt.kinv[p.log2K];
    const uint32_t C = 1u << lc;
    const int ls = lk - lm;  // log2 S
    uint32_t *out = a.C + (size_t)prod * a.sC;
    for (uint32_t r = threadIdx.x; r < (1u << lm); r += 256u) {
        const uint64_t n0 = ((uint64_t)r << ls) + ct.col0;
        unsigned __int128 pending = 0;
        for (uint32_t g = 0; g < C; ++g) {
            pending += recombine(mul2(lds[lds_index((g << lm) + r)], kinv));
            const uint64_t n = n0 + g;
            if (n < a.Lc) out[n] = (uint32_t)pending;
            pending >>= 32;
        }
        a.spill[(size_t)prod * (K >> lc) + (n0 >> lc)] = (uint64_t)pending;
    }
}

// K <= 4096: whole products inside one tile (G = 4096 / K per block).
__global__ void __launch_bounds__(256) ntt_local_kernel(PassArgs a) {
    extern __shared__ double2 dyn[];
    __shared__ uint64_t spill_s[256];
    __shared__ uint32_t wave_flags[4];
    double2 *la = dyn;
    double2 *lb = dyn + kLdsSlots;
    const int lm = a.plan.log2K;
    const uint32_t G = kTile >> lm;
    const uint32_t prod0 = blockIdx.x * G;
#pragma unroll 4
    for (int u = 0; u < 16; ++u) {
        const uint32_t i = threadIdx.x + 256u * u;
        const uint32_t prod = prod0 + (i >> lm), r = i & ((1u << lm) - 1u);
        double xa = 0.0, xb = 0.0;
        if (prod < a.N) {
            if (r < a.La) xa = (double)a.A[(size_t)prod * a.sA + r];
            if (r < a.Lb) xb = (double)a.B[(size_t)prod * a.sB + r];
        }
        la[lds_index(i)] = make_double2(xa, xa);
        lb[lds_index(i)] = make_double2(xb, xb);
    }
    __syncthreads();
    tile_dft<false>(la, lm, a.tables + kOffTw);
    tile_dft<false>(lb, lm, a.tables + kOffTw);
#pragma unroll 4
    for (int u = 0; u < 16; ++u) {
        const uint32_t i = lds_index(threadIdx.x + 256u * u);
        la[i] = mul2(la[i], lb[i]);
    }
    __syncthreads();
    tile_dft<true>(la, lm, a.tables + kOffTwInv);

    // each thread: 16 consecutive coefficients of one product
    const double2 kinv = c_ntt.kinv[lm];
    const uint32_t i0 = threadIdx.x * 16u;
    const uint32_t prod = prod0 + (i0 >> lm);
    const uint32_t r0 = i0 & ((1u << lm) - 1u);
    uint32_t limb[16];
    unsigned __int128 pending = 0;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        pending += recombine(mul2(la[lds_index(i0 + j)], kinv));
        limb[j] = (uint32_t)pending;
        pending >>= 32;
    }
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
    const bool last_of_product = (r0 + 16u) == (1u << lm);
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
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const uint64_t x = (uint64_t)limb[j] + ci;
            ci = (uint32_t)(x >> 32);
            if (r0 + j < a.Lc) out[r0 + j] = (uint32_t)x;
        }
    }
}

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

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

bool g_attr_done = false;

// Optional per-launch timing of the NTT kernels (HIP events on the launch
// stream) with each launch's algorithmic bytes, for live roofline numbers.
enum KernelKind { kColFwdLimbs, kColFwd, kRowFused, kColInv, kColInvLimbs, kLocal, kKindCount };
const char *const kKindNames[kKindCount] = {"ntt_col_fwd_kernel<limbs>", "ntt_col_fwd_kernel",
                                            "ntt_row_fused_kernel", "ntt_col_inv_kernel",
                                            "ntt_col_inv_kernel<limbs>", "ntt_local_kernel"};
struct StatRecord {
    int kind;
    hipEvent_t start, stop;
    double bytes;
};
struct Stats {
    bool enabled = false;
    std::vector<StatRecord> records;
    std::vector<hipEvent_t> pool;
} g_stats;

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

// Records `launch` between two events when stats are enabled.
template <typename F>
hipError_t timed(int kind, double bytes, hipStream_t stream, F &&launch) {
    if (!g_stats.enabled) {
        launch();
        return hipGetLastError();
    }
    StatRecord r{kind, stats_event(), stats_event(), bytes};
    if (r.start == nullptr || r.stop == nullptr) return hipErrorOutOfMemory;
    GPUPI_RETURN_IF(hipEventRecord(r.start, stream));
    launch();
    GPUPI_RETURN_IF(hipGetLastError());
    GPUPI_RETURN_IF(hipEventRecord(r.stop, stream));
    g_stats.records.push_back(r);
    return hipSuccess;
}

}  // namespace

bool ntt_supports(uint32_t L_a, uint32_t L_b) {
    return L_a > 0 && L_b > 0 && transform_log2(L_a, L_b) <= kMaxLog2;
}

size_t ntt_workspace_bytes(uint32_t N, uint32_t L_a, uint32_t L_b) {
    if (!ntt_supports(L_a, L_b) || N == 0) return 0;
    const int k = transform_log2(L_a, L_b);
    const Plan p = make_plan(k);
    if (p.levels == 1) return 0;
    const size_t nc = chunk_products(N, k);
    const size_t K = size_t(1) << k;
    const size_t runs = K >> final_run_log2(p);
    return align256(nc * K * sizeof(double2)) * 2 + align256(nc * runs * sizeof(uint64_t)) +
           align256(carry_flags_bytes((uint32_t)nc, L_a + L_b));
}

hipError_t ntt_tables_init(NTTPrecomputedTables *tables) {
    const uint64_t primes[2] = {kPrime0, kPrime1};
    const uint64_t gens[2] = {kGen0, kGen1};
    std::vector<double2> host(kTableLen);
    NttConsts cst{};
    for (int q = 0; q < 2; ++q) {
        const uint64_t P = primes[q];
        const uint64_t w30 = powmod_u64(gens[q], (P - 1) >> kMaxLog2, P);
        const uint64_t w30i = powmod_u64(w30, P - 2, P);
        auto put = [&](size_t idx, uint64_t v) {
            double *d = reinterpret_cast<double *>(&host[idx]);
            d[q] = centred(v, P);
        };
        const uint64_t w4096 = powmod_u64(w30, 1ull << (kMaxLog2 - kTileLog2), P);
        const uint64_t w4096i = powmod_u64(w4096, P - 2, P);
        uint64_t x = 1, y = 1;
        for (size_t e = 0; e < 4096; ++e) {
            put(kOffTw + e, x);
            put(kOffTwInv + e, y);
            x = mulmod_u64(x, w4096, P);
            y = mulmod_u64(y, w4096i, P);
        }
        const uint64_t step_hi = powmod_u64(w30, 32768, P), step_hi_i = powmod_u64(w30i, 32768, P);
        uint64_t lo = 1, loi = 1, hi = 1, hii = 1;
        for (size_t e = 0; e < 32768; ++e) {
            put(kOffLo + e, lo);
            put(kOffLoInv + e, loi);
            put(kOffHi + e, hi);
            put(kOffHiInv + e, hii);
            lo = mulmod_u64(lo, w30, P);
            loi = mulmod_u64(loi, w30i, P);
            hi = mulmod_u64(hi, step_hi, P);
            hii = mulmod_u64(hii, step_hi_i, P);
        }
        const uint64_t w16 = powmod_u64(w30, 1ull << (kMaxLog2 - 4), P);
        const uint64_t w16i = powmod_u64(w16, P - 2, P);
        for (int e = 0; e < 8; ++e) {
            reinterpret_cast<double *>(&cst.w16[e])[q] = centred(powmod_u64(w16, e, P), P);
            reinterpret_cast<double *>(&cst.w16i[e])[q] = centred(powmod_u64(w16i, e, P), P);
        }
        for (int k = 0; k <= kMaxLog2; ++k) {
            reinterpret_cast<double *>(&cst.kinv[k])[q] = centred(powmod_u64(powmod_u64(2, k, P), P - 2, P), P);
        }
    }
    cst.garner = centred(powmod_u64(kPrime0 % kPrime1, kPrime1 - 2, kPrime1), kPrime1);

    double2 *d = nullptr;
    GPUPI_RETURN_IF(hipMalloc(&d, kTableLen * sizeof(double2)));
    hipError_t err = hipMemcpy(d, host.data(), kTableLen * sizeof(double2), hipMemcpyHostToDevice);
    if (err == hipSuccess) err = hipMemcpyToSymbol(HIP_SYMBOL(c_ntt), &cst, sizeof(cst));
    if (err != hipSuccess) {
        (void)hipFree(d);
        return err;
    }
    if (!g_attr_done) {
        const int dyn = (int)(2 * kLdsSlots * sizeof(double2));
        (void)hipFuncSetAttribute((const void *)ntt_row_fused_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, dyn);
        (void)hipFuncSetAttribute((const void *)ntt_local_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, dyn);
        g_attr_done = true;
    }
    tables->roots = reinterpret_cast<double *>(d);
    tables->max_log2 = kMaxLog2;
    return hipSuccess;
}

void ntt_tables_release(NTTPrecomputedTables *tables) {
    if (tables->roots) (void)hipFree(tables->roots);
    tables->roots = nullptr;
    tables->max_log2 = 0;
}

hipError_t launch_mul_ntt(const uint32_t *A, const uint32_t *B, uint32_t *C, void *workspace,
                          size_t workspace_bytes, const NTTPrecomputedTables &tables, uint32_t N,
                          uint32_t L_a, uint32_t L_b, uint32_t stride_A, uint32_t stride_B,
                          uint32_t stride_C, hipStream_t stream) {
    if (N == 0) return hipSuccess;
    if (!ntt_supports(L_a, L_b) || tables.roots == nullptr) return hipErrorInvalidValue;
    const int k = transform_log2(L_a, L_b);
    const Plan p = make_plan(k);
    const uint32_t Lc = L_a + L_b;
    const size_t K = size_t(1) << k;
    const size_t dyn = 2 * kLdsSlots * sizeof(double2);

    PassArgs a{};
    a.La = L_a;
    a.Lb = L_b;
    a.sA = stride_A;
    a.sB = stride_B;
    a.Lc = Lc;
    a.sC = stride_C;
    a.plan = p;
    a.tables = reinterpret_cast<const double2 *>(tables.roots);

    if (p.levels == 1) {
        const uint32_t G = kTile >> k;
        const uint64_t blocks = ceil_div64(N, G);
        if (blocks > 0x7fffffffull) return hipErrorInvalidValue;
        a.A = A;
        a.B = B;
        a.C = C;
        a.N = N;
        const double bytes = 4.0 * N * ((double)L_a + L_b + Lc);
        return timed(kLocal, bytes, stream, [&] {
            hipLaunchKernelGGL(ntt_local_kernel, dim3((uint32_t)blocks), dim3(256), dyn, stream, a);
        });
    }
    if (workspace == nullptr || workspace_bytes < ntt_workspace_bytes(N, L_a, L_b)) return hipErrorInvalidValue;
    const size_t nc = chunk_products(N, k);
    const size_t runs = K >> final_run_log2(p);
    char *ws = static_cast<char *>(workspace);
    a.TA = reinterpret_cast<double2 *>(ws);
    a.TB = reinterpret_cast<double2 *>(ws + align256(nc * K * sizeof(double2)));
    a.spill = reinterpret_cast<uint64_t *>(ws + 2 * align256(nc * K * sizeof(double2)));
    uint8_t *flags = reinterpret_cast<uint8_t *>(ws + 2 * align256(nc * K * sizeof(double2)) +
                                                 align256(nc * runs * sizeof(uint64_t)));
    const uint32_t tiles = tiles_per_product(p);
    for (size_t r0 = 0; r0 < N; r0 += nc) {
        const uint32_t n = (uint32_t)((N - r0) < nc ? (N - r0) : nc);
        a.A = A + r0 * stride_A;
        a.B = B + r0 * stride_B;
        a.C = C + r0 * stride_C;
        a.N = n;
        const uint64_t fwd_blocks = (uint64_t)n * 2 * tiles, blocks = (uint64_t)n * tiles;
        if (fwd_blocks > 0x7fffffffull) return hipErrorInvalidValue;
        // algorithmic bytes: residue pairs are 16 B, limbs 4 B
        const double tb = 16.0 * (double)n * (double)K;
        GPUPI_RETURN_IF(timed(kColFwdLimbs, 4.0 * n * ((double)L_a + L_b) + 2 * tb, stream, [&] {
            hipLaunchKernelGGL(ntt_col_fwd_kernel<true>, dim3((uint32_t)fwd_blocks), dim3(256), 0, stream, a, 0);
        }));
        if (p.levels == 3) {
            GPUPI_RETURN_IF(timed(kColFwd, 4 * tb, stream, [&] {
                hipLaunchKernelGGL(ntt_col_fwd_kernel<false>, dim3((uint32_t)fwd_blocks), dim3(256), 0, stream, a, 1);
            }));
        }
        GPUPI_RETURN_IF(timed(kRowFused, 3 * tb, stream, [&] {
            hipLaunchKernelGGL(ntt_row_fused_kernel, dim3((uint32_t)blocks), dim3(256), dyn, stream, a);
        }));
        if (p.levels == 3) {
            GPUPI_RETURN_IF(timed(kColInv, 2 * tb, stream, [&] {
                hipLaunchKernelGGL(ntt_col_inv_kernel<false>, dim3((uint32_t)blocks), dim3(256), 0, stream, a, 1);
            }));
        }
        GPUPI_RETURN_IF(timed(kColInvLimbs, tb + 4.0 * n * Lc + 8.0 * n * runs, stream, [&] {
            hipLaunchKernelGGL(ntt_col_inv_kernel<true>, dim3((uint32_t)blocks), dim3(256), 0, stream, a, 0);
        }));
        GPUPI_RETURN_IF(launch_add_spills(a.C, a.spill, n, Lc, stride_C, 1u << final_run_log2(p), (uint32_t)runs,
                                          flags, stream));
    }
    return hipSuccess;
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
    for (auto &r : g_stats.records) {
        g_stats.pool.push_back(r.start);
        g_stats.pool.push_back(r.stop);
    }
    g_stats.records.clear();
    g_stats.enabled = enable != 0;
}

int gpupi_kernel_stats_query(GpupiKernelStat *out, int max_entries) {
    using namespace gpupi;
    GpupiKernelStat agg[kKindCount] = {};
    for (int k = 0; k < kKindCount; ++k) {
        std::snprintf(agg[k].name, sizeof(agg[k].name), "%s", kKindNames[k]);
    }
    for (auto &r : g_stats.records) {
        if (hipEventSynchronize(r.stop) != hipSuccess) return -1;
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, r.start, r.stop) != hipSuccess) return -1;
        agg[r.kind].launches += 1;
        agg[r.kind].total_ms += ms;
        agg[r.kind].total_bytes += r.bytes;
    }
    int n = 0;
    for (int k = 0; k < kKindCount && n < max_entries; ++k)
        if (agg[k].launches > 0) out[n++] = agg[k];
    return n;
}

}  // extern "C"
