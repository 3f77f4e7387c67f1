// Carry-propagating batched ops: add, sub, add_small, sub_small, mul_small,
// shift_add, shift_sub, and the spill-add used to finish NTT products.
//
// Every op is "per-limb digit + one carry/borrow chain": limb j of the result
// is d_j +/- c_j where (d_j, generate_j) come from the inputs and c_j is the
// incoming carry.  The chain is resolved wave-wide with two 64-bit ballots
// (generate and propagate masks) and one 64-bit add (resolve_carries), so a
// wave settles 64 limbs per step instead of the reference's 32-lane
// Kogge-Stone scans (reference batch_mul_addsub_warp.h:5-85).
//
// Layout: one wave owns a chunk of SEG*64 consecutive limbs of one row
// (coalesced 256-byte segments).  Rows longer than a chunk get a second,
// tiny kernel: each chunk looks back at its predecessors' (generate,
// all-propagate) flags and, only if its carry-in is 1, ripples it through
// its own limbs.  Rows of at most kThreadRowMax limbs use one thread per row.
#include "common.hpp"

namespace gpupi {
namespace {

constexpr uint32_t kThreadRowMax = 16;

struct Digit {
    uint32_t d;
    uint32_t g;
};

struct SpillArgs {
    const uint64_t *spill;
    uint32_t run_log2;  // runs of 2^run_log2 limbs
    uint32_t runs_per_row;
};

template <AddSubKind K>
struct OpTraits {
    static constexpr bool kSub = (K == AddSubKind::Sub || K == AddSubKind::SubSmall ||
                                  K == AddSubKind::ShiftSub);
};

__device__ __forceinline__ uint32_t load_or_zero(const uint32_t *p, uint32_t row, uint32_t stride,
                                                 int64_t j, uint32_t len) {
    return (j >= 0 && j < (int64_t)len) ? p[(size_t)row * stride + (size_t)j] : 0u;
}

template <AddSubKind K>
__device__ __forceinline__ Digit digit_at(const AddSubArgs &a, const SpillArgs &sp, uint32_t row,
                                          uint32_t j) {
    uint32_t x, y;
    if constexpr (K == AddSubKind::Add || K == AddSubKind::Sub) {
        x = load_or_zero(a.A, row, a.stride_A, j, a.L_a);
        y = load_or_zero(a.B, row, a.stride_B, j, a.L_b);
    } else if constexpr (K == AddSubKind::AddSmall || K == AddSubKind::SubSmall) {
        x = load_or_zero(a.A, row, a.stride_A, j, a.L_a);
        y = (j == 0) ? a.scalar : 0u;
    } else if constexpr (K == AddSubKind::MulSmall) {
        const uint64_t p = (uint64_t)load_or_zero(a.A, row, a.stride_A, j, a.L_a) * a.scalar;
        const uint64_t q = (uint64_t)load_or_zero(a.A, row, a.stride_A, (int64_t)j - 1, a.L_a) * a.scalar;
        const uint32_t lo = (uint32_t)p;
        const uint32_t d = lo + (uint32_t)(q >> 32);
        return {d, (uint32_t)(d < lo)};
    } else if constexpr (K == AddSubKind::ShiftAdd || K == AddSubKind::ShiftSub) {
        const int64_t i = (int64_t)j - (int64_t)(a.shift >> 5);
        const uint32_t r = a.shift & 31u;
        const uint32_t hi = load_or_zero(a.A, row, a.stride_A, i, a.L_a);
        x = hi;
        if (r != 0) {
            const uint32_t lo = load_or_zero(a.A, row, a.stride_A, i - 1, a.L_a);
            x = (hi << r) | (lo >> (32u - r));
        }
        y = load_or_zero(a.B, row, a.stride_B, j, a.L_b);
    } else {
        // spill-add: C[j] += word of spill[run] at (run + 1) 2^run_log2 (+1)
        x = load_or_zero(a.A, row, a.stride_A, j, a.L_a);
        y = 0;
        const uint32_t q = j >> sp.run_log2, rem = j & ((1u << sp.run_log2) - 1u);
        if (q >= 1 && rem < 2) {
            const uint64_t s = sp.spill[(size_t)row * sp.runs_per_row + (q - 1)];
            y = rem == 0 ? (uint32_t)s : (uint32_t)(s >> 32);
        }
    }
    if constexpr (OpTraits<K>::kSub) {
        return {x - y, (uint32_t)(x < y)};
    } else {
        const uint32_t d = x + y;
        return {d, (uint32_t)(d < x)};
    }
}

template <AddSubKind K>
__global__ void __launch_bounds__(256) addsub_thread_kernel(AddSubArgs a, SpillArgs sp) {
    const uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= a.N) return;
    constexpr bool kSub = OpTraits<K>::kSub;
    uint32_t c = 0;
    uint32_t *out = a.C + (size_t)row * a.stride_C;
    for (uint32_t j = 0; j < a.L_c; ++j) {
        const Digit t = digit_at<K>(a, sp, row, j);
        const uint32_t r = kSub ? t.d - c : t.d + c;
        c = t.g | (c & (uint32_t)(kSub ? t.d == 0u : t.d == 0xffffffffu));
        out[j] = r;
    }
}

// One wave per chunk of SEG*64 limbs.  flags[chunk] = carry-out (bit 0) |
// all-propagate (bit 1), both for a carry-in of zero.
template <AddSubKind K, int SEG>
__global__ void __launch_bounds__(256) addsub_chunk_kernel(AddSubArgs a, SpillArgs sp,
                                                           uint32_t chunks_per_row,
                                                           uint64_t total_chunks, uint8_t *flags) {
    constexpr bool kSub = OpTraits<K>::kSub;
    const uint64_t wave = (uint64_t)blockIdx.x * (blockDim.x / kWave) + (threadIdx.x / kWave);
    if (wave >= total_chunks) return;
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t row = (uint32_t)(wave / chunks_per_row);
    const uint32_t chunk = (uint32_t)(wave - (uint64_t)row * chunks_per_row);
    const uint32_t base = chunk * (uint32_t)(SEG * kWave);

    uint32_t d[SEG];
    uint64_t G[SEG], P[SEG];
#pragma unroll
    for (int s = 0; s < SEG; ++s) {
        const uint32_t j = base + s * kWave + lane;
        const bool valid = j < a.L_c;
        Digit t = {0u, 0u};
        if (valid) t = digit_at<K>(a, sp, row, j);
        d[s] = t.d;
        G[s] = ballot64(valid && t.g);
        P[s] = ballot64(valid && (kSub ? t.d == 0u : t.d == 0xffffffffu));
    }
    uint32_t cin = 0;
    bool all_prop = true;
    uint32_t *out = a.C + (size_t)row * a.stride_C;
#pragma unroll
    for (int s = 0; s < SEG; ++s) {
        uint32_t cout;
        const uint64_t cm = resolve_carries(G[s], P[s], cin, &cout);
        const uint32_t bit = (uint32_t)(cm >> lane) & 1u;
        const uint32_t j = base + s * kWave + lane;
        if (j < a.L_c) out[j] = kSub ? d[s] - bit : d[s] + bit;
        all_prop = all_prop && (P[s] == ~0ull);
        cin = cout;
    }
    if (flags != nullptr && lane == 0) flags[wave] = (uint8_t)(cin | (all_prop ? 2u : 0u));
}

// One wave per chunk >= 1 of each row: carry-in by look-back over the
// flags, then an in-chunk ripple if the carry-in is 1.
template <bool kSub>
__global__ void __launch_bounds__(256) carry_fixup_kernel(uint32_t *C, uint32_t stride_C, uint32_t L_c,
                                                          uint32_t chunk_len, uint32_t chunks_per_row,
                                                          uint64_t total, const uint8_t *flags) {
    const uint64_t wave = (uint64_t)blockIdx.x * (blockDim.x / kWave) + (threadIdx.x / kWave);
    if (wave >= total) return;
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t per_row = chunks_per_row - 1;
    const uint32_t row = (uint32_t)(wave / per_row);
    const uint32_t chunk = 1u + (uint32_t)(wave - (uint64_t)row * per_row);
    const uint8_t *rf = flags + (size_t)row * chunks_per_row;

    uint32_t cin = 0;
    for (int64_t hi = (int64_t)chunk - 1; hi >= 0; hi -= kWave) {
        const int64_t k = hi - (int64_t)lane;
        const uint32_t f = (k >= 0) ? rf[k] : 0u;  // k < 0: start of row, carry-in 0
        const uint64_t stop = ballot64((f & 2u) == 0u);
        if (stop != 0) {
            const int first = __builtin_ctzll(stop);
            cin = __shfl(f, first) & 1u;
            break;
        }
    }
    if (cin == 0) return;
    uint32_t *out = C + (size_t)row * stride_C;
    const uint32_t start = chunk * chunk_len;
    const uint32_t end = min(start + chunk_len, L_c);
    for (uint32_t seg = start; seg < end; seg += kWave) {
        const uint32_t j = seg + lane;
        const bool valid = j < end;
        const uint32_t v = valid ? out[j] : 0u;
        const bool prop = valid && (kSub ? v == 0u : v == 0xffffffffu);
        const uint64_t m = ballot64(!prop);
        const uint32_t first = m ? (uint32_t)__builtin_ctzll(m) : 64u;
        if (valid && lane < first) out[j] = kSub ? 0xffffffffu : 0u;
        if (valid && lane == first) out[j] = kSub ? v - 1u : v + 1u;
        if (m != 0) break;
    }
}

// The spill-add of NTT products with four limbs per lane (16-byte loads and
// stores): one wave per chunk of 1024 limbs of one row, the same chunk and
// flag layout as addsub_chunk_kernel<SpillAdd, 16>.  Each lane settles its
// four limbs for a carry-in of zero, the lanes' (generate, propagate) pairs
// go through one ballot resolution per 256 limbs, and the lane adds its
// carry-in.  Every row must start 16-byte aligned.
__global__ void __launch_bounds__(256) spill_add_vec_kernel(uint32_t *C, uint32_t stride_C, uint32_t L_c,
                                                            SpillArgs sp, uint32_t chunks_per_row,
                                                            uint64_t total_chunks, uint8_t *flags) {
    constexpr int kSeg = 4;  // 4 x 256 limbs per wave
    const uint64_t wave = (uint64_t)blockIdx.x * (blockDim.x / kWave) + (threadIdx.x / kWave);
    if (wave >= total_chunks) return;
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t row = (uint32_t)(wave / chunks_per_row);
    const uint32_t chunk = (uint32_t)(wave - (uint64_t)row * chunks_per_row);
    uint32_t *out = C + (size_t)row * stride_C;
    const uint64_t *rsp = sp.spill + (size_t)row * sp.runs_per_row;
    const uint32_t run_mask = (1u << sp.run_log2) - 1u;
    uint32_t x[kSeg][4];
#pragma unroll
    for (int s = 0; s < kSeg; ++s) {
        const uint32_t j0 = chunk * (uint32_t)(kSeg * 4 * kWave) + (uint32_t)s * (4 * kWave) + 4 * lane;
        if (j0 + 4 <= L_c) {
            const uint4 v = *reinterpret_cast<const uint4 *>(out + j0);
            x[s][0] = v.x;
            x[s][1] = v.y;
            x[s][2] = v.z;
            x[s][3] = v.w;
        } else {
#pragma unroll
            for (int t = 0; t < 4; ++t) x[s][t] = j0 + t < L_c ? out[j0 + t] : 0u;
        }
    }
    uint32_t cin = 0;
    bool all_prop = true;
#pragma unroll
    for (int s = 0; s < kSeg; ++s) {
        const uint32_t j0 = chunk * (uint32_t)(kSeg * 4 * kWave) + (uint32_t)s * (4 * kWave) + 4 * lane;
        uint32_t r[4];
        uint32_t c = 0;
        bool ones = true;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const uint32_t j = j0 + t;
            uint32_t y = 0;
            const uint32_t q = j >> sp.run_log2, rem = j & run_mask;
            if (q >= 1 && rem < 2 && j < L_c) {
                const uint64_t w = rsp[q - 1];
                y = rem == 0 ? (uint32_t)w : (uint32_t)(w >> 32);
            }
            const uint32_t d = x[s][t] + y;
            const uint32_t g = d < y ? 1u : 0u;
            r[t] = d + c;
            c = g | (c & (d == 0xffffffffu ? 1u : 0u));
            ones = ones && r[t] == 0xffffffffu && j < L_c;
        }
        const bool valid = j0 < L_c;
        const uint64_t G = ballot64(valid && c != 0), P = ballot64(valid && ones);
        uint32_t cout;
        const uint64_t cm = resolve_carries(G, P, cin, &cout);
        uint32_t ci = (uint32_t)(cm >> lane) & 1u;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const uint32_t v = r[t] + ci;
            ci = ci & (r[t] == 0xffffffffu ? 1u : 0u);
            r[t] = v;
        }
        if (j0 + 4 <= L_c) {
            *reinterpret_cast<uint4 *>(out + j0) = make_uint4(r[0], r[1], r[2], r[3]);
        } else {
#pragma unroll
            for (int t = 0; t < 4; ++t)
                if (j0 + t < L_c) out[j0 + t] = r[t];
        }
        all_prop = all_prop && (P == ~0ull);
        cin = cout;
    }
    if (flags != nullptr && lane == 0) flags[wave] = (uint8_t)(cin | (all_prop ? 2u : 0u));
}

// The assembly holding launch chunk g (assemblies in chunk order).
__device__ __forceinline__ RaggedAssembly assembly_of(const RaggedAssembly *as, uint32_t n, uint32_t g) {
    uint32_t lo = 0, hi = n;  // the last assembly with chunk0 <= g
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (as[mid].chunk0 <= g) lo = mid;
        else hi = mid;
    }
    return as[lo];
}

// Ragged assembly (RaggedAssembly): one wave per 1024-limb chunk, the
// layout of addsub_vec_kernel (flags per chunk for the fixup).  Limb x of C
// is piece i's staged word x - i s plus, when x - i s < S, piece i - 1's
// word x - i s + s, i = min(x / s, m - 1).
__global__ void __launch_bounds__(256) ragged_assemble_kernel(const uint32_t *staging, uint64_t K,
                                                              const RaggedAssembly *as, uint32_t n,
                                                              uint32_t chunks, uint8_t *flags) {
    constexpr int kSeg = 4;  // 4 x 256 limbs per wave
    const uint32_t g = blockIdx.x * (blockDim.x / kWave) + (threadIdx.x / kWave);
    if (g >= chunks) return;
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const RaggedAssembly a = assembly_of(as, n, g);
    const uint32_t chunk = g - a.chunk0;
    uint32_t cin = 0;
    bool all_prop = true;
#pragma unroll
    for (int sg = 0; sg < kSeg; ++sg) {
        const uint32_t j0 = chunk * kAssembleChunk + (uint32_t)sg * (4 * kWave) + 4 * lane;
        uint32_t d[4], gen[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const uint32_t x = j0 + t;
            d[t] = 0;
            gen[t] = 0;
            if (x < a.c) {
                uint32_t i, off;
                bool below;
                assembly_source(a, x, &i, &off, &below);
                const uint32_t u = staging[(size_t)(a.row0 + i) * K + off];
                const uint32_t v = below ? staging[(size_t)(a.row0 + i - 1) * K + off + a.s] : 0u;
                d[t] = u + v;
                gen[t] = d[t] < u ? 1u : 0u;
            }
        }
        // the lane's four limbs for a carry-in of zero
        uint32_t r[4], c = 0;
        bool prop = true;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            r[t] = d[t] + c;
            c = gen[t] | (c & (d[t] == 0xffffffffu ? 1u : 0u));
            prop = prop && r[t] == 0xffffffffu && j0 + t < a.c;
        }
        const bool valid = j0 < a.c;
        const uint64_t G = ballot64(valid && c != 0), P = ballot64(valid && prop);
        uint32_t cout;
        const uint64_t cm = resolve_carries(G, P, cin, &cout);
        uint32_t ci = (uint32_t)(cm >> lane) & 1u;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const uint32_t v = r[t] + ci;
            ci = ci & (r[t] == 0xffffffffu ? 1u : 0u);
            if (j0 + t < a.c) a.C[j0 + t] = v;
        }
        all_prop = all_prop && (P == ~0ull);
        cin = cout;
    }
    if (lane == 0) flags[g] = (uint8_t)(cin | (all_prop ? 2u : 0u));
}

// One wave per chunk >= 1 of each assembly: carry-in by look-back over the
// assembly's flags, then an in-chunk ripple if it is 1 (carry_fixup_kernel
// over ragged rows).
__global__ void __launch_bounds__(256) ragged_assembly_fixup_kernel(const RaggedAssembly *as, uint32_t n,
                                                                    uint32_t chunks, const uint8_t *flags) {
    const uint32_t g = blockIdx.x * (blockDim.x / kWave) + (threadIdx.x / kWave);
    if (g >= chunks) return;
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const RaggedAssembly a = assembly_of(as, n, g);
    const uint32_t chunk = g - a.chunk0;
    if (chunk == 0) return;
    const uint8_t *rf = flags + a.chunk0;
    uint32_t cin = 0;
    for (int64_t hi = (int64_t)chunk - 1; hi >= 0; hi -= kWave) {
        const int64_t k = hi - (int64_t)lane;
        const uint32_t f = (k >= 0) ? rf[k] : 0u;  // k < 0: start of the assembly, carry-in 0
        const uint64_t stop = ballot64((f & 2u) == 0u);
        if (stop != 0) {
            cin = __shfl(f, __builtin_ctzll(stop)) & 1u;
            break;
        }
    }
    if (cin == 0) return;
    const uint32_t start = chunk * kAssembleChunk, end = min(start + kAssembleChunk, a.c);
    for (uint32_t seg = start; seg < end; seg += kWave) {
        const uint32_t j = seg + lane;
        const bool valid = j < end;
        const uint32_t v = valid ? a.C[j] : 0u;
        const uint64_t m = ballot64(!(valid && v == 0xffffffffu));
        const uint32_t first = m ? (uint32_t)__builtin_ctzll(m) : 64u;
        if (valid && lane < first) a.C[j] = 0u;
        if (valid && lane == first) a.C[j] = v + 1u;
        if (m != 0) break;
    }
}

// Row-aligned 16-byte access: the base is 16-byte aligned and, with more
// than one row, so is every row.
bool rows_aligned(const void *p, uint32_t N, uint32_t stride) {
    return p != nullptr && (((uintptr_t)p) & 15u) == 0 && (N <= 1 || (stride & 3u) == 0);
}

// Four limbs x[0..4) of row `row` from limb j0 on (zero outside [0, L)).
__device__ __forceinline__ void load4(const uint32_t *p, uint32_t row, uint32_t stride, int64_t j0, uint32_t L,
                                      bool vec, uint32_t *x) {
    const uint32_t *r = p + (size_t)row * stride;
    if (vec && j0 >= 0 && j0 + 4 <= (int64_t)L) {
        const uint4 v = *reinterpret_cast<const uint4 *>(r + j0);
        x[0] = v.x;
        x[1] = v.y;
        x[2] = v.z;
        x[3] = v.w;
    } else {
#pragma unroll
        for (int t = 0; t < 4; ++t) x[t] = (j0 + t >= 0 && j0 + t < (int64_t)L) ? r[j0 + t] : 0u;
    }
}

// Add, Sub, ShiftAdd and ShiftSub with four limbs per lane: one wave per
// chunk of 1024 limbs of one row, the chunk and flag layout of
// addsub_chunk_kernel<K, 16> (so carry_fixup_kernel completes long rows).
// C's rows (and B's and, for Add/Sub, A's when a_vec/b_vec) start 16-byte
// aligned; the shifted A of ShiftAdd/ShiftSub is read a limb at a time.
template <AddSubKind K>
__global__ void __launch_bounds__(256) addsub_vec_kernel(AddSubArgs a, bool a_vec, bool b_vec, uint32_t chunks_per_row,
                                                         uint64_t total_chunks, uint8_t *flags) {
    constexpr bool kSub = OpTraits<K>::kSub;
    constexpr bool kShift = K == AddSubKind::ShiftAdd || K == AddSubKind::ShiftSub;
    constexpr int kSeg = 4;  // 4 x 256 limbs per wave
    const uint64_t wave = (uint64_t)blockIdx.x * (blockDim.x / kWave) + (threadIdx.x / kWave);
    if (wave >= total_chunks) return;
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t row = (uint32_t)(wave / chunks_per_row);
    const uint32_t chunk = (uint32_t)(wave - (uint64_t)row * chunks_per_row);
    uint32_t *out = a.C + (size_t)row * a.stride_C;
    const int64_t ws = kShift ? (int64_t)(a.shift >> 5) : 0;
    const uint32_t bs = kShift ? (a.shift & 31u) : 0u;
    uint32_t d[kSeg][4], g[kSeg][4];
#pragma unroll
    for (int s = 0; s < kSeg; ++s) {
        const int64_t j0 = (int64_t)chunk * (kSeg * 4 * kWave) + s * (4 * kWave) + 4 * (int64_t)lane;
        uint32_t x[4], y[4];
        if constexpr (kShift) {
            // x_t = (A << shift) at limb j0 + t: A[i] << bs | A[i - 1] >> (32 - bs), i = j0 + t - ws
            uint32_t w[5];
            const uint32_t *ra = a.A + (size_t)row * a.stride_A;
#pragma unroll
            for (int t = 0; t < 5; ++t) {
                const int64_t i = j0 - ws - 1 + t;
                w[t] = (i >= 0 && i < (int64_t)a.L_a) ? ra[i] : 0u;
            }
#pragma unroll
            for (int t = 0; t < 4; ++t) x[t] = bs ? (w[t + 1] << bs) | (w[t] >> (32u - bs)) : w[t + 1];
        } else {
            load4(a.A, row, a.stride_A, j0, a.L_a, a_vec, x);
        }
        load4(a.B, row, a.stride_B, j0, a.L_b, b_vec, y);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            d[s][t] = kSub ? x[t] - y[t] : x[t] + y[t];
            g[s][t] = kSub ? (x[t] < y[t] ? 1u : 0u) : (d[s][t] < x[t] ? 1u : 0u);
        }
    }
    uint32_t cin = 0;
    bool all_prop = true;
#pragma unroll
    for (int s = 0; s < kSeg; ++s) {
        const int64_t j0 = (int64_t)chunk * (kSeg * 4 * kWave) + s * (4 * kWave) + 4 * (int64_t)lane;
        // the lane's four limbs for a carry-in of zero
        uint32_t r[4], c = 0;
        bool prop = true;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            r[t] = kSub ? d[s][t] - c : d[s][t] + c;
            c = g[s][t] | (c & (kSub ? (d[s][t] == 0u ? 1u : 0u) : (d[s][t] == 0xffffffffu ? 1u : 0u)));
            prop = prop && (kSub ? r[t] == 0u : r[t] == 0xffffffffu) && j0 + t < (int64_t)a.L_c;
        }
        const bool valid = j0 < (int64_t)a.L_c;
        const uint64_t G = ballot64(valid && c != 0), P = ballot64(valid && prop);
        uint32_t cout;
        const uint64_t cm = resolve_carries(G, P, cin, &cout);
        uint32_t ci = (uint32_t)(cm >> lane) & 1u;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const uint32_t v = kSub ? r[t] - ci : r[t] + ci;
            ci = ci & (kSub ? (r[t] == 0u ? 1u : 0u) : (r[t] == 0xffffffffu ? 1u : 0u));
            r[t] = v;
        }
        if (j0 + 4 <= (int64_t)a.L_c) {
            *reinterpret_cast<uint4 *>(out + j0) = make_uint4(r[0], r[1], r[2], r[3]);
        } else {
#pragma unroll
            for (int t = 0; t < 4; ++t)
                if (j0 + t < (int64_t)a.L_c) out[j0 + t] = r[t];
        }
        all_prop = all_prop && (P == ~0ull);
        cin = cout;
    }
    if (flags != nullptr && lane == 0) flags[wave] = (uint8_t)(cin | (all_prop ? 2u : 0u));
}

struct ChunkPlan {
    bool thread_rows;
    int seg;
    uint32_t chunk_len;
    uint32_t chunks_per_row;
};

ChunkPlan plan_for(uint32_t L_c) {
    ChunkPlan p{};
    if (L_c <= kThreadRowMax) {
        p.thread_rows = true;
        p.seg = 0;
        p.chunk_len = L_c;
        p.chunks_per_row = 1;
        return p;
    }
    int seg = 1;
    while (seg < 16 && (uint64_t)seg * kWave < L_c) seg *= 2;
    p.thread_rows = false;
    p.seg = seg;
    p.chunk_len = (uint32_t)seg * kWave;
    p.chunks_per_row = ceil_div(L_c, p.chunk_len);
    return p;
}

template <AddSubKind K>
hipError_t run_addsub(const AddSubArgs &a, const SpillArgs &sp, uint8_t *flags, hipStream_t stream) {
    if (a.N == 0 || a.L_c == 0) return hipSuccess;
    const ChunkPlan p = plan_for(a.L_c);
    if (p.thread_rows) {
        const uint32_t blocks = ceil_div(a.N, 256);
        hipLaunchKernelGGL(addsub_thread_kernel<K>, dim3(blocks), dim3(256), 0, stream, a, sp);
        return hipGetLastError();
    }
    const uint64_t total = (uint64_t)a.N * p.chunks_per_row;
    const bool need_fix = p.chunks_per_row > 1;
    if (need_fix && flags == nullptr) return hipErrorInvalidValue;
    const uint64_t blocks = ceil_div64(total, 4);
    if (blocks > 0x7fffffffull) return hipErrorInvalidValue;
    uint8_t *fl = need_fix ? flags : nullptr;
    constexpr bool kVecKind = K == AddSubKind::Add || K == AddSubKind::Sub || K == AddSubKind::ShiftAdd ||
                              K == AddSubKind::ShiftSub;
    if (kVecKind && p.chunk_len == 4u * 4u * kWave && rows_aligned(a.C, a.N, a.stride_C)) {
        // four limbs per lane (in place is fine: a lane reads the limbs it
        // writes before writing them); 16-byte B (and A) loads where their
        // rows allow
        const bool a_vec = rows_aligned(a.A, a.N, a.stride_A), b_vec = rows_aligned(a.B, a.N, a.stride_B);
        hipLaunchKernelGGL(addsub_vec_kernel<K>, dim3((uint32_t)blocks), dim3(256), 0, stream, a, a_vec, b_vec,
                           p.chunks_per_row, total, fl);
    } else {
    switch (p.seg) {
        case 1: hipLaunchKernelGGL((addsub_chunk_kernel<K, 1>), dim3((uint32_t)blocks), dim3(256), 0, stream, a, sp, p.chunks_per_row, total, fl); break;
        case 2: hipLaunchKernelGGL((addsub_chunk_kernel<K, 2>), dim3((uint32_t)blocks), dim3(256), 0, stream, a, sp, p.chunks_per_row, total, fl); break;
        case 4: hipLaunchKernelGGL((addsub_chunk_kernel<K, 4>), dim3((uint32_t)blocks), dim3(256), 0, stream, a, sp, p.chunks_per_row, total, fl); break;
        case 8: hipLaunchKernelGGL((addsub_chunk_kernel<K, 8>), dim3((uint32_t)blocks), dim3(256), 0, stream, a, sp, p.chunks_per_row, total, fl); break;
        default: hipLaunchKernelGGL((addsub_chunk_kernel<K, 16>), dim3((uint32_t)blocks), dim3(256), 0, stream, a, sp, p.chunks_per_row, total, fl); break;
    }
    }
    GPUPI_RETURN_IF(hipGetLastError());
    if (need_fix) {
        const uint64_t fix_total = (uint64_t)a.N * (p.chunks_per_row - 1);
        const uint64_t fblocks = ceil_div64(fix_total, 4);
        if (fblocks > 0x7fffffffull) return hipErrorInvalidValue;
        hipLaunchKernelGGL(carry_fixup_kernel<OpTraits<K>::kSub>, dim3((uint32_t)fblocks), dim3(256), 0,
                           stream, a.C, a.stride_C, a.L_c, p.chunk_len, p.chunks_per_row, fix_total,
                           (const uint8_t *)flags);
        GPUPI_RETURN_IF(hipGetLastError());
    }
    return hipSuccess;
}

}  // namespace

size_t carry_flags_bytes(uint32_t N, uint32_t L_c) {
    const ChunkPlan p = plan_for(L_c);
    if (p.chunks_per_row <= 1) return 0;
    return (size_t)N * p.chunks_per_row;
}

hipError_t launch_addsub(const AddSubArgs &a, uint8_t *flags, hipStream_t stream) {
    const SpillArgs sp{nullptr, 0, 0};
    switch (a.kind) {
        case AddSubKind::Add: return run_addsub<AddSubKind::Add>(a, sp, flags, stream);
        case AddSubKind::Sub: return run_addsub<AddSubKind::Sub>(a, sp, flags, stream);
        case AddSubKind::AddSmall: return run_addsub<AddSubKind::AddSmall>(a, sp, flags, stream);
        case AddSubKind::SubSmall: return run_addsub<AddSubKind::SubSmall>(a, sp, flags, stream);
        case AddSubKind::MulSmall: return run_addsub<AddSubKind::MulSmall>(a, sp, flags, stream);
        case AddSubKind::ShiftAdd: return run_addsub<AddSubKind::ShiftAdd>(a, sp, flags, stream);
        case AddSubKind::ShiftSub: return run_addsub<AddSubKind::ShiftSub>(a, sp, flags, stream);
        case AddSubKind::SpillAdd: break;  // needs spill words: launch_add_spills
    }
    return hipErrorInvalidValue;
}

hipError_t launch_add_spills(uint32_t *C, const uint64_t *spill, uint32_t N, uint32_t L_c,
                             uint32_t stride_C, uint32_t run_len, uint32_t runs_per_row,
                             uint8_t *flags, hipStream_t stream) {
    AddSubArgs a{};
    a.kind = AddSubKind::SpillAdd;
    a.A = C;
    a.C = C;
    a.N = N;
    a.L_a = L_c;
    a.L_c = L_c;
    a.stride_A = stride_C;
    a.stride_C = stride_C;
    if (run_len == 0 || (run_len & (run_len - 1)) != 0) return hipErrorInvalidValue;
    const SpillArgs sp{spill, (uint32_t)__builtin_ctz(run_len), runs_per_row};
    const ChunkPlan p = plan_for(L_c);
    if (N == 0 || L_c == 0) return hipSuccess;
    if (p.thread_rows || p.chunk_len != 4u * 4u * kWave || (((uintptr_t)C) & 15u) != 0 ||
        (N > 1 && (stride_C & 3u) != 0))
        return run_addsub<AddSubKind::SpillAdd>(a, sp, flags, stream);
    const uint64_t total = (uint64_t)N * p.chunks_per_row;
    const bool need_fix = p.chunks_per_row > 1;
    if (need_fix && flags == nullptr) return hipErrorInvalidValue;
    const uint64_t blocks = ceil_div64(total, 4);
    if (blocks > 0x7fffffffull) return hipErrorInvalidValue;
    hipLaunchKernelGGL(spill_add_vec_kernel, dim3((uint32_t)blocks), dim3(256), 0, stream, C, stride_C, L_c, sp,
                       p.chunks_per_row, total, need_fix ? flags : nullptr);
    GPUPI_RETURN_IF(hipGetLastError());
    if (need_fix) {
        const uint64_t fix_total = (uint64_t)N * (p.chunks_per_row - 1);
        const uint64_t fblocks = ceil_div64(fix_total, 4);
        if (fblocks > 0x7fffffffull) return hipErrorInvalidValue;
        hipLaunchKernelGGL(carry_fixup_kernel<false>, dim3((uint32_t)fblocks), dim3(256), 0, stream, C, stride_C, L_c,
                           p.chunk_len, p.chunks_per_row, fix_total, (const uint8_t *)flags);
        GPUPI_RETURN_IF(hipGetLastError());
    }
    return hipSuccess;
}

hipError_t launch_ragged_assemble(const uint32_t *staging, uint64_t K, const RaggedAssembly *as, uint32_t n,
                                  uint32_t chunks, uint8_t *flags, hipStream_t stream) {
    if (n == 0 || chunks == 0) return hipSuccess;
    if (staging == nullptr || as == nullptr || flags == nullptr) return hipErrorInvalidValue;
    const uint32_t blocks = (chunks + 3) / 4;
    hipLaunchKernelGGL(ragged_assemble_kernel, dim3(blocks), dim3(256), 0, stream, staging, K, as, n, chunks, flags);
    GPUPI_RETURN_IF(hipGetLastError());
    hipLaunchKernelGGL(ragged_assembly_fixup_kernel, dim3(blocks), dim3(256), 0, stream, as, n, chunks,
                       (const uint8_t *)flags);
    return hipGetLastError();
}

}  // namespace gpupi

// ---------------------------------------------------------------------------
// C entry points (reference per-op headers); null stream, caller workspace.
// ---------------------------------------------------------------------------
using gpupi::AddSubArgs;
using gpupi::AddSubKind;

static void addsub_entry(AddSubKind kind, const uint32_t *A, const uint32_t *B, uint32_t scalar,
                         uint32_t shift, uint32_t *C, uint32_t *ws, uint32_t N, uint32_t L_a,
                         uint32_t L_b, uint32_t L_c, uint32_t sA, uint32_t sB, uint32_t sC) {
    AddSubArgs a{kind, A, B, C, scalar, shift, N, L_a, L_b, L_c, sA, sB, sC};
    (void)gpupi::launch_addsub(a, (uint8_t *)ws, 0);
}

extern "C" {

void batch_add_simple(uint32_t *A, uint32_t *B, uint32_t *C, uint32_t *workspace, uint32_t N,
                      uint32_t L_a, uint32_t L_b, uint32_t L_c, uint32_t stride_A,
                      uint32_t stride_B, uint32_t stride_C) {
    addsub_entry(AddSubKind::Add, A, B, 0, 0, C, workspace, N, L_a, L_b, L_c, stride_A, stride_B, stride_C);
}
size_t batch_add_simple_workspace_size(uint32_t N, uint32_t, uint32_t, uint32_t L_c) {
    return gpupi::carry_flags_bytes(N, L_c);
}

void batch_sub_simple(uint32_t *A, uint32_t *B, uint32_t *C, uint32_t *workspace, uint32_t N,
                      uint32_t L_a, uint32_t L_b, uint32_t L_c, uint32_t stride_A,
                      uint32_t stride_B, uint32_t stride_C) {
    addsub_entry(AddSubKind::Sub, A, B, 0, 0, C, workspace, N, L_a, L_b, L_c, stride_A, stride_B, stride_C);
}
size_t batch_sub_simple_workspace_size(uint32_t N, uint32_t, uint32_t, uint32_t L_c) {
    return gpupi::carry_flags_bytes(N, L_c);
}

void batch_add_small(const uint32_t *A, uint32_t B, uint32_t *C, uint32_t *workspace, uint32_t N,
                     uint32_t L_a, uint32_t L_c, uint32_t stride_A, uint32_t stride_C) {
    addsub_entry(AddSubKind::AddSmall, A, nullptr, B, 0, C, workspace, N, L_a, 0, L_c, stride_A, 0, stride_C);
}
size_t batch_add_small_workspace_size(uint32_t N, uint32_t, uint32_t L_c) {
    return gpupi::carry_flags_bytes(N, L_c);
}

void batch_sub_small(const uint32_t *A, uint32_t B, uint32_t *C, uint32_t *workspace, uint32_t N,
                     uint32_t L_a, uint32_t L_c, uint32_t stride_A, uint32_t stride_C) {
    addsub_entry(AddSubKind::SubSmall, A, nullptr, B, 0, C, workspace, N, L_a, 0, L_c, stride_A, 0, stride_C);
}
size_t batch_sub_small_workspace_size(uint32_t N, uint32_t, uint32_t L_c) {
    return gpupi::carry_flags_bytes(N, L_c);
}

void batch_mul_small(const uint32_t *A, uint32_t B, uint32_t *C, uint32_t *workspace, uint32_t N,
                     uint32_t L_a, uint32_t L_c, uint32_t stride_A, uint32_t stride_C) {
    addsub_entry(AddSubKind::MulSmall, A, nullptr, B, 0, C, workspace, N, L_a, 0, L_c, stride_A, 0, stride_C);
}
size_t batch_mul_small_workspace_size(uint32_t N, uint32_t, uint32_t L_c) {
    return gpupi::carry_flags_bytes(N, L_c);
}

void batch_shift_add_simple(const uint32_t *A, const uint32_t *B, uint32_t *C, uint32_t *workspace,
                            uint32_t shift, uint32_t N, uint32_t L_a, uint32_t L_b, uint32_t L_c,
                            uint32_t stride_A, uint32_t stride_B, uint32_t stride_C) {
    addsub_entry(AddSubKind::ShiftAdd, A, B, 0, shift, C, workspace, N, L_a, L_b, L_c, stride_A, stride_B, stride_C);
}
size_t batch_shift_add_simple_workspace_size(uint32_t N, uint32_t, uint32_t, uint32_t L_c, uint32_t) {
    return gpupi::carry_flags_bytes(N, L_c);
}

void batch_shift_sub_simple(const uint32_t *A, const uint32_t *B, uint32_t *C, uint32_t *workspace,
                            uint32_t shift, uint32_t N, uint32_t L_a, uint32_t L_b, uint32_t L_c,
                            uint32_t stride_A, uint32_t stride_B, uint32_t stride_C) {
    addsub_entry(AddSubKind::ShiftSub, A, B, 0, shift, C, workspace, N, L_a, L_b, L_c, stride_A, stride_B, stride_C);
}
size_t batch_shift_sub_simple_workspace_size(uint32_t N, uint32_t, uint32_t, uint32_t L_c, uint32_t) {
    return gpupi::carry_flags_bytes(N, L_c);
}

}  // extern "C"
