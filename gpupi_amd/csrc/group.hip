// The 17-term group stage of the binary-splitting tree (pi_bs_gpu.cu:555-779)
// as one kernel: each wave builds one group's P, Q, R from its 16 body terms
// and its head term entirely in LDS, and writes the three rows once.
//
// The op sequence is the reference's: four pairwise levels (R = R_e Q_o
// 2^(15 t) -+ P_e R_o, subtracting at the first level), the head merge
// (R = R Q_0 2^15 + P R_0) and the exact division by 13^2 17^3 7^6 11^3 (here
// one Hensel pass by the product 830297 * 156590819, which leaves the same
// quotient).  Lengths are the allocation bounds of the batched form, so the
// rows (51, 68 and 83 limbs) are exactly those the per-op form leaves
// before compaction.
//
// Within a level, the four products of every pair are laid out end to end
// (each in exactly L_a + L_b limbs) and computed column-parallel: lane l owns
// column c0 + l, accumulates it as x + 2^32 h, and the columns become limbs
// with two neighbour shifts and one ballot carry resolution per 64 columns.
// Carries never cross from one product into the next because every product
// fits its slot; the shifted add/sub of R works the same way.
#include <atomic>

#include "common.hpp"
#include "gpupi_pi.h"
#include "exact_div.hpp"
#include "leaf_terms.hpp"

namespace gpupi {
namespace {

// Groups per block (one wave each): a launch parameter
// (gpupi_group_stage_set_shape), 4 built in.
constexpr uint32_t kGroupsPerBlockDefault = 4;
std::atomic<uint32_t> g_groups_per_block{kGroupsPerBlockDefault};

// LDS words per group: leaf terms (16 x 13) | head term (13) | two product
// areas | two R areas.
constexpr int kLeaf = 0, kHead = 208, kT0 = 224, kT1 = kT0 + 264, kR0 = kT1 + 264, kR1 = kR0 + 88;
constexpr int kGroupWords = kR1 + 88;

// A set of nodes: node n's P, Q, R at P + n sP, Q + n sQ, R + n sR.
struct Nodes {
    const uint32_t *P, *Q, *R;
    uint32_t sP, sQ, sR;
};

// One level: PAIRS pairs (even node, odd node), products into T, R into Rout.
// ODD_NEXT: the odd node of pair j is node 2j + 1 of `ev` (pairwise levels);
// otherwise node j of `od` (the head merge).
template <int PAIRS, int LPe, int LQe, int LRe, int LPo, int LQo, int LRo, int SHIFT, bool SUB, int RLEN,
          bool ODD_NEXT>
__device__ __forceinline__ void group_level(uint32_t lane, const Nodes &ev, const Nodes &od, uint32_t *T,
                                            uint32_t *Rout) {
    constexpr int LP2 = LPe + LPo, LQ2 = LQe + LQo, L1 = LRe + LQo, L2 = LPe + LRo;
    constexpr int S = LP2 + LQ2 + L1 + L2;
    constexpr int TOT = PAIRS * S;
    // products: columns of [P_e P_o | Q_e Q_o | R_e Q_o | P_e R_o] per pair
    uint32_t prev_y = 0, prev_z62 = 0, prev_z63 = 0, prev_cc = 0, bin = 0;
    for (int c0 = 0; c0 < TOT; c0 += kWave) {
        const int c = c0 + (int)lane;
        uint32_t x = 0;
        uint64_t h = 0;
        if (c < TOT) {
            const int j = c / S, off = c - j * S;
            const int e = ODD_NEXT ? 2 * j : j, o = ODD_NEXT ? 2 * j + 1 : j;
            const Nodes &on = ODD_NEXT ? ev : od;
            const uint32_t *a, *b;
            int la, lb, k;
            if (off < LP2) {
                a = ev.P + e * ev.sP; la = LPe; b = on.P + o * on.sP; lb = LPo; k = off;
            } else if (off < LP2 + LQ2) {
                a = ev.Q + e * ev.sQ; la = LQe; b = on.Q + o * on.sQ; lb = LQo; k = off - LP2;
            } else if (off < LP2 + LQ2 + L1) {
                a = ev.R + e * ev.sR; la = LRe; b = on.Q + o * on.sQ; lb = LQo; k = off - LP2 - LQ2;
            } else {
                a = ev.P + e * ev.sP; la = LPe; b = on.R + o * on.sR; lb = LRo; k = off - LP2 - LQ2 - L1;
            }
            const int ilo = k - lb + 1 > 0 ? k - lb + 1 : 0;
            const int ihi = k < la - 1 ? k : la - 1;
            for (int i = ilo; i <= ihi; ++i) {
                const uint64_t t = (uint64_t)a[i] * b[k - i] + x;
                x = (uint32_t)t;
                h += t >> 32;
            }
        }
        // column c = x + 2^32 y + 2^64 z  ->  limbs (pi: x_c + y_{c-1} + z_{c-2})
        const uint32_t y = (uint32_t)h, z = (uint32_t)(h >> 32);
        uint32_t y1 = __shfl(y, (int)lane - 1), z2 = __shfl(z, (int)lane - 2);
        if (lane == 0) y1 = prev_y;
        if (lane == 0) z2 = prev_z62;
        if (lane == 1) z2 = prev_z63;
        const uint64_t s = (uint64_t)x + y1 + z2;
        const uint32_t d = (uint32_t)s, cc = (uint32_t)(s >> 32);
        uint32_t cc1 = __shfl(cc, (int)lane - 1);
        if (lane == 0) cc1 = prev_cc;
        const uint64_t u = (uint64_t)d + cc1;
        const uint32_t el = (uint32_t)u, gen = (uint32_t)(u >> 32);
        const uint64_t G = ballot64(gen != 0), Pm = ballot64(gen == 0 && el == 0xffffffffu);
        uint32_t cout;
        const uint64_t cm = resolve_carries(G, Pm, bin, &cout);
        if (c < TOT) T[c] = el + ((uint32_t)(cm >> lane) & 1u);
        prev_y = __shfl(y, 63);
        prev_z62 = __shfl(z, 62);
        prev_z63 = __shfl(z, 63);
        prev_cc = __shfl(cc, 63);
        bin = cout;
    }
    __syncthreads();
    // R = (R_e Q_o << SHIFT) -+ P_e R_o, RLEN limbs per pair
    constexpr int TOTR = PAIRS * RLEN;
    constexpr int WS = SHIFT / 32, BS = SHIFT % 32;
    bin = 0;
    for (int c0 = 0; c0 < TOTR; c0 += kWave) {
        const int c = c0 + (int)lane;
        uint32_t el = 0, gen = 0, prop = 0;
        if (c < TOTR) {
            const int j = c / RLEN, k = c - j * RLEN;
            const uint32_t *R1 = T + j * S + LP2 + LQ2, *R2 = R1 + L1;
            const int i1 = k - WS;
            uint32_t v = (i1 >= 0 && i1 < L1) ? (R1[i1] << BS) : 0u;
            if (BS != 0 && i1 - 1 >= 0 && i1 - 1 < L1) v |= R1[i1 - 1] >> ((32 - BS) & 31);
            const uint32_t w = k < L2 ? R2[k] : 0u;
            if (SUB) {
                el = v - w;
                gen = v < w;
                prop = v == w;
            } else {
                const uint64_t t = (uint64_t)v + w;
                el = (uint32_t)t;
                gen = (uint32_t)(t >> 32);
                prop = gen == 0 && el == 0xffffffffu;
            }
        }
        const uint64_t G = ballot64(gen != 0), Pm = ballot64(prop != 0);
        uint32_t cout;
        const uint64_t cm = resolve_carries(G, Pm, bin, &cout);
        const uint32_t ci = (uint32_t)(cm >> lane) & 1u;
        if (c < TOTR) Rout[c] = SUB ? el - ci : el + ci;
        bin = cout;
    }
    __syncthreads();
}

// Exact division of a[0..n) by the block divisor, in place (hensel_step
// from the low limb up).  Lanes divide different rows in one loop of
// NMAX >= n steps (a lane idles past its n), and the next limb is read from
// LDS while the current one is divided.
template <int NMAX>
__device__ __forceinline__ void exact_div_block(uint32_t *a, int n) {
    uint64_t c = 0;
    uint32_t next = a[0];
    for (int i = 0; i < NMAX; ++i) {
        const uint32_t ai = next;
        if (i + 1 < n) next = a[i + 1];
        uint64_t c1 = c;
        const uint32_t q = hensel_step(ai, c1);
        if (i < n) {
            c = c1;
            a[i] = q;
        }
    }
}

template <uint32_t GPB>
__global__ void __launch_bounds__(GPB * kWave) group_stage_kernel(uint32_t N, uint32_t *P, uint32_t *Q, uint32_t *R) {
    __shared__ uint32_t smem[GPB * kGroupWords];
    const uint32_t wave = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
    const uint32_t g = blockIdx.x * GPB + wave;
    const bool live = g < N;
    uint32_t *sm = smem + wave * kGroupWords;
    // leaf terms: body i -> term 17 g + i, head -> term 17 g + 16
    if (lane < 17) {
        const int64_t term = (int64_t)g * 17 + lane;
        uint32_t *dst = lane < 16 ? sm + kLeaf + 13 * lane : sm + kHead;
        if (live) {
            leaf_term(term, dst, dst + 3, dst + 7);
        } else {
            for (int w = 0; w < 13; ++w) dst[w] = 0;
        }
    }
    __syncthreads();
    const Nodes leaf{sm + kLeaf, sm + kLeaf + 3, sm + kLeaf + 7, 13, 13, 13};
    const Nodes head{sm + kHead, sm + kHead + 3, sm + kHead + 7, 0, 0, 0};
    // level 1: 8 pairs of (3, 4, 6)-limb terms, shift 15, subtract
    group_level<8, 3, 4, 6, 3, 4, 6, 15, true, 11, true>(lane, leaf, leaf, sm + kT0, sm + kR0);
    const Nodes n1{sm + kT0, sm + kT0 + 6, sm + kR0, 33, 33, 11};
    group_level<4, 6, 8, 11, 6, 8, 11, 30, false, 21, true>(lane, n1, n1, sm + kT1, sm + kR1);
    const Nodes n2{sm + kT1, sm + kT1 + 12, sm + kR1, 64, 64, 21};
    group_level<2, 12, 16, 21, 12, 16, 21, 60, false, 40, true>(lane, n2, n2, sm + kT0, sm + kR0);
    const Nodes n3{sm + kT0, sm + kT0 + 24, sm + kR0, 126, 126, 40};
    group_level<1, 24, 32, 40, 24, 32, 40, 120, false, 77, true>(lane, n3, n3, sm + kT1, sm + kR1);
    const Nodes n4{sm + kT1, sm + kT1 + 48, sm + kR1, 0, 0, 0};
    // head merge: P P0 | Q Q0 | R Q0 | P R0, R = (R Q0 << 15) + P R0
    group_level<1, 48, 64, 77, 3, 4, 6, 15, false, 83, false>(lane, n4, head, sm + kT0, sm + kR0);
    // exact division of P (51), Q (68), R (83): one lane each, side by side
    if (lane < 3) {
        const int off = lane == 0 ? kT0 : lane == 1 ? kT0 + 51 : kR0;
        const int n = lane == 0 ? 51 : lane == 1 ? 68 : 83;
        exact_div_block<83>(sm + off, n);
    }
    __syncthreads();
    if (live) {
        uint32_t *po = P + (size_t)g * 51, *qo = Q + (size_t)g * 68, *ro = R + (size_t)g * 83;
        for (uint32_t i = lane; i < 51; i += kWave) po[i] = sm[kT0 + i];
        for (uint32_t i = lane; i < 68; i += kWave) qo[i] = sm[kT0 + 51 + i];
        for (uint32_t i = lane; i < 83; i += kWave) ro[i] = sm[kR0 + i];
    }
}

}  // namespace

hipError_t launch_group_stage(uint32_t N, uint32_t *P, uint32_t *Q, uint32_t *R, hipStream_t stream) {
    if (N == 0) return hipSuccess;
    const uint32_t gpb = g_groups_per_block.load(std::memory_order_relaxed);
    const auto launch = [&](auto kernel, uint32_t per_block) {
        hipLaunchKernelGGL(kernel, dim3(ceil_div(N, per_block)), dim3(per_block * kWave), 0, stream, N, P, Q, R);
    };
    switch (gpb) {
        case 1: launch(group_stage_kernel<1>, 1); break;
        case 2: launch(group_stage_kernel<2>, 2); break;
        case 8: launch(group_stage_kernel<8>, 8); break;
        case 16: launch(group_stage_kernel<16>, 16); break;
        default: launch(group_stage_kernel<4>, 4); break;
    }
    return hipGetLastError();
}

}  // namespace gpupi

extern "C" {

hipError_t gpupi_group_stage_set_shape(uint32_t groups_per_block) {
    if (groups_per_block != 1 && groups_per_block != 2 && groups_per_block != 4 && groups_per_block != 8 &&
        groups_per_block != 16)
        return hipErrorInvalidValue;
    gpupi::g_groups_per_block.store(groups_per_block);
    return hipSuccess;
}

uint32_t gpupi_group_stage_get_shape(void) { return gpupi::g_groups_per_block.load(); }

}  // extern "C"
