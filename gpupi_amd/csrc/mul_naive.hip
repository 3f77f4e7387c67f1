// Schoolbook batched products for short operands (reference
// batch_mul_naive.cu:8-273 covers the same role with warp-32 kernels).
//
// Two wave64 designs:
//  * thread per product (short operand <= 16 limbs): the short operand and a
//    sliding window of partial sums live in registers; the long operand
//    streams through, one 32x32+32+32 -> 64-bit multiply-add per limb pair.
//  * wave per product: operands staged in LDS; lane l owns output columns
//    256g + 4l .. +3 and accumulates each column as lo32 + hi64 (no carry
//    handling in the inner loop: t = a*b + lo fits 64 bits).  Columns are
//    turned into limbs with two lane shifts (y_{k-1}, z_{k-2}) and one
//    ballot carry resolution per 256-limb group.
#include <mutex>

#include "common.hpp"

namespace gpupi {
namespace {

constexpr uint32_t kThreadMaxShort = 16;
constexpr uint32_t kWaveMaxTotal = BATCH_MUL_NAIVE_L_MAX;

template <int LB>
__global__ void __launch_bounds__(256) mul_thread_kernel(const uint32_t *A, const uint32_t *B, uint32_t *C,
                                                         uint32_t N, uint32_t LA, uint32_t sA,
                                                         uint32_t sB, uint32_t sC) {
    const uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= N) return;
    const uint32_t *a = A + (size_t)row * sA;
    const uint32_t *b = B + (size_t)row * sB;
    uint32_t *c = C + (size_t)row * sC;
    uint32_t bv[LB], w[LB];
#pragma unroll
    for (int j = 0; j < LB; ++j) {
        bv[j] = b[j];
        w[j] = 0;
    }
    for (uint32_t i = 0; i < LA; ++i) {
        const uint32_t ai = a[i];
        uint32_t carry = 0;
#pragma unroll
        for (int j = 0; j < LB; ++j) {
            const uint64_t t = (uint64_t)ai * bv[j] + w[j] + carry;
            w[j] = (uint32_t)t;
            carry = (uint32_t)(t >> 32);
        }
        c[i] = w[0];
#pragma unroll
        for (int j = 0; j + 1 < LB; ++j) w[j] = w[j + 1];
        w[LB - 1] = carry;
    }
#pragma unroll
    for (int j = 0; j < LB; ++j) c[LA + j] = w[j];
}

constexpr uint32_t kPad = 4;

// Dynamic LDS: a[LA] | 4 zero words | b[LB] | 4 zero words (b indexed -4..LB+3).
// One wave per product; lane l computes the CPL columns k0 = g0 + CPL l ..
// (CPL = 2 for products of at most 128 limbs, so that every lane works; 4
// otherwise), sliding a CPL-wide window of b down the a loop.  Column k =
// x + 2^32 y + 2^64 z becomes limbs as x_k + y_(k-1) + z_(k-2), and the
// remaining carries go through one ballot resolution per 64 CPL columns.
template <int CPL>
__global__ void __launch_bounds__(64) mul_wave_kernel(const uint32_t *A, const uint32_t *B, uint32_t *C,
                                                      uint32_t LA, uint32_t LB, uint32_t sA,
                                                      uint32_t sB, uint32_t sC) {
    static_assert(CPL == 2 || CPL == 4, "two or four columns per lane");
    extern __shared__ uint32_t lds[];
    const uint32_t row = blockIdx.x;
    const uint32_t lane = threadIdx.x;
    uint32_t *a_s = lds;
    uint32_t *b_s = lds + LA + kPad;  // b_s[-4..-1] are the padding after a
    const uint32_t *a = A + (size_t)row * sA;
    const uint32_t *b = B + (size_t)row * sB;
    for (uint32_t i = lane; i < LA; i += kWave) a_s[i] = a[i];
    for (uint32_t i = lane; i < LB; i += kWave) b_s[i] = b[i];
    if (lane < kPad) {
        b_s[(int)lane - (int)kPad] = 0;
        b_s[LB + lane] = 0;
    }
    __syncthreads();

    const uint32_t Lc = LA + LB;
    uint32_t *c = C + (size_t)row * sC;
    // state handed from group to group (lane 63's top columns)
    uint32_t prev_y = 0, prev_z2 = 0, prev_z1 = 0, prev_c = 0, bin = 0;
    for (uint32_t g0 = 0; g0 < Lc; g0 += CPL * kWave) {
        const int64_t k0 = (int64_t)g0 + CPL * lane;
        uint32_t x[CPL];
        uint64_t h[CPL];
#pragma unroll
        for (int j = 0; j < CPL; ++j) {
            x[j] = 0;
            h[j] = 0;
        }
        int64_t ilo = k0 - (int64_t)LB + 1;
        if (ilo < 0) ilo = 0;
        int64_t ihi = k0 + CPL - 1;
        if (ihi > (int64_t)LA - 1) ihi = (int64_t)LA - 1;
        if (ilo <= ihi) {
            // window bw[j] = b[k0 + j - i]
            uint32_t bw[CPL];
#pragma unroll
            for (int j = 0; j < CPL; ++j) bw[j] = b_s[k0 + j - ilo];
            for (int64_t i = ilo; i <= ihi; ++i) {
                const uint32_t ai = a_s[i];
#pragma unroll
                for (int j = 0; j < CPL; ++j) {
                    const uint64_t t = (uint64_t)ai * bw[j] + x[j];
                    x[j] = (uint32_t)t;
                    h[j] += t >> 32;
                }
#pragma unroll
                for (int j = CPL - 1; j > 0; --j) bw[j] = bw[j - 1];
                bw[0] = b_s[k0 - i - 1];
            }
        }
        uint32_t y[CPL], z[CPL];
#pragma unroll
        for (int j = 0; j < CPL; ++j) {
            y[j] = (uint32_t)h[j];
            z[j] = (uint32_t)(h[j] >> 32);
        }
        // the previous lane's top columns (lane 0 takes the previous group's lane 63)
        uint32_t ly = __shfl_up(y[CPL - 1], 1), lz2 = __shfl_up(z[CPL - 2], 1), lz1 = __shfl_up(z[CPL - 1], 1);
        if (lane == 0) {
            ly = prev_y;
            lz2 = prev_z2;
            lz1 = prev_z1;
        }
        uint64_t s[CPL];
        s[0] = (uint64_t)x[0] + ly + lz2;
        s[1] = (uint64_t)x[1] + y[0] + lz1;
#pragma unroll
        for (int j = 2; j < CPL; ++j) s[j] = (uint64_t)x[j] + y[j - 1] + z[j - 2];
        uint32_t d[CPL], cc[CPL];
#pragma unroll
        for (int j = 0; j < CPL; ++j) {
            d[j] = (uint32_t)s[j];
            cc[j] = (uint32_t)(s[j] >> 32);
        }
        uint32_t lc = __shfl_up(cc[CPL - 1], 1);
        if (lane == 0) lc = prev_c;
        uint32_t e[CPL], gen[CPL];
#pragma unroll
        for (int j = 0; j < CPL; ++j) {
            const uint64_t u = (uint64_t)d[j] + (j == 0 ? lc : cc[j - 1]);
            e[j] = (uint32_t)u;
            gen[j] = (uint32_t)(u >> 32);
        }
        // lane aggregate of its CPL limbs for carry-in 0
        uint32_t lg = 0;
        bool lp = true;
#pragma unroll
        for (int j = 0; j < CPL; ++j) {
            const bool p = e[j] == 0xffffffffu;
            lg = gen[j] | (lg & (uint32_t)p);
            lp = lp && p;
        }
        const uint64_t G = ballot64(lg != 0), P = ballot64(lp && lg == 0);
        uint32_t cout;
        const uint64_t cm = resolve_carries(G, P, bin, &cout);
        uint32_t ci = (uint32_t)(cm >> lane) & 1u;
#pragma unroll
        for (int j = 0; j < CPL; ++j) {
            const uint32_t r = e[j] + ci;
            const uint32_t co = gen[j] | (ci & (uint32_t)(e[j] == 0xffffffffu));
            const int64_t k = k0 + j;
            if (k < (int64_t)Lc) c[k] = r;
            ci = co;
        }
        prev_y = __shfl(y[CPL - 1], 63);
        prev_z2 = __shfl(z[CPL - 2], 63);
        prev_z1 = __shfl(z[CPL - 1], 63);
        prev_c = __shfl(cc[CPL - 1], 63);
        bin = cout;
    }
}

__global__ void zero_rows_kernel(uint32_t *C, uint32_t N, uint32_t L, uint32_t sC) {
    const uint64_t total = (uint64_t)N * L;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (uint64_t)blockDim.x * gridDim.x) {
        const uint64_t r = i / L;
        C[r * sC + (i - r * L)] = 0;
    }
}

template <int LB>
void launch_thread_variant(const uint32_t *A, const uint32_t *B, uint32_t *C, uint32_t N, uint32_t LA,
                           uint32_t sA, uint32_t sB, uint32_t sC, hipStream_t stream) {
    hipLaunchKernelGGL(mul_thread_kernel<LB>, dim3(ceil_div(N, 256)), dim3(256), 0, stream, A, B, C, N, LA,
                       sA, sB, sC);
}

typedef void (*ThreadLauncher)(const uint32_t *, const uint32_t *, uint32_t *, uint32_t, uint32_t, uint32_t,
                               uint32_t, uint32_t, hipStream_t);
const ThreadLauncher kThreadLaunchers[kThreadMaxShort + 1] = {
    nullptr, launch_thread_variant<1>, launch_thread_variant<2>, launch_thread_variant<3>,
    launch_thread_variant<4>, launch_thread_variant<5>, launch_thread_variant<6>, launch_thread_variant<7>,
    launch_thread_variant<8>, launch_thread_variant<9>, launch_thread_variant<10>,
    launch_thread_variant<11>, launch_thread_variant<12>, launch_thread_variant<13>,
    launch_thread_variant<14>, launch_thread_variant<15>, launch_thread_variant<16>};

}  // namespace

namespace {
// mul_wave_kernel<4> may take up to 160 KB of dynamic LDS; the attribute is
// a property of the current device, set once per device (serving threads
// launch concurrently).
std::mutex g_lds_attr_mu;
uint64_t g_lds_attr_done = 0;  // bit d: set on device d

hipError_t allow_large_lds() {
    int dev = 0;
    GPUPI_RETURN_IF(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_lds_attr_mu);
    const uint64_t bit = dev >= 0 && dev < 64 ? 1ull << dev : 0ull;
    if (bit != 0 && (g_lds_attr_done & bit) != 0) return hipSuccess;
    GPUPI_RETURN_IF(hipFuncSetAttribute((const void *)mul_wave_kernel<4>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                        160 * 1024));
    g_lds_attr_done |= bit;
    return hipSuccess;
}
}  // namespace

bool mul_naive_supports(uint32_t L_a, uint32_t L_b) {
    return (uint64_t)L_a + L_b <= kWaveMaxTotal || (L_a <= kThreadMaxShort || L_b <= kThreadMaxShort);
}

hipError_t launch_mul_naive(const uint32_t *A, const uint32_t *B, uint32_t *C, uint32_t N, uint32_t L_a,
                            uint32_t L_b, uint32_t stride_A, uint32_t stride_B, uint32_t stride_C,
                            hipStream_t stream) {
    if (N == 0) return hipSuccess;
    if (L_a < L_b) {
        const uint32_t *t = A;
        A = B;
        B = t;
        uint32_t tl = L_a;
        L_a = L_b;
        L_b = tl;
        tl = stride_A;
        stride_A = stride_B;
        stride_B = tl;
    }
    if (L_b == 0) {
        if (L_a == 0) return hipSuccess;
        hipLaunchKernelGGL(zero_rows_kernel, dim3(ceil_div((uint64_t)N * L_a, 256) > 65536 ? 65536 : ceil_div((uint64_t)N * L_a, 256)),
                           dim3(256), 0, stream, C, N, L_a, stride_C);
        return hipGetLastError();
    }
    const uint64_t total = (uint64_t)L_a + L_b;
    const bool thread_ok = L_b <= kThreadMaxShort;
    const bool wave_ok = total <= kWaveMaxTotal;
    if (thread_ok && (!wave_ok || L_a <= 128 || N >= 16384)) {
        kThreadLaunchers[L_b](A, B, C, N, L_a, stride_A, stride_B, stride_C, stream);
        return hipGetLastError();
    }
    if (!wave_ok) return hipErrorInvalidValue;
    if (N > 0x7fffffffu) return hipErrorInvalidValue;
    const size_t lds = (size_t)(L_a + kPad + L_b + kPad) * sizeof(uint32_t);
    if (lds > 65536) GPUPI_RETURN_IF(allow_large_lds());
    if (total <= 2u * kWave) {
        hipLaunchKernelGGL(mul_wave_kernel<2>, dim3(N), dim3(kWave), lds, stream, A, B, C, L_a, L_b, stride_A,
                           stride_B, stride_C);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(mul_wave_kernel<4>, dim3(N), dim3(kWave), lds, stream, A, B, C, L_a, L_b, stride_A,
                       stride_B, stride_C);
    return hipGetLastError();
}

}  // namespace gpupi

extern "C" void batch_mul_naive(uint32_t *A, uint32_t *B, uint32_t *ret, int N, int L_a, int L_b,
                                int stride_A, int stride_B, int stride_ret) {
    if (N <= 0 || L_a < 0 || L_b < 0) return;
    (void)gpupi::launch_mul_naive(A, B, ret, (uint32_t)N, (uint32_t)L_a, (uint32_t)L_b, (uint32_t)stride_A,
                                  (uint32_t)stride_B, (uint32_t)stride_ret, 0);
}
