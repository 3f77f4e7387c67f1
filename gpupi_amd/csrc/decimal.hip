// Binary fraction -> decimal digits: the leaf converter (reference
// batch_decimal_naive.cu:16-217) and the equal-split radix conversion
// (reference convert_decimal.cu:67-474).
//
// Leaf: X / 2^(32 L) -> digits by repeated multiplication by 10^9; the word
// that overflows each multiplication is the next nine digits.  Rows of at
// most 64 limbs run one thread per row with the row held in LDS
// (word-interleaved across the wave), and the block's digits are staged in
// LDS so the global writes are contiguous.  Longer rows (up to 2047 limbs)
// run one wave per row: each x10^9 step is a wave-wide multiply-by-word with
// a ballot carry chain.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "common.hpp"
#include "convert_decimal.h"

namespace gpupi {
namespace {

constexpr uint32_t kBillion = 1000000000u;
constexpr uint32_t kLeafThreadMaxLimbs = 64;
constexpr uint32_t kLeafRowsPerBlock = 64;
constexpr uint32_t kLeafMaxDigitsStaged = 640;  // digits per row staged in LDS

__device__ __forceinline__ void emit_nine(char *dst, uint32_t v) {
    // dst[0..8] = nine decimal digits of v, most significant first
#pragma unroll
    for (int k = 8; k >= 0; --k) {
        dst[k] = (char)('0' + v % 10u);
        v /= 10u;
    }
}

// Digit j (natural index, 0 = 10^-1 place) goes to out[j] (natural) or
// out[L_dec - 1 - j] (reference order).
__device__ __forceinline__ void put_group(char *out, uint32_t L_dec, uint32_t start, uint32_t value,
                                          bool natural) {
    char buf[9];
    emit_nine(buf, value);
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const uint32_t j = start + (uint32_t)k;
        if (j < L_dec) out[natural ? j : (L_dec - 1u - j)] = buf[k];
    }
}

__global__ void __launch_bounds__(64) decimal_leaf_thread_kernel(const uint32_t *A, char *B, uint32_t N,
                                                                 uint32_t L, uint32_t sA, uint32_t L_dec,
                                                                 uint32_t sB, int natural, int staged) {
    __shared__ uint32_t words[kLeafThreadMaxLimbs * kLeafRowsPerBlock];
    __shared__ char digits[kLeafRowsPerBlock * kLeafMaxDigitsStaged];
    const uint32_t t = threadIdx.x;
    const uint32_t row0 = blockIdx.x * kLeafRowsPerBlock;
    const uint32_t rows = min(kLeafRowsPerBlock, N - row0);
    for (uint32_t idx = t; idx < rows * L; idx += blockDim.x) {
        const uint32_t r = idx / L, i = idx - r * L;
        words[i * kLeafRowsPerBlock + r] = A[(size_t)(row0 + r) * sA + i];
    }
    __syncthreads();
    if (t < rows) {
        char *out = staged ? digits + t * L_dec : B + (size_t)(row0 + t) * sB;
        for (uint32_t start = 0; start < L_dec; start += 9) {
            uint32_t carry = 0;
            for (uint32_t i = 0; i < L; ++i) {
                uint32_t *w = &words[i * kLeafRowsPerBlock + t];
                const uint64_t v = (uint64_t)(*w) * kBillion + carry;
                *w = (uint32_t)v;
                carry = (uint32_t)(v >> 32);
            }
            put_group(out, L_dec, start, carry, natural != 0);
        }
    }
    if (staged) {
        __syncthreads();
        if (sB == L_dec) {
            // rows are contiguous in the output: one flat copy
            char *dst = B + (size_t)row0 * sB;
            for (uint32_t idx = t; idx < rows * L_dec; idx += blockDim.x) dst[idx] = digits[idx];
        } else {
            for (uint32_t idx = t; idx < rows * L_dec; idx += blockDim.x) {
                const uint32_t r = idx / L_dec, j = idx - r * L_dec;
                B[(size_t)(row0 + r) * sB + j] = digits[idx];
            }
        }
    }
}

// One wave per row; the row lives in LDS (<= 2047 words).
__global__ void __launch_bounds__(64) decimal_leaf_wave_kernel(const uint32_t *A, char *B, uint32_t L,
                                                               uint32_t sA, uint32_t L_dec, uint32_t sB,
                                                               int natural) {
    extern __shared__ uint32_t w[];
    const uint32_t row = blockIdx.x;
    const uint32_t lane = threadIdx.x;
    for (uint32_t i = lane; i < L; i += kWave) w[i] = A[(size_t)row * sA + i];
    __syncthreads();
    char *out = B + (size_t)row * sB;
    for (uint32_t start = 0; start < L_dec; start += 9) {
        uint32_t hi_prev = 0;  // hi word of the product at limb seg-1
        uint32_t cin = 0;
        for (uint32_t seg = 0; seg < L; seg += kWave) {
            const uint32_t i = seg + lane;
            const bool valid = i < L;
            const uint64_t p = valid ? (uint64_t)w[i] * kBillion : 0ull;
            uint32_t hl = __shfl_up((uint32_t)(p >> 32), 1);
            if (lane == 0) hl = hi_prev;
            const uint32_t lo = (uint32_t)p;
            const uint32_t d = lo + hl;
            const bool g = valid && d < lo;
            const bool pr = valid && d == 0xffffffffu;
            uint32_t cout;
            const uint64_t cm = resolve_carries(ballot64(g), ballot64(pr), cin, &cout);
            if (valid) w[i] = d + ((uint32_t)(cm >> lane) & 1u);
            if (seg + kWave >= L) {
                // last segment: the integer part is hi(limb L-1) plus the
                // carry out of limb L-1 (lane `last`); it is below 10^9
                const uint32_t last = (L - 1u) - seg;
                const uint32_t hl_last = __shfl((uint32_t)(p >> 32), last);
                const uint32_t gl = __shfl((uint32_t)g, last);
                const uint32_t pl = __shfl((uint32_t)pr, last);
                const uint32_t cil = (uint32_t)(cm >> last) & 1u;
                hi_prev = hl_last + (gl | (pl & cil));
            } else {
                hi_prev = __shfl((uint32_t)(p >> 32), 63);
                cin = cout;
            }
        }
        __syncthreads();
        if (lane == 0) put_group(out, L_dec, start, hi_prev, natural != 0);
        __syncthreads();
    }
}

// convert_decimal.cu:67-96: dst row (2r + off) = window of src row r >> shift
__global__ void __launch_bounds__(256) extract_shifted_rows_kernel(const uint32_t *src, uint32_t src_stride,
                                                                   uint32_t src_length, uint32_t *dst,
                                                                   uint32_t dst_stride, uint32_t dst_length,
                                                                   uint32_t rows, uint32_t shift_bits,
                                                                   uint32_t dst_row_offset, uint32_t dst_row_step) {
    // rows * dst_length < 2^32 (checked at launch): 32-bit index arithmetic
    const uint32_t total = rows * dst_length;
    const uint32_t ws = shift_bits >> 5, bs = shift_bits & 31u;
    for (uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += blockDim.x * gridDim.x) {
        const uint32_t r = idx / dst_length;
        const uint32_t j = idx - r * dst_length;
        const uint64_t k = (uint64_t)j + ws;
        const uint32_t *s = src + (size_t)r * src_stride;
        uint64_t v = 0;
        if (k < src_length) {
            v = (uint64_t)s[k] >> bs;
            if (bs != 0 && k + 1 < src_length) v |= (uint64_t)s[k + 1] << (32u - bs);
        }
        dst[(size_t)(r * dst_row_step + dst_row_offset) * dst_stride + j] = (uint32_t)v;
    }
}

// convert_decimal.cu:98-119: low valid_bits of each row
__global__ void __launch_bounds__(256) extract_low_bits_rows_kernel(const uint32_t *src, uint32_t src_stride,
                                                                    uint32_t *dst, uint32_t dst_stride,
                                                                    uint32_t dst_length, uint32_t rows,
                                                                    uint32_t top_mask) {
    // rows * dst_length < 2^32 (checked at launch): 32-bit index arithmetic
    const uint32_t total = rows * dst_length;
    for (uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += blockDim.x * gridDim.x) {
        const uint32_t r = idx / dst_length;
        const uint32_t j = idx - r * dst_length;
        uint32_t v = src[(size_t)r * src_stride + j];
        if (j + 1u == dst_length) v &= top_mask;
        dst[(size_t)r * dst_stride + j] = v;
    }
}

// Limb `index` of each row ANDed with `mask`, in place.
__global__ void __launch_bounds__(256) mask_limb_rows_kernel(uint32_t *a, uint32_t stride, uint32_t rows,
                                                             uint32_t index, uint32_t mask) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < rows) a[(size_t)r * stride + index] &= mask;
}

uint32_t grid_for(uint64_t total) {
    uint64_t g = ceil_div64(total, 256);
    if (g > 1048576ull) g = 1048576ull;
    return (uint32_t)(g ? g : 1);
}

}  // namespace

hipError_t launch_decimal_leaf(const uint32_t *A, char *B, uint32_t N, uint32_t L_limb, uint32_t stride_A,
                               uint32_t L_dec, uint32_t stride_B, bool natural_order, hipStream_t stream) {
    if (N == 0 || L_limb == 0 || L_dec == 0) return hipSuccess;
    if (L_limb > BATCH_DECIMAL_NAIVE_L_LIMB_MAX) return hipErrorInvalidValue;
    if (L_limb <= kLeafThreadMaxLimbs) {
        const int staged = L_dec <= kLeafMaxDigitsStaged ? 1 : 0;
        hipLaunchKernelGGL(decimal_leaf_thread_kernel, dim3(ceil_div(N, kLeafRowsPerBlock)),
                           dim3(kLeafRowsPerBlock), 0, stream, A, B, N, L_limb, stride_A, L_dec, stride_B,
                           natural_order ? 1 : 0, staged);
    } else {
        if (N > 0x7fffffffu) return hipErrorInvalidValue;
        hipLaunchKernelGGL(decimal_leaf_wave_kernel, dim3(N), dim3(kWave), L_limb * sizeof(uint32_t), stream,
                           A, B, L_limb, stride_A, L_dec, stride_B, natural_order ? 1 : 0);
    }
    return hipGetLastError();
}

hipError_t launch_extract_shifted(const BatchMPArray &src, BatchMPArray &dst, uint32_t rows, uint32_t shift_bits,
                                  uint32_t dst_row_offset, uint32_t dst_row_step, hipStream_t stream) {
    if (dst.length == 0 || rows == 0) return hipSuccess;
    if ((uint64_t)rows * dst.length > 0xffffffffull) return hipErrorInvalidValue;
    hipLaunchKernelGGL(extract_shifted_rows_kernel, dim3(grid_for((uint64_t)rows * dst.length)), dim3(256), 0,
                       stream, src.data, src.stride, src.length, dst.data, dst.stride, dst.length, rows,
                       shift_bits, dst_row_offset, dst_row_step);
    return hipGetLastError();
}

hipError_t launch_keep_low_bits(BatchMPArray &a, uint32_t rows, uint32_t valid_bits, hipStream_t stream) {
    const uint32_t len = (uint32_t)ceil_div64(valid_bits, 32);
    if (len > a.length) return hipErrorInvalidValue;
    a.length = len;
    if (len == 0 || rows == 0 || (valid_bits & 31u) == 0) return hipSuccess;
    hipLaunchKernelGGL(mask_limb_rows_kernel, dim3(ceil_div(rows, 256)), dim3(256), 0, stream, a.data, a.stride, rows,
                       len - 1, (1u << (valid_bits & 31u)) - 1u);
    return hipGetLastError();
}

hipError_t launch_extract_low(const BatchMPArray &src, BatchMPArray &dst, uint32_t rows, uint32_t valid_bits,
                              hipStream_t stream) {
    if (dst.length == 0 || rows == 0) return hipSuccess;
    if ((uint64_t)rows * dst.length > 0xffffffffull) return hipErrorInvalidValue;
    uint32_t mask = 0xffffffffu;
    if (valid_bits & 31u) mask = (1u << (valid_bits & 31u)) - 1u;
    hipLaunchKernelGGL(extract_low_bits_rows_kernel, dim3(grid_for((uint64_t)rows * dst.length)), dim3(256), 0,
                       stream, src.data, src.stride, dst.data, dst.stride, dst.length, rows, mask);
    return hipGetLastError();
}

}  // namespace gpupi

extern "C" void batch_decimal_naive(const uint32_t *A, char *B, uint32_t N, int L_limb, uint32_t stride_A,
                                    int L_dec, uint32_t stride_B) {
    if (N == 0 || L_limb <= 0 || L_dec <= 0) return;
    (void)gpupi::launch_decimal_leaf(A, B, N, (uint32_t)L_limb, stride_A, (uint32_t)L_dec, stride_B, false, 0);
}
