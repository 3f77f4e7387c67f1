// Bit shift, exact division by an odd word, and batch length reductions.
#include <algorithm>

#include "common.hpp"

namespace gpupi {
namespace {

// Word j of (A << s) or (A >> -s) for one row (reference batch_shift.h:5-15).
__device__ __forceinline__ uint32_t shifted_word(const uint32_t *a, uint32_t L_in, uint64_t j,
                                                 int64_t shift) {
    if (shift >= 0) {
        const uint64_t w = (uint64_t)shift >> 5;
        const uint32_t r = (uint32_t)shift & 31u;
        if (j < w) return 0u;
        const uint64_t i = j - w;
        const uint32_t hi = i < L_in ? a[i] : 0u;
        if (r == 0) return hi;
        const uint32_t lo = (i >= 1 && i - 1 < L_in) ? a[i - 1] : 0u;
        return (hi << r) | (lo >> (32u - r));
    }
    const uint64_t s = (uint64_t)(-shift);
    const uint64_t i = j + (s >> 5);
    const uint32_t r = (uint32_t)s & 31u;
    const uint32_t lo = i < L_in ? a[i] : 0u;
    if (r == 0) return lo;
    const uint32_t hi = i + 1 < L_in ? a[i + 1] : 0u;
    return (lo >> r) | (hi << (32u - r));
}

__global__ void __launch_bounds__(256) shift_bits_kernel(const uint32_t *A, uint32_t *B, uint32_t L_in,
                                                         uint32_t L_out, uint32_t stride_in,
                                                         uint32_t stride_out, int64_t shift,
                                                         uint64_t total) {
    for (uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (uint64_t)blockDim.x * gridDim.x) {
        const uint64_t row = idx / L_out;
        const uint64_t j = idx - row * L_out;
        B[row * stride_out + j] = shifted_word(A + row * stride_in, L_in, j, shift);
    }
}

// 64 rows per block; rows staged through LDS so global accesses stay
// contiguous, then one thread runs each row's Hensel division
// (q_i = (a_i - borrow) * b^-1 mod 2^32, borrow = hi(q_i b) + underflow).
constexpr int kDivRows = 64;
constexpr int kDivMaxL = 128;

__global__ void __launch_bounds__(64) exactdiv_kernel(uint32_t *A, uint32_t b, uint32_t binv, uint32_t N,
                                                      uint32_t L, uint32_t stride_A) {
    __shared__ uint32_t tile[kDivRows * (kDivMaxL + 1)];
    const uint32_t r0 = blockIdx.x * kDivRows;
    const uint32_t rows = min((uint32_t)kDivRows, N - r0);
    const uint32_t pitch = L + 1;
    for (uint32_t idx = threadIdx.x; idx < rows * L; idx += blockDim.x) {
        const uint32_t r = idx / L, j = idx - r * L;
        tile[r * pitch + j] = A[(size_t)(r0 + r) * stride_A + j];
    }
    __syncthreads();
    if (threadIdx.x < rows) {
        uint32_t *t = tile + threadIdx.x * pitch;
        uint32_t borrow = 0;
        for (uint32_t j = 0; j < L; ++j) {
            const uint32_t a = t[j];
            const uint32_t s = a - borrow;
            const uint32_t under = a < borrow;
            const uint32_t q = s * binv;
            t[j] = q;
            borrow = __umulhi(q, b) + under;
        }
    }
    __syncthreads();
    for (uint32_t idx = threadIdx.x; idx < rows * L; idx += blockDim.x) {
        const uint32_t r = idx / L, j = idx - r * L;
        A[(size_t)(r0 + r) * stride_A + j] = tile[r * pitch + j];
    }
}

__device__ __forceinline__ uint32_t wave_max(uint32_t v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = max(v, (uint32_t)__shfl_xor(v, off));
    return v;
}

__device__ __forceinline__ uint32_t length_of(uint32_t top_index_plus_one, uint32_t top_word, bool bits) {
    if (top_index_plus_one == 0) return 0u;
    if (!bits) return top_index_plus_one;
    return (top_index_plus_one - 1) * 32u + (32u - (uint32_t)__clz(top_word));
}

// Batch length reductions: a grid-stride loop over rows with the maximum
// reduced per block (one atomic per block).  A thread or wave stops as soon
// as it has seen the largest possible value.
__device__ __forceinline__ void block_max_atomic(uint32_t v, uint32_t *result) {
    __shared__ uint32_t wmax[4];
    v = wave_max(v);
    const uint32_t wave = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) wmax[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t m = wmax[0];
        for (uint32_t w = 1; w < blockDim.x / kWave; ++w) m = max(m, wmax[w]);
        if (m != 0) atomicMax(result, m);
    }
}

// Rows of at most 64 limbs: one thread per row scanning down from the top.
__global__ void __launch_bounds__(256) length_thread_kernel(const uint32_t *A, uint32_t N, uint32_t L,
                                                            uint32_t stride_A, int bits, uint32_t *result) {
    const uint32_t top = bits ? 32u * L : L;
    uint32_t best = 0;
    for (uint32_t row = blockIdx.x * blockDim.x + threadIdx.x; row < N && best < top;
         row += blockDim.x * gridDim.x) {
        const uint32_t *a = A + (size_t)row * stride_A;
        uint32_t j = L;
        while (j > 0 && a[j - 1] == 0u) --j;
        best = max(best, length_of(j, j ? a[j - 1] : 0u, bits != 0));
    }
    block_max_atomic(best, result);
}

// Longer rows: one wave per row, 64 limbs per step from the top down.
__global__ void __launch_bounds__(256) length_wave_kernel(const uint32_t *A, uint32_t N, uint32_t L,
                                                          uint32_t stride_A, int bits, uint32_t *result) {
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t waves = gridDim.x * (blockDim.x / kWave);
    const uint32_t topv = bits ? 32u * L : L;
    uint32_t best = 0;
    for (uint32_t row = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; row < N && best < topv;
         row += waves) {
        const uint32_t *a = A + (size_t)row * stride_A;
        for (int64_t top = L; top > 0; top -= kWave) {
            // the rest of the row lies below this wave's best length so far
            if ((bits ? 32 * top : top) <= (int64_t)best) break;
            const int64_t j = top - kWave + (int64_t)lane;
            const uint32_t v = (j >= 0) ? a[j] : 0u;
            const uint64_t nz = ballot64(v != 0u);
            if (nz != 0) {
                const int hi_lane = 63 - __builtin_clzll(nz);
                const uint32_t w = __shfl(v, hi_lane);
                best = max(best, length_of((uint32_t)(top - kWave + hi_lane + 1), w, bits != 0));
                break;
            }
        }
    }
    block_max_atomic(best, result);
}

// Long rows: 4096-limb chunks counted from the top of each row.  Block g of
// a row scans chunks g, g + G, g + 2G, ... (G blocks per row), each only
// above the best length found so far, and stops at the first chunk that lies
// wholly below it, so the work is about the limbs above the best length in
// each row plus the zero tail above the top limb of the longest rows.  Rows are taken last first: in the binary-splitting tree the
// last rows (the largest terms) are the longest, so the best length is known
// before the shorter rows' zero tails are reached.  G grows for few rows, so
// that a lone row's zero tail (a Newton residual) is scanned by many blocks.
constexpr uint32_t kLenChunk = 4096;
constexpr uint32_t kLenMinBlocksPerRow = 1;
constexpr uint32_t kLenBlocks = 4096;  // blocks per launch when rows are few

__global__ void __launch_bounds__(256) length_chunk_kernel(const uint32_t *A, uint32_t N, uint32_t L,
                                                           uint32_t stride_A, uint32_t chunks_per_row,
                                                           uint32_t blocks_per_row, int bits, uint32_t *result) {
    __shared__ int64_t start;
    const uint32_t row = N - 1u - blockIdx.x / blocks_per_row;
    const uint32_t g = blockIdx.x % blocks_per_row;
    const uint32_t *a = A + (size_t)row * stride_A;
    for (uint32_t c = g; c < chunks_per_row; c += blocks_per_row) {
        const int64_t hi = (int64_t)L - (int64_t)c * kLenChunk;  // chunk covers [hi - 4096, hi)
        const int64_t lo = hi - (int64_t)kLenChunk > 0 ? hi - (int64_t)kLenChunk : 0;
        if (threadIdx.x == 0) {
            // a relaxed load: a stale value only scans more.  A limb below
            // index cur (limbs) or cur / 32 (bits) cannot beat the best so far.
            const uint32_t cur = __hip_atomic_load(result, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const int64_t first = bits ? (int64_t)(cur >> 5) : (int64_t)cur;
            start = first > lo ? first : lo;
        }
        __syncthreads();
        const int64_t from = start;
        if (from >= hi) return;  // lower chunks lie lower still
        uint32_t best = 0;
        for (int64_t j = from + threadIdx.x; j < hi; j += blockDim.x) {
            const uint32_t v = a[j];
            if (v != 0u) best = max(best, length_of((uint32_t)j + 1u, v, bits != 0));
        }
        // its barrier also orders this chunk's reads of `start` before the next write
        block_max_atomic(best, result);
    }
}

__global__ void fill_u32_kernel(uint32_t *p, uint64_t count, uint32_t value) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (uint64_t)blockDim.x * gridDim.x)
        p[i] = value;
}

uint32_t grid_for(uint64_t total, uint32_t block) {
    uint64_t g = ceil_div64(total, block);
    if (g > 65536ull * 16) g = 65536ull * 16;
    return (uint32_t)(g ? g : 1);
}

uint32_t odd_inverse_u32(uint32_t b) {
    uint32_t x = b;  // correct to 3 bits for odd b
    for (int i = 0; i < 5; ++i) x *= 2u - b * x;
    return x;
}

}  // namespace

hipError_t launch_shift_bits(const uint32_t *A, uint32_t *B, uint32_t N, uint32_t L_in, uint32_t L_out,
                             uint32_t stride_in, uint32_t stride_out, int64_t shift_bits,
                             hipStream_t stream) {
    const uint64_t total = (uint64_t)N * L_out;
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(shift_bits_kernel, dim3(grid_for(total, 256)), dim3(256), 0, stream, A, B, L_in,
                       L_out, stride_in, stride_out, shift_bits, total);
    return hipGetLastError();
}

hipError_t launch_exactdiv_small(uint32_t *A, uint32_t B, uint32_t N, uint32_t L, uint32_t stride_A,
                                 hipStream_t stream) {
    if (N == 0 || L == 0) return hipSuccess;
    if (L > (uint32_t)kDivMaxL || (B & 1u) == 0u) return hipErrorInvalidValue;
    hipLaunchKernelGGL(exactdiv_kernel, dim3(ceil_div(N, kDivRows)), dim3(kDivRows), 0, stream, A, B,
                       odd_inverse_u32(B), N, L, stride_A);
    return hipGetLastError();
}

hipError_t launch_length_max(const uint32_t *A, uint32_t N, uint32_t L, uint32_t stride_A, bool bits,
                             uint32_t *d_result, hipStream_t stream) {
    if (N == 0 || L == 0) return hipSuccess;
    if (L > 4 * kLenChunk) {
        const uint32_t cpr = ceil_div(L, kLenChunk);
        const uint32_t bpr = std::min(cpr, std::max(kLenMinBlocksPerRow, ceil_div(kLenBlocks, N)));
        if ((uint64_t)N * bpr > 0x7fffffffull) return hipErrorInvalidValue;
        hipLaunchKernelGGL(length_chunk_kernel, dim3(N * bpr), dim3(256), 0, stream, A, N, L, stride_A, cpr, bpr,
                           bits ? 1 : 0, d_result);
    } else if (L <= 64) {
        hipLaunchKernelGGL(length_thread_kernel, dim3(std::min<uint32_t>(ceil_div(N, 256), 2048)), dim3(256), 0,
                           stream, A, N, L, stride_A, bits ? 1 : 0, d_result);
    } else {
        hipLaunchKernelGGL(length_wave_kernel, dim3(std::min<uint32_t>(ceil_div(N, 4), 4096)), dim3(256), 0,
                           stream, A, N, L, stride_A, bits ? 1 : 0, d_result);
    }
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t *p, uint64_t count, uint32_t value, hipStream_t stream) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_for(count, 256)), dim3(256), 0, stream, p, count, value);
    return hipGetLastError();
}

}  // namespace gpupi

// ---------------------------------------------------------------------------
// C entry points (reference batch_shift.h, batch_exactdiv_small.h,
// batch_bitlength.h); null stream.
// ---------------------------------------------------------------------------
extern "C" {

void batch_shift_bits(const uint32_t *A, uint32_t *B, uint32_t N, uint32_t L_in, uint32_t L_out,
                      uint32_t stride_in, uint32_t stride_out, int32_t shift_bits) {
    (void)gpupi::launch_shift_bits(A, B, N, L_in, L_out, stride_in, stride_out, shift_bits, 0);
}

void batch_exactdiv_small(uint32_t *A, uint32_t B, uint32_t N, uint32_t L, uint32_t stride_A) {
    (void)gpupi::launch_exactdiv_small(A, B, N, L, stride_A, 0);
}

size_t batch_bitlength_workspace_size(uint32_t, uint32_t) { return sizeof(uint32_t); }

static uint32_t length_entry(const uint32_t *A, uint32_t *ws, uint32_t N, uint32_t L, uint32_t sA,
                             bool bits) {
    uint32_t h = 0;
    if (ws == nullptr) return 0;
    if (gpupi::launch_fill_u32(ws, 1, 0u, 0) != hipSuccess) return 0;
    if (gpupi::launch_length_max(A, N, L, sA, bits, ws, 0) != hipSuccess) return 0;
    if (hipMemcpy(&h, ws, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess) return 0;
    return h;
}

uint32_t batch_bitlength_max(const uint32_t *A, uint32_t *workspace, uint32_t N, uint32_t L,
                             uint32_t stride_A) {
    return length_entry(A, workspace, N, L, stride_A, true);
}

uint32_t batch_limblength_max(const uint32_t *A, uint32_t *workspace, uint32_t N, uint32_t L,
                              uint32_t stride_A) {
    return length_entry(A, workspace, N, L, stride_A, false);
}

}  // extern "C"
