// Context, arrays, validation and op dispatch: the C ABI of batch_arith.h
// (reference batch_arith.cu:18-424).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include <new>

#include "common.hpp"
#include "gpupi_pi.h"

namespace {

// reference batch_arith.cu:89-100
bool valid_array(const BatchMPArray &a) {
    if (a.batch_size == 0) return a.length <= a.stride;
    if (a.data == nullptr) return false;
    if (a.stride == 0) return true;
    return a.length <= a.stride;
}

// Outputs must not broadcast (every row would be written at one address).
bool valid_output(const BatchMPArray &c) {
    return valid_array(c) && (c.batch_size <= 1 || c.stride != 0);
}

bool same_batch(const BatchMPArray &a, const BatchMPArray &b) { return a.batch_size == b.batch_size; }

// In-place use of an operand is accepted only with C's exact view.
bool alias_ok(const BatchMPArray &x, const BatchMPArray &c) {
    if (x.data != c.data || x.length == 0) return true;  // a zero-length operand is never read
    return x.length == c.length && x.stride == c.stride;
}

hipError_t ensure(BatchMPContext *ctx, size_t bytes) {
    if (bytes <= ctx->workspace_size_bytes) return hipSuccess;
    if (ctx->cache->pool() != nullptr) {
        // stream order: the old workspace returns to the pool after the work
        // queued on it, the new one is ready for the work queued next
        if (ctx->workspace != nullptr) GPUPI_RETURN_IF(hipFreeAsync(ctx->workspace, ctx->stream));
        ctx->workspace = nullptr;
        ctx->workspace_size_bytes = 0;
        void *fresh = nullptr;
        GPUPI_RETURN_IF(hipMallocFromPoolAsync(&fresh, bytes, ctx->cache->pool(), ctx->stream));
        ctx->workspace = static_cast<uint32_t *>(fresh);
        ctx->workspace_size_bytes = bytes;
        return hipSuccess;
    }
    uint32_t *fresh = nullptr;
    GPUPI_RETURN_IF(hipMalloc(&fresh, bytes));
    if (ctx->workspace != nullptr) {
        hipError_t err = hipFree(ctx->workspace);
        if (err != hipSuccess) {
            (void)hipFree(fresh);
            return err;
        }
    }
    ctx->workspace = fresh;
    ctx->workspace_size_bytes = bytes;
    return hipSuccess;
}

hipError_t run_addsub(BatchMPContext *ctx, gpupi::AddSubKind kind, const BatchMPArray &A, const BatchMPArray *B,
                      uint32_t scalar, uint32_t shift, const BatchMPArray &C) {
    gpupi::AddSubArgs a{};
    a.kind = kind;
    a.A = A.data;
    a.B = B ? B->data : nullptr;
    a.C = C.data;
    a.scalar = scalar;
    a.shift = shift;
    a.N = C.batch_size;
    a.L_a = A.length;
    a.L_b = B ? B->length : 0;
    a.L_c = C.length;
    a.stride_A = A.stride;
    a.stride_B = B ? B->stride : 0;
    a.stride_C = C.stride;
    if (a.N == 0 || a.L_c == 0) return hipSuccess;
    GPUPI_RETURN_IF(ensure(ctx, gpupi::carry_flags_bytes(a.N, a.L_c)));
    return gpupi::launch_addsub(a, reinterpret_cast<uint8_t *>(ctx->workspace), ctx->stream);
}

hipError_t length_max(BatchMPContext *ctx, const BatchMPArray &A, bool bits, uint32_t *result) {
    if (ctx == nullptr || !valid_array(A) || result == nullptr) return hipErrorInvalidValue;
    if (A.batch_size == 0 || A.length == 0) {
        *result = 0;
        return hipSuccess;
    }
    GPUPI_RETURN_IF(gpupi::launch_fill_u32(ctx->d_scalar, 1, 0u, ctx->stream));
    GPUPI_RETURN_IF(gpupi::launch_length_max(A.data, A.batch_size, A.length, A.stride, bits, ctx->d_scalar,
                                             ctx->stream));
    GPUPI_RETURN_IF(hipMemcpyAsync(ctx->h_scalar, ctx->d_scalar, sizeof(uint32_t), hipMemcpyDeviceToHost,
                                   ctx->stream));
    GPUPI_RETURN_IF(hipStreamSynchronize(ctx->stream));
    *result = *ctx->h_scalar;
    return hipSuccess;
}

// Default schoolbook -> NTT crossover on L_a + L_b (tuned by measurement).
constexpr int kDefaultNttMinLimbs = 384;

// Scratch of the allocation-free multiply (batch_mp_mul_noalloc): the NTT's
// residue planes, spill words and carry flags.  Each device holds its own
// copy from module load to process end.  Host code never passes `scratch`
// itself to a kernel (on the host it is a stand-in, not a device address):
// it uses the address hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) uint8_t scratch[GPUPI_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices];

// This device's scratch address, looked up once per device (with that
// device current) and kept.
hipError_t scratch_of_current_device(void **out) {
    int dev = 0;
    GPUPI_RETURN_IF(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void *p = nullptr;
        const hipError_t err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (err != hipSuccess) return err;
        if (p == nullptr) return hipErrorInvalidSymbol;
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

}  // namespace

namespace gpupi {

namespace {
// Requests round up to a power of two below 1 MiB and to 2 MiB steps above,
// so blocks of neighbouring requests are interchangeable.
size_t round_block(size_t bytes) {
    if (bytes <= (1u << 20)) {
        size_t b = 256;
        while (b < bytes) b <<= 1;
        return b;
    }
    const size_t step = 2u << 20;
    return (bytes + step - 1) / step * step;
}
}  // namespace

hipError_t compact_arrays(BatchMPContext *ctx, BatchMPArray *const *arrays, int n) {
    if (ctx == nullptr || n < 0 || n > kScalarWords) return hipErrorInvalidValue;
    for (int i = 0; i < n; ++i)
        if (arrays[i] == nullptr || !valid_array(*arrays[i])) return hipErrorInvalidValue;
    GPUPI_RETURN_IF(launch_fill_u32(ctx->d_scalar, (uint64_t)n, 0u, ctx->stream));
    for (int i = 0; i < n; ++i) {
        const BatchMPArray &A = *arrays[i];
        if (A.batch_size == 0 || A.length == 0) continue;
        GPUPI_RETURN_IF(launch_length_max(A.data, A.batch_size, A.length, A.stride, false, ctx->d_scalar + i,
                                          ctx->stream));
    }
    GPUPI_RETURN_IF(hipMemcpyAsync(ctx->h_scalar, ctx->d_scalar, (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                   ctx->stream));
    GPUPI_RETURN_IF(hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < n; ++i) arrays[i]->length = ctx->h_scalar[i];
    return hipSuccess;
}

hipError_t BlockCache::acquire(hipStream_t stream, size_t bytes, void **ptr, size_t *got) {
    *ptr = nullptr;
    *got = 0;
    if (pool_ != nullptr) {
        const size_t want = round_block(bytes);
        GPUPI_RETURN_IF(hipMallocFromPoolAsync(ptr, want, pool_, stream));
        ++allocations_;
        *got = want;
        return hipSuccess;
    }
    if (!bound_ || stream != stream_) {
        // blocks last used on another stream: let the device drain first
        if (!free_.empty()) GPUPI_RETURN_IF(hipDeviceSynchronize());
        stream_ = stream;
        bound_ = true;
    }
    const size_t want = round_block(bytes);
    // best fit, wasting at most half of the block
    auto it = free_.lower_bound(want);
    if (it != free_.end() && it->first <= 2 * want) {
        *ptr = it->second;
        *got = it->first;
        cached_ -= it->first;
        free_.erase(it);
        return hipSuccess;
    }
    hipError_t err = hipMalloc(ptr, want);
    if (err == hipErrorOutOfMemory && !free_.empty()) {
        (void)hipGetLastError();
        GPUPI_RETURN_IF(trim());
        err = hipMalloc(ptr, want);
    }
    if (err != hipSuccess) {
        *ptr = nullptr;
        return err;
    }
    ++allocations_;
    *got = want;
    return hipSuccess;
}

void BlockCache::give_back(void *ptr, size_t bytes, hipStream_t stream) {
    if (ptr == nullptr) return;
    if (pool_ != nullptr) {
        (void)hipFreeAsync(ptr, stream);
        return;
    }
    free_.emplace(bytes, ptr);
    cached_ += bytes;
}

hipError_t ctx_alloc(BatchMPContext *ctx, void **ptr, size_t bytes) {
    if (ctx->cache->pool() != nullptr) return hipMallocFromPoolAsync(ptr, bytes, ctx->cache->pool(), ctx->stream);
    return hipMalloc(ptr, bytes);
}

void ctx_free(BatchMPContext *ctx, void *ptr) {
    if (ptr == nullptr) return;
    if (ctx->cache->pool() != nullptr) (void)hipFreeAsync(ptr, ctx->stream);
    else (void)hipFree(ptr);
}

hipError_t BlockCache::trim() {
    hipError_t first = hipSuccess;
    for (auto &kv : free_) {
        const hipError_t e = hipFree(kv.second);
        if (first == hipSuccess) first = e;
    }
    free_.clear();
    cached_ = 0;
    return first;
}

}  // namespace gpupi

extern "C" {

BatchMPArray batch_mp_array_create(uint32_t batch_size, uint32_t length, uint32_t stride) {
    BatchMPArray a{};
    a.length = length;
    a.batch_size = batch_size;
    a.stride = stride == 0 ? length : stride;
    if (a.length > a.stride) {
        a.stride = 0;
        a.length = 0;
        a.batch_size = 0;
        return a;
    }
    const size_t bytes = (size_t)a.batch_size * a.stride * sizeof(uint32_t);
    if (bytes == 0) return a;
    if (hipMalloc(&a.data, bytes) != hipSuccess) a = BatchMPArray{};
    return a;
}

void batch_mp_array_release(BatchMPArray array) {
    if (array.data != nullptr) (void)hipFree(array.data);
}

BatchMPContext *batch_mp_init(void) {
    BatchMPContext *ctx = new (std::nothrow) BatchMPContext{};
    if (ctx == nullptr) return nullptr;
    ctx->stream = nullptr;
    ctx->ntt_min_limbs = kDefaultNttMinLimbs;
    ctx->cache = new (std::nothrow) gpupi::BlockCache();
    if (ctx->cache == nullptr || gpupi::ntt_tables_init(&ctx->ntt_tables) != hipSuccess ||
        hipMalloc(&ctx->d_scalar, gpupi::kScalarWords * sizeof(uint32_t)) != hipSuccess ||
        hipHostMalloc(&ctx->h_scalar, gpupi::kScalarWords * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess) {
        batch_mp_destroy(ctx);
        return nullptr;
    }
    void *scratch_ptr = nullptr;
    if (hipDeviceSynchronize() != hipSuccess || scratch_of_current_device(&scratch_ptr) != hipSuccess) {
        batch_mp_destroy(ctx);
        return nullptr;
    }
    return ctx;
}

void batch_mp_destroy(BatchMPContext *ctx) {
    if (ctx == nullptr) return;
    if (ctx->workspace) {
        if (ctx->cache != nullptr && ctx->cache->pool() != nullptr) {
            (void)hipFreeAsync(ctx->workspace, ctx->stream);
            (void)hipStreamSynchronize(ctx->stream);
        } else {
            (void)hipFree(ctx->workspace);
        }
    }
    if (ctx->d_scalar) (void)hipFree(ctx->d_scalar);
    if (ctx->h_scalar) (void)hipHostFree(ctx->h_scalar);
    gpupi::ntt_tables_release(&ctx->ntt_tables);
    stage_profiler_free(ctx->profiler);
    tree_graph_free(ctx->tree_graph);
    delete ctx->cache;
    delete ctx;
}

size_t batch_mp_workspace_size(const BatchMPContext *ctx) { return ctx ? ctx->workspace_size_bytes : 0; }

hipError_t batch_mp_ensure_workspace(BatchMPContext *ctx, size_t required_bytes) {
    if (ctx == nullptr) return hipErrorInvalidValue;
    return ensure(ctx, required_bytes);
}

hipError_t batch_mp_set_stream(BatchMPContext *ctx, hipStream_t stream) {
    if (ctx == nullptr) return hipErrorInvalidValue;
    ctx->stream = stream;
    return hipSuccess;
}

hipStream_t batch_mp_get_stream(const BatchMPContext *ctx) { return ctx ? ctx->stream : nullptr; }

// Extension: schoolbook/NTT crossover (L_a + L_b at or above uses the NTT).
hipError_t batch_mp_set_ntt_threshold(BatchMPContext *ctx, uint32_t min_total_limbs) {
    if (ctx == nullptr) return hipErrorInvalidValue;
    ctx->ntt_min_limbs = (int)min_total_limbs;
    return hipSuccess;
}

namespace {

// One-row products whose transform is mostly padding, or whose shorter
// operand is far shorter than the longer (L the longer, S the shorter,
// K = 2^k >= L + S the transform), run in pieces of the longer operand that
// share the shorter one's transform.  Transform work counts 3 K for the
// product as it is, and (2 m + 1) K' for m pieces of K' points (m forward,
// m inverse, one of S):
//   two pieces at K' = K/2 when L + 2 S <= K (2.5 K): C[0, s + S) =
//     L[0, s) S directly, C[s, L + S) += L[s, L) S, s = K/2 - S;
//   m >= 3 pieces of s = K' - S >= S limbs at a smaller K' when that costs
//     at most 0.8 of 3 K: the pieces' products (m - 1 rows against S
//     broadcast, one batched product, and the last piece) go to even and odd
//     accumulators (pieces of one parity never overlap since s >= S), which
//     one full-length add sums.
// Transforms of 2^22 points and up for two pieces, 2^20 for more (smaller
// products are bound by launches, not by transform size).
constexpr int kSplitMinLog2 = 22, kSplitManyMinLog2 = 20;

struct SplitPlan {
    uint32_t m = 0;  // pieces: 0 = none
    uint32_t s = 0;  // piece length (the last may be shorter)
};

SplitPlan split_plan(uint32_t L, uint32_t S) {
    SplitPlan p;
    const int k = gpupi::ntt_transform_log2(L, S);
    const uint64_t K = 1ull << k;
    double best = 3.0 * (double)K;
    if (k >= kSplitMinLog2 && (uint64_t)L + 2ull * S <= K && (uint64_t)S < K / 2) {
        p.m = 2;
        p.s = (uint32_t)(K / 2 - S);
        best = 2.5 * (double)K;
    }
    if (k < kSplitManyMinLog2) return p;
    for (uint64_t Kp = K / 2; Kp >= 2ull * S && Kp >= (1ull << 13); Kp /= 2) {
        const uint64_t s = Kp - S;
        const uint64_t m = ((uint64_t)L + s - 1) / s;
        const double cost = (2.0 * (double)m + 1.0) * (double)Kp;
        if (m >= 3 && cost <= 0.8 * 3.0 * (double)K && cost < best) {
            best = cost;
            p.m = (uint32_t)m;
            p.s = (uint32_t)s;
        }
    }
    return p;
}

hipError_t mul_split_two(BatchMPContext *ctx, const BatchMPArray &L, const BatchMPArray &S, uint32_t s,
                         const BatchMPArray &C) {
    const uint32_t hi = L.length - s, lt = hi + S.length;
    const BatchMPArray L0{L.data, s, 1, s}, L1{L.data + s, hi, 1, hi};
    const BatchMPArray C0{C.data, s + S.length, 1, s + S.length}, Ct{C.data + s, lt, 1, lt};
    void *t = nullptr;
    size_t got = 0;
    GPUPI_RETURN_IF(ctx->cache->acquire(ctx->stream, (size_t)lt * sizeof(uint32_t), &t, &got));
    const BatchMPArray T{static_cast<uint32_t *>(t), lt, 1, lt};
    hipError_t err = gpupi::launch_fill_u32(C.data + s + S.length, hi, 0u, ctx->stream);
    if (err == hipSuccess) err = batch_mp_mul_shared(ctx, L0, L1, S, C0, T);
    if (err == hipSuccess) err = batch_mp_add(ctx, Ct, T, Ct);
    ctx->cache->give_back(t, got, ctx->stream);
    return err;
}

hipError_t mul_split_many(BatchMPContext *ctx, const BatchMPArray &L, const BatchMPArray &S, uint32_t m,
                          uint32_t s, const BatchMPArray &C) {
    const uint32_t w = s + S.length;                  // a piece's product
    const uint32_t last = L.length - (m - 1) * s;     // the last piece, 1..s limbs
    const size_t p_words = (size_t)m * w, c_words = C.length;
    void *pp = nullptr, *op = nullptr;
    size_t pgot = 0, ogot = 0;
    GPUPI_RETURN_IF(ctx->cache->acquire(ctx->stream, p_words * sizeof(uint32_t), &pp, &pgot));
    hipError_t err = ctx->cache->acquire(ctx->stream, c_words * sizeof(uint32_t), &op, &ogot);
    if (err != hipSuccess) {
        ctx->cache->give_back(pp, pgot, ctx->stream);
        return err;
    }
    uint32_t *P = static_cast<uint32_t *>(pp), *O = static_cast<uint32_t *>(op);
    // pieces 0..m-2 against S broadcast (S transformed once), then the last
    const BatchMPArray Ar{L.data, s, m - 1, s}, Sb{S.data, S.length, m - 1, 0}, Pr{P, w, m - 1, w};
    const BatchMPArray Al{L.data + (size_t)(m - 1) * s, last, 1, last};
    const BatchMPArray Pl{P + (size_t)(m - 1) * w, last + S.length, 1, last + S.length};
    err = batch_mp_mul(ctx, Ar, Sb, Pr);
    if (err == hipSuccess) err = batch_mp_mul(ctx, Al, S, Pl);
    // even pieces into C, odd ones into O, each at 32 s i bits
    if (err == hipSuccess) err = gpupi::launch_fill_u32(C.data, c_words, 0u, ctx->stream);
    if (err == hipSuccess) err = gpupi::launch_fill_u32(O, c_words, 0u, ctx->stream);
    const uint32_t full = m - 1;  // pieces copied as whole rows of w limbs
    for (uint32_t parity = 0; parity < 2 && err == hipSuccess; ++parity) {
        const uint32_t rows = full > parity ? (full - parity + 1) / 2 : 0;
        uint32_t *dst = parity ? O : C.data;
        if (rows > 0)
            err = hipMemcpy2DAsync(dst + (size_t)parity * s, (size_t)2 * s * sizeof(uint32_t), P + (size_t)parity * w,
                                   (size_t)2 * w * sizeof(uint32_t), (size_t)w * sizeof(uint32_t), rows,
                                   hipMemcpyDeviceToDevice, ctx->stream);
    }
    if (err == hipSuccess)
        err = hipMemcpyAsync(((m - 1) & 1 ? O : C.data) + (size_t)(m - 1) * s, Pl.data,
                             (size_t)Pl.length * sizeof(uint32_t), hipMemcpyDeviceToDevice, ctx->stream);
    const BatchMPArray Cf{C.data, C.length, 1, C.length}, Of{O, C.length, 1, C.length};
    if (err == hipSuccess) err = batch_mp_add(ctx, Cf, Of, Cf);
    ctx->cache->give_back(op, ogot, ctx->stream);
    ctx->cache->give_back(pp, pgot, ctx->stream);
    return err;
}

}  // namespace

hipError_t batch_mp_mul(BatchMPContext *ctx, BatchMPArray A, BatchMPArray B, BatchMPArray C) {
    if (ctx == nullptr || !valid_array(A) || !valid_array(B) || !valid_output(C)) return hipErrorInvalidValue;
    if (!same_batch(A, B) || !same_batch(A, C)) return hipErrorInvalidValue;
    if ((uint64_t)C.length != (uint64_t)A.length + B.length) return hipErrorInvalidValue;
    if (C.batch_size > 0 && (C.data == A.data || C.data == B.data)) return hipErrorInvalidValue;
    if (C.batch_size == 0 || C.length == 0) return hipSuccess;
    const uint32_t total = A.length + B.length;
    const bool small = A.length == 0 || B.length == 0 ||
                       ((int64_t)total < ctx->ntt_min_limbs && gpupi::mul_naive_supports(A.length, B.length));
    if (small || !gpupi::ntt_supports(A.length, B.length)) {
        if (!gpupi::mul_naive_supports(A.length, B.length)) return hipErrorInvalidValue;
        return gpupi::launch_mul_naive(A.data, B.data, C.data, C.batch_size, A.length, B.length, A.stride,
                                       B.stride, C.stride, ctx->stream);
    }
    if (C.batch_size == 1 && ctx->pipeline_form != GPUPI_FORM_PLAIN) {
        const bool a_long = A.length >= B.length;
        const BatchMPArray &L = a_long ? A : B, &S = a_long ? B : A;
        const SplitPlan sp = split_plan(L.length, S.length);
        if (sp.m == 2) return mul_split_two(ctx, L, S, sp.s, C);
        if (sp.m >= 3) return mul_split_many(ctx, L, S, sp.m, sp.s, C);
    }
    GPUPI_RETURN_IF(ensure(ctx, gpupi::ntt_workspace_bytes(C.batch_size, A.length, B.length)));
    return gpupi::launch_mul_ntt(A.data, B.data, C.data, ctx->workspace, ctx->workspace_size_bytes,
                                 ctx->ntt_tables, C.batch_size, A.length, B.length, A.stride, B.stride,
                                 C.stride, ctx->stream);
}

// batch_mp_mul without allocating: the NTT's scratch is this device's static
// array, rows run in chunks that fit it; the context's workspace and cache
// are not touched.
hipError_t batch_mp_mul_noalloc(BatchMPContext *ctx, BatchMPArray A, BatchMPArray B, BatchMPArray C) {
    if (ctx == nullptr || !valid_array(A) || !valid_array(B) || !valid_output(C)) return hipErrorInvalidValue;
    if (!same_batch(A, B) || !same_batch(A, C)) return hipErrorInvalidValue;
    if ((uint64_t)C.length != (uint64_t)A.length + B.length) return hipErrorInvalidValue;
    if (C.batch_size > 0 && (C.data == A.data || C.data == B.data)) return hipErrorInvalidValue;
    if (C.batch_size == 0 || C.length == 0) return hipSuccess;
    const uint32_t total = A.length + B.length;
    const bool small = A.length == 0 || B.length == 0 ||
                       ((int64_t)total < ctx->ntt_min_limbs && gpupi::mul_naive_supports(A.length, B.length));
    if (small || !gpupi::ntt_supports(A.length, B.length)) {
        if (!gpupi::mul_naive_supports(A.length, B.length)) return hipErrorInvalidValue;
        return gpupi::launch_mul_naive(A.data, B.data, C.data, C.batch_size, A.length, B.length, A.stride,
                                       B.stride, C.stride, ctx->stream);
    }
    void *ws = nullptr;
    GPUPI_RETURN_IF(scratch_of_current_device(&ws));
    return gpupi::launch_mul_ntt_bounded(A.data, B.data, C.data, ws, GPUPI_NOALLOC_SCRATCH_BYTES, ctx->ntt_tables,
                                         C.batch_size, A.length, B.length, A.stride, B.stride, C.stride,
                                         ctx->stream);
}

hipError_t batch_mp_mul_shared(BatchMPContext *ctx, BatchMPArray A1, BatchMPArray A2, BatchMPArray B,
                               BatchMPArray C1, BatchMPArray C2) {
    if (ctx == nullptr || !valid_array(A1) || !valid_array(A2) || !valid_array(B) || !valid_output(C1) ||
        !valid_output(C2))
        return hipErrorInvalidValue;
    if (!same_batch(A1, C1) || !same_batch(A2, C2)) return hipErrorInvalidValue;
    const uint32_t rows = std::max(C1.batch_size, C2.batch_size);
    if (B.stride != 0 && B.batch_size < rows) return hipErrorInvalidValue;
    if (B.stride == 0 && rows > 0 && B.batch_size == 0) return hipErrorInvalidValue;
    if ((uint64_t)C1.length != (uint64_t)A1.length + B.length || (uint64_t)C2.length != (uint64_t)A2.length + B.length)
        return hipErrorInvalidValue;
    for (const BatchMPArray *c : {&C1, &C2})
        if (c->batch_size > 0 && (c->data == A1.data || c->data == A2.data || c->data == B.data))
            return hipErrorInvalidValue;
    if (C1.batch_size > 0 && C2.batch_size > 0 && C1.data == C2.data) return hipErrorInvalidValue;
    const auto ntt_sized = [&](uint32_t La) {
        const uint32_t total = La + B.length;
        return La > 0 && B.length > 0 &&
               !((int64_t)total < ctx->ntt_min_limbs && gpupi::mul_naive_supports(La, B.length));
    };
    if (C1.batch_size > 0 && C2.batch_size > 0 && C1.length > 0 && C2.length > 0 && ntt_sized(A1.length) &&
        ntt_sized(A2.length) && gpupi::ntt_shared_supports(A1.length, A2.length, B.length)) {
        GPUPI_RETURN_IF(ensure(ctx, gpupi::ntt_shared_workspace_bytes(A1.length, C1.batch_size, A2.length,
                                                                      C2.batch_size, B.length, B.stride)));
        return gpupi::launch_mul_ntt_shared(A1.data, A1.length, A1.stride, C1.batch_size, C1.data, C1.stride,
                                            A2.data, A2.length, A2.stride, C2.batch_size, C2.data, C2.stride,
                                            B.data, B.length, B.stride, ctx->workspace, ctx->workspace_size_bytes,
                                            ctx->ntt_tables, ctx->stream);
    }
    BatchMPArray B1 = B, B2 = B;
    B1.batch_size = C1.batch_size;
    B2.batch_size = C2.batch_size;
    GPUPI_RETURN_IF(batch_mp_mul(ctx, A1, B1, C1));
    return batch_mp_mul(ctx, A2, B2, C2);
}

hipError_t gpupi_mul_ragged(BatchMPContext *ctx, const GpupiMulItem *items, uint32_t count) {
    if (ctx == nullptr || (count > 0 && items == nullptr)) return hipErrorInvalidValue;
    // validate everything before launching anything
    std::vector<std::pair<int, uint32_t>> big;  // (log2 K, item) of the NTT products
    for (uint32_t i = 0; i < count; ++i) {
        const GpupiMulItem &it = items[i];
        const uint64_t total = (uint64_t)it.L_a + it.L_b;
        if (total == 0) continue;
        if (it.C == nullptr || total > 0xffffffffull || (it.L_a && it.A == nullptr) || (it.L_b && it.B == nullptr))
            return hipErrorInvalidValue;
        if (it.C == it.A || it.C == it.B) return hipErrorInvalidValue;
        const bool small = it.L_a == 0 || it.L_b == 0 ||
                           ((int64_t)total < ctx->ntt_min_limbs && gpupi::mul_naive_supports(it.L_a, it.L_b));
        if (small) {
            if (!gpupi::mul_naive_supports(it.L_a, it.L_b) && it.L_a && it.L_b) return hipErrorInvalidValue;
        } else {
            if (!gpupi::ntt_supports(it.L_a, it.L_b)) return hipErrorInvalidValue;
            big.emplace_back(gpupi::ntt_transform_log2(it.L_a, it.L_b), i);
        }
    }
    for (uint32_t i = 0; i < count; ++i) {
        const GpupiMulItem &it = items[i];
        const uint64_t total = (uint64_t)it.L_a + it.L_b;
        if (total == 0) continue;
        const bool small = it.L_a == 0 || it.L_b == 0 ||
                           ((int64_t)total < ctx->ntt_min_limbs && gpupi::mul_naive_supports(it.L_a, it.L_b));
        if (small)
            GPUPI_RETURN_IF(gpupi::launch_mul_naive(it.A, it.B, it.C, 1, it.L_a, it.L_b, it.L_a, it.L_b,
                                                    (uint32_t)total, ctx->stream));
    }
    // products that gain from running in pieces (batch_mp_mul splits them),
    // before the grouped launches below take the workspace
    if (ctx->pipeline_form != GPUPI_FORM_PLAIN) {
        std::vector<std::pair<int, uint32_t>> rest;
        for (const auto &b : big) {
            const GpupiMulItem &it = items[b.second];
            if (split_plan(std::max(it.L_a, it.L_b), std::min(it.L_a, it.L_b)).m >= 2) {
                const uint32_t total = it.L_a + it.L_b;
                GPUPI_RETURN_IF(batch_mp_mul(ctx, BatchMPArray{const_cast<uint32_t *>(it.A), it.L_a, 1, it.L_a},
                                             BatchMPArray{const_cast<uint32_t *>(it.B), it.L_b, 1, it.L_b},
                                             BatchMPArray{it.C, total, 1, total}));
            } else {
                rest.push_back(b);
            }
        }
        big.swap(rest);
    }
    if (big.empty()) return hipSuccess;
    std::stable_sort(big.begin(), big.end(),
                     [](const std::pair<int, uint32_t> &x, const std::pair<int, uint32_t> &y) { return x.first < y.first; });
    std::vector<GpupiMulItem> sorted(big.size());
    for (size_t i = 0; i < big.size(); ++i) sorted[i] = items[big[i].second];
    // workspace: the sorted descriptors, then the largest group's NTT scratch
    const size_t desc_bytes = (sorted.size() * sizeof(GpupiMulItem) + 255) & ~(size_t)255;
    size_t ntt_bytes = 0;
    for (size_t g0 = 0; g0 < big.size();) {
        size_t g1 = g0;
        while (g1 < big.size() && big[g1].first == big[g0].first) ++g1;
        ntt_bytes = std::max(ntt_bytes, gpupi::ntt_ragged_workspace_bytes((uint32_t)(g1 - g0), big[g0].first));
        g0 = g1;
    }
    GPUPI_RETURN_IF(ensure(ctx, desc_bytes + ntt_bytes));
    GpupiMulItem *desc = reinterpret_cast<GpupiMulItem *>(ctx->workspace);
    char *ntt_ws = reinterpret_cast<char *>(ctx->workspace) + desc_bytes;
    GPUPI_RETURN_IF(hipMemcpyAsync(desc, sorted.data(), sorted.size() * sizeof(GpupiMulItem), hipMemcpyHostToDevice,
                                   ctx->stream));
    GPUPI_RETURN_IF(hipStreamSynchronize(ctx->stream));
    for (size_t g0 = 0; g0 < big.size();) {
        size_t g1 = g0;
        while (g1 < big.size() && big[g1].first == big[g0].first) ++g1;
        GPUPI_RETURN_IF(gpupi::launch_mul_ntt_ragged(desc + g0, sorted.data() + g0, (uint32_t)(g1 - g0), big[g0].first,
                                                     ntt_ws, ntt_bytes, ctx->ntt_tables, ctx->stream));
        g0 = g1;
    }
    return hipSuccess;
}

hipError_t batch_mp_mul_small(BatchMPContext *ctx, BatchMPArray A, uint32_t B, BatchMPArray C) {
    if (ctx == nullptr || !valid_array(A) || !valid_output(C) || !same_batch(A, C)) return hipErrorInvalidValue;
    if (C.batch_size > 0 && A.data == C.data) return hipErrorInvalidValue;  // reference asserts A != C
    return run_addsub(ctx, gpupi::AddSubKind::MulSmall, A, nullptr, B, 0, C);
}

hipError_t batch_mp_exactdiv_small(BatchMPContext *ctx, BatchMPArray A, uint32_t B) {
    if (ctx == nullptr || !valid_output(A)) return hipErrorInvalidValue;
    if (B == 0u || (B & 1u) == 0u || A.length > 128u) return hipErrorInvalidValue;
    if (A.batch_size == 0 || A.length == 0) return hipSuccess;
    return gpupi::launch_exactdiv_small(A.data, B, A.batch_size, A.length, A.stride, ctx->stream);
}

hipError_t batch_mp_add(BatchMPContext *ctx, BatchMPArray A, BatchMPArray B, BatchMPArray C) {
    if (ctx == nullptr || !valid_array(A) || !valid_array(B) || !valid_output(C)) return hipErrorInvalidValue;
    if (!same_batch(A, B) || !same_batch(A, C) || !alias_ok(A, C) || !alias_ok(B, C)) return hipErrorInvalidValue;
    return run_addsub(ctx, gpupi::AddSubKind::Add, A, &B, 0, 0, C);
}

hipError_t batch_mp_add_small(BatchMPContext *ctx, BatchMPArray A, uint32_t B, BatchMPArray C) {
    if (ctx == nullptr || !valid_array(A) || !valid_output(C) || !same_batch(A, C) || !alias_ok(A, C))
        return hipErrorInvalidValue;
    return run_addsub(ctx, gpupi::AddSubKind::AddSmall, A, nullptr, B, 0, C);
}

hipError_t batch_mp_shift_add(BatchMPContext *ctx, BatchMPArray A, BatchMPArray B, uint32_t shift, BatchMPArray C) {
    if (ctx == nullptr || !valid_array(A) || !valid_array(B) || !valid_output(C)) return hipErrorInvalidValue;
    if (!same_batch(A, B) || !same_batch(A, C)) return hipErrorInvalidValue;
    if (C.data == A.data || C.data == B.data) return hipErrorInvalidValue;
    return run_addsub(ctx, gpupi::AddSubKind::ShiftAdd, A, &B, 0, shift, C);
}

hipError_t batch_mp_shift_sub(BatchMPContext *ctx, BatchMPArray A, BatchMPArray B, uint32_t shift, BatchMPArray C) {
    if (ctx == nullptr || !valid_array(A) || !valid_array(B) || !valid_output(C)) return hipErrorInvalidValue;
    if (!same_batch(A, B) || !same_batch(A, C)) return hipErrorInvalidValue;
    if (C.data == A.data || C.data == B.data) return hipErrorInvalidValue;
    return run_addsub(ctx, gpupi::AddSubKind::ShiftSub, A, &B, 0, shift, C);
}

hipError_t batch_mp_sub(BatchMPContext *ctx, BatchMPArray A, BatchMPArray B, BatchMPArray C) {
    if (ctx == nullptr || !valid_array(A) || !valid_array(B) || !valid_output(C)) return hipErrorInvalidValue;
    if (!same_batch(A, B) || !same_batch(A, C) || !alias_ok(A, C) || !alias_ok(B, C)) return hipErrorInvalidValue;
    return run_addsub(ctx, gpupi::AddSubKind::Sub, A, &B, 0, 0, C);
}

hipError_t batch_mp_sub_small(BatchMPContext *ctx, BatchMPArray A, uint32_t B, BatchMPArray C) {
    if (ctx == nullptr || !valid_array(A) || !valid_output(C) || !same_batch(A, C) || !alias_ok(A, C))
        return hipErrorInvalidValue;
    return run_addsub(ctx, gpupi::AddSubKind::SubSmall, A, nullptr, B, 0, C);
}

hipError_t batch_mp_shift_bits(BatchMPContext *ctx, BatchMPArray A, BatchMPArray B, int32_t shift_bits) {
    if (ctx == nullptr || !valid_array(A) || !valid_output(B) || !same_batch(A, B)) return hipErrorInvalidValue;
    if (B.batch_size > 0 && A.data == B.data && shift_bits != 0) return hipErrorInvalidValue;
    if (B.batch_size == 0 || B.length == 0) return hipSuccess;
    return gpupi::launch_shift_bits(A.data, B.data, B.batch_size, A.length, B.length, A.stride, B.stride,
                                    shift_bits, ctx->stream);
}

hipError_t batch_mp_bitlength_max(BatchMPContext *ctx, BatchMPArray A, uint32_t *result) {
    return length_max(ctx, A, true, result);
}

hipError_t batch_mp_limblength_max(BatchMPContext *ctx, BatchMPArray A, uint32_t *result) {
    return length_max(ctx, A, false, result);
}

hipError_t batch_mp_array_compact(BatchMPContext *ctx, BatchMPArray *array) {
    if (array == nullptr) return hipErrorInvalidValue;
    uint32_t len = 0;
    GPUPI_RETURN_IF(batch_mp_limblength_max(ctx, *array, &len));
    array->length = len;
    return hipSuccess;
}

}  // extern "C"
