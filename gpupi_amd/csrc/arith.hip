// Context, arrays, validation and op dispatch: the C ABI of batch_arith.h
// (reference batch_arith.cu:18-424).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
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
//   two pieces at K' = K/2 when L + 2 S <= K (2.5 K), s = K/2 - S;
//   m >= 3 pieces of s = K' - S >= S limbs at a smaller K' when that costs
//     at most 0.8 of 3 K.
// Two pieces are two products sharing S's transform, the first written in
// place.  More are the rows of one ragged launch sequence (piece_row): even
// pieces' products land in C and odd ones in an accumulator, each at its
// offset with zeros to the next piece of its parity.  Either way one add
// finishes.  Transforms of 2^22 points and up for two pieces, 2^20 for more
// (smaller products are bound by launches, not by transform size).
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

// Two pieces: C[0, s + S) = L[0, s) S directly, T = L[s, L) S, both one
// launch sequence sharing S's transform, then C[s..) += T.
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

// C = L * S in the pieces of `sp`: one launch sequence, S transformed once,
// the pieces' products assembled into C.
hipError_t mul_pieces(BatchMPContext *ctx, const BatchMPArray &L, const BatchMPArray &S, const SplitPlan &sp,
                      const BatchMPArray &C) {
    GPUPI_RETURN_IF(ensure(ctx, gpupi::ntt_pieces_workspace_bytes(L.length, S.length, sp.m, sp.s)));
    return gpupi::launch_mul_ntt_pieces(L.data, L.length, S.data, S.length, C.data, sp.m, sp.s, ctx->workspace,
                                        ctx->workspace_size_bytes, ctx->ntt_tables, ctx->stream);
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
        if (sp.m >= 3) return mul_pieces(ctx, L, S, sp, C);
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

namespace {
// The NTT launches of a ragged call: rows of whole products, and the
// pieces of split ones (plans[q].m >= 2) with their assemblies, grouped by
// transform size into launches filled to ntt_ragged_max_rows rows (a
// product's pieces stay in one launch), each with the set of distinct B
// operands its rows use.
hipError_t ragged_launches(BatchMPContext *ctx, const GpupiMulItem *items,
                           const std::vector<std::pair<int, uint32_t>> &big, const std::vector<SplitPlan> &plans) {
    // groups of rows: a whole product (one row) or a split one (its pieces)
    struct Group {
        int k;
        size_t q;
    };
    std::vector<Group> groups;
    for (size_t q = 0; q < big.size(); ++q) {
        const GpupiMulItem &it = items[big[q].second];
        const int k = plans[q].m < 2 ? big[q].first
                                     : gpupi::ntt_transform_log2(plans[q].s, std::min(it.L_a, it.L_b));
        groups.push_back(Group{k, q});
    }
    std::stable_sort(groups.begin(), groups.end(), [](const Group &x, const Group &y) { return x.k < y.k; });
    struct Launch {
        int k;
        size_t r0, n, b0, nb, a0, na;
        uint32_t chunks;
    };
    std::vector<Launch> launches;
    std::vector<gpupi::RaggedRow> host_rows;
    std::vector<gpupi::RaggedB> host_bset;
    std::vector<gpupi::RaggedAssembly> host_asms;
    std::map<std::pair<const uint32_t *, uint32_t>, uint32_t> bidx;
    const auto b_of = [&](Launch &l, const uint32_t *B, uint32_t LB) {
        auto f = bidx.find(std::make_pair(B, LB));
        if (f == bidx.end()) {
            f = bidx.emplace(std::make_pair(B, LB), (uint32_t)l.nb++).first;
            host_bset.push_back(gpupi::RaggedB{B, LB});
        }
        return f->second;
    };
    size_t ntt_bytes = 0;
    for (size_t g0 = 0; g0 < groups.size();) {
        const int k = groups[g0].k;
        const size_t cap = gpupi::ntt_ragged_max_rows(k);
        Launch l{k, host_rows.size(), 0, host_bset.size(), 0, host_asms.size(), 0, 0};
        bidx.clear();
        size_t g1 = g0;
        for (; g1 < groups.size() && groups[g1].k == k; ++g1) {
            const size_t q = groups[g1].q;
            const GpupiMulItem &it = items[big[q].second];
            const uint32_t m = plans[q].m < 2 ? 1u : plans[q].m;
            if (l.n > 0 && l.n + m > cap) break;
            if (m == 1) {
                host_rows.push_back(gpupi::RaggedRow{it.A, it.B, it.C, it.L_a, it.L_b, b_of(l, it.B, it.L_b), 0});
            } else {
                const bool a_long = it.L_a >= it.L_b;
                const uint32_t *Lp = a_long ? it.A : it.B, *Sp = a_long ? it.B : it.A;
                const uint32_t L = a_long ? it.L_a : it.L_b, S = a_long ? it.L_b : it.L_a;
                const uint32_t b = b_of(l, Sp, S);
                host_asms.push_back(gpupi::RaggedAssembly{it.C, L + S, plans[q].s, S, m, (uint32_t)l.n, l.chunks});
                l.chunks += gpupi::assembly_chunks(L + S);
                for (uint32_t i = 0; i < m; ++i) host_rows.push_back(gpupi::piece_row(i, Lp, L, Sp, S, m, plans[q].s, b));
            }
            l.n += m;
        }
        l.na = host_asms.size() - l.a0;
        ntt_bytes = std::max(ntt_bytes, gpupi::ntt_ragged_workspace_bytes((uint32_t)l.n, (uint32_t)l.nb, l.chunks, k));
        launches.push_back(l);
        g0 = g1;
    }
    // the descriptors at the head of the workspace, then the largest
    // launch's NTT scratch
    const size_t row_bytes = (host_rows.size() * sizeof(gpupi::RaggedRow) + 255) & ~(size_t)255;
    const size_t b_bytes = (host_bset.size() * sizeof(gpupi::RaggedB) + 255) & ~(size_t)255;
    const size_t a_bytes = (host_asms.size() * sizeof(gpupi::RaggedAssembly) + 255) & ~(size_t)255;
    GPUPI_RETURN_IF(ensure(ctx, row_bytes + b_bytes + a_bytes + ntt_bytes));
    char *ws = reinterpret_cast<char *>(ctx->workspace);
    auto *d_rows = reinterpret_cast<gpupi::RaggedRow *>(ws);
    auto *d_bset = reinterpret_cast<gpupi::RaggedB *>(ws + row_bytes);
    auto *d_asms = reinterpret_cast<gpupi::RaggedAssembly *>(ws + row_bytes + b_bytes);
    char *ntt_ws = ws + row_bytes + b_bytes + a_bytes;
    GPUPI_RETURN_IF(hipMemcpyAsync(d_rows, host_rows.data(), host_rows.size() * sizeof(gpupi::RaggedRow),
                                   hipMemcpyHostToDevice, ctx->stream));
    GPUPI_RETURN_IF(hipMemcpyAsync(d_bset, host_bset.data(), host_bset.size() * sizeof(gpupi::RaggedB),
                                   hipMemcpyHostToDevice, ctx->stream));
    if (!host_asms.empty())
        GPUPI_RETURN_IF(hipMemcpyAsync(d_asms, host_asms.data(), host_asms.size() * sizeof(gpupi::RaggedAssembly),
                                       hipMemcpyHostToDevice, ctx->stream));
    GPUPI_RETURN_IF(hipStreamSynchronize(ctx->stream));
    for (const Launch &l : launches)
        GPUPI_RETURN_IF(gpupi::launch_mul_ntt_ragged(
            d_rows + l.r0, host_rows.data() + l.r0, (uint32_t)l.n, d_bset + l.b0, host_bset.data() + l.b0,
            (uint32_t)l.nb, d_asms + l.a0, host_asms.data() + l.a0, (uint32_t)l.na, l.chunks, l.k, ntt_ws,
            ntt_bytes, ctx->ntt_tables, ctx->stream));
    return hipSuccess;
}
}  // namespace

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
    if (big.empty()) return hipSuccess;
    // NTT rows, one launch sequence per transform size (and row chunk):
    // whole products, and in the batched form the pieces of the products
    // that gain from running in pieces (split_plan), which share the
    // shorter operand's transform and are assembled from their staged
    // products into C.
    std::vector<SplitPlan> plans(big.size());
    if (ctx->pipeline_form != GPUPI_FORM_PLAIN)
        for (size_t q = 0; q < big.size(); ++q) {
            const GpupiMulItem &it = items[big[q].second];
            plans[q] = split_plan(std::max(it.L_a, it.L_b), std::min(it.L_a, it.L_b));
        }
    return ragged_launches(ctx, items, big, plans);
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
