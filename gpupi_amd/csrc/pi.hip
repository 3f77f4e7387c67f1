// The Chudnovsky pipeline on the device: leaf terms, the batched binary
// splitting tree, sqrt(10005), the inverse, the final division and the
// equal-split radix conversion.  The sequence of batched operations (and so
// every intermediate integer) follows the reference pi_bs_gpu.cu and
// convert_decimal.cu line for line where the values depend on it:
// truncations keyed on compacted lengths, the 2^15-per-term shift, the
// divisors of the 17-term block, the Pell doubling, and the approximate
// inverse.  Buffers are sized on the host from the compacted lengths, which
// every compaction (a synchronizing length reduction) brings back.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <initializer_list>
#include <memory>
#include <new>
#include <vector>

#include "common.hpp"
#include "convert_decimal.h"
#include "gpupi_pi.h"
#include "leaf_terms.hpp"

namespace gpupi {
hipError_t launch_extract_shifted(const BatchMPArray &src, BatchMPArray &dst, uint32_t rows, uint32_t shift_bits,
                                  uint32_t dst_row_offset, uint32_t dst_row_step, hipStream_t stream);
hipError_t launch_extract_low(const BatchMPArray &src, BatchMPArray &dst, uint32_t rows, uint32_t valid_bits,
                              hipStream_t stream);
hipError_t launch_keep_low_bits(BatchMPArray &a, uint32_t rows, uint32_t valid_bits, hipStream_t stream);
}  // namespace gpupi

namespace {

using gpupi::ceil_div;

// ------------------------------------------------------------------ leaf
// Head rows are terms 17t + 16, body rows terms t + t/16 (pi_bs_gpu.cu:40-46).
__global__ void __launch_bounds__(256) leaf17_kernel(int is_head, uint32_t *P, uint32_t *Q, uint32_t *R,
                                                     uint32_t count) {
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < count; t += blockDim.x * gridDim.x) {
        const int64_t i = is_head ? (int64_t)t * 17 + 16 : (int64_t)t + (int64_t)(t >> 4);
        gpupi::leaf_term(i, P + (size_t)t * 3, Q + (size_t)t * 4, R + (size_t)t * 6);
    }
}

// ------------------------------------------------------- buffer helpers

// A grow-only device buffer holding one batch array.  Attached to a context
// in the batched form, it takes its blocks from the context's block cache
// and hands them back (work on the context stream only); otherwise (no
// context, or the plain form) it allocates and frees them.
struct Buf {
    uint32_t *ptr = nullptr;
    size_t cap = 0;       // words
    bool frozen = false;  // set while a graph is being captured: no reallocation
    bool cached = false;  // ptr is a block of ctx's cache
    BatchMPContext *ctx = nullptr;
    Buf() = default;
    explicit Buf(BatchMPContext *c) : ctx(c) {}
    Buf(const Buf &) = delete;
    Buf &operator=(const Buf &) = delete;
    ~Buf() { drop(); }
    bool use_cache() const { return ctx != nullptr && ctx->pipeline_form != GPUPI_FORM_PLAIN; }
    void drop() {
        if (ptr == nullptr) return;
        if (cached) ctx->cache->give_back(ptr, cap * sizeof(uint32_t), ctx->stream);
        else (void)hipFree(ptr);
        ptr = nullptr;
        cap = 0;
        cached = false;
    }
    // Takes ownership of a block of `words` words handed over by a stage
    // entry point (a block of this context's cache in the batched form).
    void adopt(uint32_t *p, size_t words) {
        drop();
        ptr = p;
        cap = words;
        cached = p != nullptr && use_cache();
    }
    hipError_t reserve(size_t words) {
        if (words <= cap) return hipSuccess;
        if (frozen) return hipErrorInvalidValue;
        if (!use_cache()) {
            drop();
            GPUPI_RETURN_IF(hipMalloc(&ptr, std::max<size_t>(words, 1) * sizeof(uint32_t)));
            cap = words;
            return hipSuccess;
        }
        // a growing buffer takes an eighth more than asked for: the tree's
        // rows grow a little at every level
        const size_t want = cap > 0 ? words + words / 8 : words;
        drop();
        void *p = nullptr;
        size_t got = 0;
        GPUPI_RETURN_IF(ctx->cache->acquire(ctx->stream, std::max<size_t>(want, 1) * sizeof(uint32_t), &p, &got));
        ptr = static_cast<uint32_t *>(p);
        cap = got / sizeof(uint32_t);
        cached = true;
        return hipSuccess;
    }
    // A contiguous array view (stride = length) of `batch` rows.
    hipError_t view(uint32_t batch, uint32_t length, BatchMPArray &out) {
        GPUPI_RETURN_IF(reserve((size_t)batch * length));
        out.data = ptr;
        out.length = length;
        out.batch_size = batch;
        out.stride = length;
        return hipSuccess;
    }
    // The same with rows 16-byte aligned (stride a multiple of 4 limbs): the
    // NTT's final carry pass reads and writes such rows four limbs per lane.
    hipError_t view_aligned(uint32_t batch, uint32_t length, BatchMPArray &out) {
        if (batch <= 1) return view(batch, length, out);
        const uint32_t stride = (length + 3u) & ~3u;
        GPUPI_RETURN_IF(reserve((size_t)batch * stride));
        out.data = ptr;
        out.length = length;
        out.batch_size = batch;
        out.stride = stride;
        return hipSuccess;
    }
    // Buffers of one context (or both unattached) exchange their blocks.
    void swap(Buf &o) {
        std::swap(ptr, o.ptr);
        std::swap(cap, o.cap);
        std::swap(cached, o.cached);
    }
    uint32_t *release() {
        uint32_t *p = ptr;
        ptr = nullptr;
        cap = 0;
        cached = false;
        return p;
    }
};

BatchMPArray view_of(uint32_t *data, uint32_t length, uint32_t batch, uint32_t stride) {
    BatchMPArray a;
    a.data = data;
    a.length = length;
    a.batch_size = batch;
    a.stride = stride;
    return a;
}

// pi_bs_gpu.cu:1095-1110
BatchMPArray make_subview(const BatchMPArray &a, uint32_t start, uint32_t length) {
    if (start > a.length) start = a.length;
    const uint32_t max_len = a.length - start;
    if (length > max_len) length = max_len;
    return view_of(a.data + start, length, a.batch_size, a.stride >= start ? a.stride - start : 0u);
}

}  // namespace

// ------------------------------------------------------- stage profiling
// pi_bs_gpu.cu:133-242: with profiling on, each stage is bracketed by events
// on the context stream and synchronized; one entry per stage.
struct StageProfiler {
    bool enabled = false;
    std::vector<GpupiStageEntry> entries;
    hipEvent_t start = nullptr, stop = nullptr;
};

void stage_profiler_free(StageProfiler *p) {
    if (p == nullptr) return;
    if (p->start) (void)hipEventDestroy(p->start);
    if (p->stop) (void)hipEventDestroy(p->stop);
    delete p;
}

namespace {

template <typename F>
hipError_t stage(BatchMPContext *ctx, const char *kind, const char *label, uint32_t n1, uint32_t batch, uint32_t lhs,
                 uint32_t rhs, bool uses_ntt, F &&fn) {
    StageProfiler *p = ctx->profiler;
    if (p == nullptr || !p->enabled) return fn();
    GPUPI_RETURN_IF(hipEventRecord(p->start, ctx->stream));
    GPUPI_RETURN_IF(fn());
    GPUPI_RETURN_IF(hipEventRecord(p->stop, ctx->stream));
    GPUPI_RETURN_IF(hipEventSynchronize(p->stop));
    GpupiStageEntry e{};
    std::snprintf(e.kind, sizeof(e.kind), "%s", kind);
    std::snprintf(e.label, sizeof(e.label), "%s", label);
    e.n1 = n1;
    e.batch_size = batch;
    e.lhs_length = lhs;
    e.rhs_length = rhs;
    e.uses_ntt = uses_ntt ? 1 : 0;
    GPUPI_RETURN_IF(hipEventElapsedTime(&e.elapsed_ms, p->start, p->stop));
    p->entries.push_back(e);
    return hipSuccess;
}

// whether batch_mp_mul takes the NTT path for these lengths
bool mul_uses_ntt(const BatchMPContext *ctx, const BatchMPArray &A, const BatchMPArray &B) {
    if (A.length == 0 || B.length == 0) return false;
    const bool small = (int64_t)A.length + B.length < ctx->ntt_min_limbs && gpupi::mul_naive_supports(A.length, B.length);
    return !small && gpupi::ntt_supports(A.length, B.length);
}

hipError_t p_mul(BatchMPContext *ctx, const char *label, uint32_t n1, const BatchMPArray &A, const BatchMPArray &B,
                 const BatchMPArray &C) {
    return stage(ctx, "mul", label, n1, A.batch_size, A.length, B.length, mul_uses_ntt(ctx, A, B),
                 [&] { return batch_mp_mul(ctx, A, B, C); });
}

// C1 = A1 B and C2 = A2 B.  When both are NTT products of one transform
// size, one shared launch sequence (B transformed once) and one stage entry
// labelled "label1,label2"; otherwise the reference's two products, with B
// as the left operand when b_left (the order of pi_bs_gpu.cu's calls).
hipError_t p_mul_pair(BatchMPContext *ctx, const char *label, const char *label1, const char *label2, uint32_t n1,
                      bool b_left, const BatchMPArray &A1, const BatchMPArray &A2, BatchMPArray B,
                      const BatchMPArray &C1, const BatchMPArray &C2) {
    if (mul_uses_ntt(ctx, A1, B) && mul_uses_ntt(ctx, A2, B) &&
        gpupi::ntt_shared_supports(A1.length, A2.length, B.length))
        return stage(ctx, "mul", label, n1, A2.batch_size, A2.length, B.length, true,
                     [&] { return batch_mp_mul_shared(ctx, A1, A2, B, C1, C2); });
    B.batch_size = A1.batch_size;
    GPUPI_RETURN_IF(b_left ? p_mul(ctx, label1, n1, B, A1, C1) : p_mul(ctx, label1, n1, A1, B, C1));
    B.batch_size = A2.batch_size;
    return b_left ? p_mul(ctx, label2, n1, B, A2, C2) : p_mul(ctx, label2, n1, A2, B, C2);
}

hipError_t compact(BatchMPContext *ctx, BatchMPArray &a, const char *label = "", uint32_t n1 = 0) {
    return stage(ctx, "compact", label, n1, a.batch_size, a.length, 0, false,
                 [&] { return batch_mp_array_compact(ctx, &a); });
}

// Compactions sharing one host round trip, for arrays whose lengths nothing
// launched in between reads (entry sizes are the first array's).
hipError_t compact_many(BatchMPContext *ctx, std::initializer_list<BatchMPArray *> arrays, const char *label,
                        uint32_t n1) {
    BatchMPArray *list[gpupi::kScalarWords];
    int n = 0;
    for (BatchMPArray *a : arrays) {
        if (n == gpupi::kScalarWords) return hipErrorInvalidValue;
        list[n++] = a;
    }
    if (n == 0) return hipSuccess;
    return stage(ctx, "compact", label, n1, list[0]->batch_size, list[0]->length, 0, false,
                 [&] { return gpupi::compact_arrays(ctx, list, n); });
}

hipError_t p_shift_addsub(BatchMPContext *ctx, const BatchMPArray &A, const BatchMPArray &B, uint32_t shift,
                          bool subtract, uint32_t n1, const BatchMPArray &C) {
    return stage(ctx, subtract ? "sub" : "add", subtract ? "R*Q<<15n-P*R" : "R*Q<<15n+P*R", n1, A.batch_size,
                 A.length, B.length, false, [&] {
                     return subtract ? batch_mp_shift_sub(ctx, A, B, shift, C) : batch_mp_shift_add(ctx, A, B, shift, C);
                 });
}

hipError_t p_exactdiv(BatchMPContext *ctx, const char *label, uint32_t n1, const BatchMPArray &A, uint32_t d) {
    return stage(ctx, "exactdiv", label, n1, A.batch_size, A.length, 0, false,
                 [&] { return batch_mp_exactdiv_small(ctx, A, d); });
}

hipError_t p_mul_small(BatchMPContext *ctx, const char *label, uint32_t n1, const BatchMPArray &A, uint32_t b,
                       const BatchMPArray &C) {
    return stage(ctx, "mul_small", label, n1, A.batch_size, A.length, 1, false,
                 [&] { return batch_mp_mul_small(ctx, A, b, C); });
}

hipError_t p_add(BatchMPContext *ctx, const char *label, uint32_t n1, const BatchMPArray &A, const BatchMPArray &B,
                 const BatchMPArray &C) {
    return stage(ctx, "add", label, n1, A.batch_size, A.length, B.length, false,
                 [&] { return batch_mp_add(ctx, A, B, C); });
}

hipError_t p_sub(BatchMPContext *ctx, const char *label, uint32_t n1, const BatchMPArray &A, const BatchMPArray &B,
                 const BatchMPArray &C) {
    return stage(ctx, "sub", label, n1, A.batch_size, A.length, B.length, false,
                 [&] { return batch_mp_sub(ctx, A, B, C); });
}

hipError_t p_add_small(BatchMPContext *ctx, const char *label, uint32_t n1, const BatchMPArray &A, uint32_t b,
                       const BatchMPArray &C) {
    return stage(ctx, "add", label, n1, A.batch_size, A.length, 1, false,
                 [&] { return batch_mp_add_small(ctx, A, b, C); });
}

hipError_t p_shift_bits(BatchMPContext *ctx, const char *label, uint32_t n1, const BatchMPArray &A,
                        const BatchMPArray &B, int32_t shift) {
    return stage(ctx, "shift", label, n1, A.batch_size, A.length, B.length, false,
                 [&] { return batch_mp_shift_bits(ctx, A, B, shift); });
}

// Host-side guard: a batch-1 view must lie inside [base, base + words).
bool inside(const BatchMPArray &v, const uint32_t *base, size_t words) {
    if (v.length == 0) return true;
    return v.data >= base && v.data + v.length <= base + words;
}

// pi_bs_gpu.cu:1068-1093
hipError_t shift_right_large(BatchMPContext *ctx, const BatchMPArray &A, uint64_t shift_bits, BatchMPArray B,
                             const char *label, uint32_t n1) {
    const uint64_t limb_shift = shift_bits / 32u;
    const uint32_t bit_shift = (uint32_t)(shift_bits % 32u);
    if (limb_shift >= A.length) return hipErrorInvalidValue;
    const BatchMPArray pre = view_of(A.data + limb_shift, A.length - (uint32_t)limb_shift, A.batch_size,
                                     A.stride - (uint32_t)limb_shift);
    return p_shift_bits(ctx, label, n1, pre, B, -(int32_t)bit_shift);
}

hipError_t launch_leaf(BatchMPContext *ctx, int is_head, uint32_t *P, uint32_t *Q, uint32_t *R, uint32_t count) {
    if (count == 0) return hipSuccess;
    uint32_t blocks = ceil_div(count, 256);
    if (blocks > 65536u * 4) blocks = 65536u * 4;
    hipLaunchKernelGGL(leaf17_kernel, dim3(blocks), dim3(256), 0, ctx->stream, is_head, P, Q, R, count);
    return hipGetLastError();
}

// --------------------------------------------- host big-integer helpers
// The inverse's base case is one division of 2^L0 (L0 <= 32768 bits) by a
// number of at most ~1024 limbs, done on the host as in the reference
// (pi_bs_gpu.cu:1237-1253, there with GMP).  Here: Knuth's algorithm D on
// 32-bit limbs.
typedef std::vector<uint32_t> Limbs;

void trim(Limbs &v) {
    while (!v.empty() && v.back() == 0) v.pop_back();
}

// q = floor(u / v), r = u mod v; v non-zero.
void divmod(const Limbs &u_in, const Limbs &v_in, Limbs &q, Limbs &r) {
    Limbs u = u_in, v = v_in;
    trim(u);
    trim(v);
    q.assign(u.size() > v.size() ? u.size() - v.size() + 1 : 1, 0);
    if (u.size() < v.size() || (u.size() == v.size() && std::lexicographical_compare(u.rbegin(), u.rend(), v.rbegin(), v.rend()))) {
        r = u;
        q.assign(1, 0);
        trim(q);
        return;
    }
    const size_t n = v.size(), m = u.size() - n;
    if (n == 1) {
        uint64_t rem = 0;
        for (size_t i = u.size(); i-- > 0;) {
            const uint64_t cur = (rem << 32) | u[i];
            q[i] = (uint32_t)(cur / v[0]);
            rem = cur % v[0];
        }
        r.assign(1, (uint32_t)rem);
        trim(q);
        trim(r);
        return;
    }
    const int s = __builtin_clz(v[n - 1]);
    Limbs vn(n), un(u.size() + 1);
    for (size_t i = n - 1; i > 0; --i) vn[i] = (v[i] << s) | (s ? (uint32_t)((uint64_t)v[i - 1] >> (32 - s)) : 0);
    vn[0] = v[0] << s;
    un[u.size()] = s ? (uint32_t)((uint64_t)u[u.size() - 1] >> (32 - s)) : 0;
    for (size_t i = u.size() - 1; i > 0; --i) un[i] = (u[i] << s) | (s ? (uint32_t)((uint64_t)u[i - 1] >> (32 - s)) : 0);
    un[0] = u[0] << s;
    q.assign(m + 1, 0);
    for (size_t j = m + 1; j-- > 0;) {
        const uint64_t num = ((uint64_t)un[j + n] << 32) | un[j + n - 1];
        uint64_t qhat = num / vn[n - 1];
        uint64_t rhat = num % vn[n - 1];
        while (qhat >= (1ull << 32) || qhat * vn[n - 2] > ((rhat << 32) | un[j + n - 2])) {
            --qhat;
            rhat += vn[n - 1];
            if (rhat >= (1ull << 32)) break;
        }
        int64_t borrow = 0;
        uint64_t carry = 0;
        for (size_t i = 0; i < n; ++i) {
            const uint64_t p = qhat * vn[i] + carry;
            carry = p >> 32;
            const int64_t t = (int64_t)un[i + j] - (int64_t)(uint32_t)p + borrow;
            un[i + j] = (uint32_t)t;
            borrow = t >> 32;
        }
        const int64_t t = (int64_t)un[j + n] - (int64_t)carry + borrow;
        un[j + n] = (uint32_t)t;
        if (t < 0) {
            --qhat;
            uint64_t c = 0;
            for (size_t i = 0; i < n; ++i) {
                const uint64_t sum = (uint64_t)un[i + j] + vn[i] + c;
                un[i + j] = (uint32_t)sum;
                c = sum >> 32;
            }
            un[j + n] += (uint32_t)c;
        }
        q[j] = (uint32_t)qhat;
    }
    r.assign(n, 0);
    for (size_t i = 0; i < n; ++i) r[i] = (un[i] >> s) | (s ? (uint32_t)((uint64_t)un[i + 1] << (32 - s)) : 0);
    trim(q);
    trim(r);
}

hipError_t download(BatchMPContext *ctx, const BatchMPArray &a, Limbs &out) {
    out.assign(a.length, 0);
    if (a.length == 0) return hipSuccess;
    GPUPI_RETURN_IF(hipMemcpyAsync(out.data(), a.data, a.length * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                   ctx->stream));
    return hipStreamSynchronize(ctx->stream);
}

// pi_bs_gpu.cu:1128-1141: write `stride` words (value zero-extended); the
// view's length becomes the used length (at least 1).
hipError_t upload(BatchMPContext *ctx, const Limbs &v, BatchMPArray &a) {
    Limbs t = v;
    trim(t);
    uint32_t used = t.empty() ? 1u : (uint32_t)t.size();
    if (used > a.stride) return hipErrorInvalidValue;
    t.resize(a.stride, 0);
    GPUPI_RETURN_IF(hipMemcpyAsync(a.data, t.data(), a.stride * sizeof(uint32_t), hipMemcpyHostToDevice,
                                   ctx->stream));
    GPUPI_RETURN_IF(hipStreamSynchronize(ctx->stream));
    a.length = used;
    return hipSuccess;
}

}  // namespace

// ====================================================================== API
namespace {

// ------------------------------------------------- the binary-split tree
struct TreeBufs {
    Buf bP0, bQ0, bR0, bP, bQ, bR, bPn, bQn, bRn, bR1, bR2;
    TreeBufs() = default;
    explicit TreeBufs(BatchMPContext *ctx) {
        for (Buf *b : {&bP0, &bQ0, &bR0, &bP, &bQ, &bR, &bPn, &bQn, &bRn, &bR1, &bR2}) b->ctx = ctx;
    }
    void freeze(bool f) {
        for (Buf *b : {&bP0, &bQ0, &bR0, &bP, &bQ, &bR, &bPn, &bQn, &bRn, &bR1, &bR2}) b->frozen = f;
    }
    void drop() {
        for (Buf *b : {&bP0, &bQ0, &bR0, &bP, &bQ, &bR, &bPn, &bQn, &bRn, &bR1, &bR2}) b->drop();
    }
};

// How a tree level fixes the lengths of its outputs.
//   kBound: no compaction, the allocation bounds stay the lengths (the
//     batched and graph forms of the group stage; nothing synchronizes);
//   kShared: the P, Q, R compactions share one host round trip, products
//     sharing an operand transform it once, rows 16-byte aligned (the
//     batched pipeline);
//   kPerOp: the reference's sequence (pi_bs_gpu.cu:637-723): P, compact P,
//     Q, compact Q, R*Q, P*R, R, compact R, each product on its own.
// All three leave the same integers.
//   kBoundShared: the batched pipeline's products and rows without the
//     compaction (lengths are the products' bounds; the tree's captured
//     levels).
enum class Lengths { kBound, kShared, kBoundShared, kPerOp };

// One pairwise level n1 of the tree (pi_bs_gpu.cu:625-723) over the rows of
// P, Q, R.
hipError_t tree_level(BatchMPContext *ctx, TreeBufs &t, uint32_t N, uint32_t n1, Lengths mode, BatchMPArray &P,
                      BatchMPArray &Q, BatchMPArray &R) {
    const uint32_t n = N * 16u;
    const bool subtract = (n1 == n) || (n1 == (n >> 4));
    const uint32_t ignore_last = (n1 <= N) ? 1u : 0u;
    const uint32_t half = n1 / 2;
    BatchMPArray Pn, Qn, Rn, R1, R2;
    const BatchMPArray Pe = view_of(P.data, P.length, half, P.stride * 2);
    const BatchMPArray Po = view_of(P.data + P.stride, P.length, half - ignore_last, P.stride * 2);
    const BatchMPArray Qe = view_of(Q.data, Q.length, half, Q.stride * 2);
    const BatchMPArray Qo = view_of(Q.data + Q.stride, Q.length, half, Q.stride * 2);
    const BatchMPArray Re = view_of(R.data, R.length, half, R.stride * 2);
    const BatchMPArray Ro = view_of(R.data + R.stride, R.length, half, R.stride * 2);
    // product and R rows 16-byte aligned in the batched pipeline (the
    // NTT's carry pass and the shift-add take four limbs per lane); the
    // other forms keep the reference's stride = length (the group stage's
    // buffers keep its preallocated sizes and may be frozen for capture)
    const bool shared = mode == Lengths::kShared || mode == Lengths::kBoundShared;
    auto out_view = [&](Buf &b, uint32_t batch, uint32_t length, BatchMPArray &v) {
        return shared ? b.view_aligned(batch, length, v) : b.view(batch, length, v);
    };
    const uint32_t terms_right = (n / n1) + ((n / n1) >> 4);
    const uint32_t shift = 15u * terms_right;
    if (mode == Lengths::kPerOp) {
        GPUPI_RETURN_IF(out_view(t.bPn, half - ignore_last, P.length * 2, Pn));
        if (Po.batch_size > 0) {
            // P_e without the last row when the last group's P is unused
            // (pi_bs_gpu.cu:645-650)
            BatchMPArray Pe_p = Pe;
            Pe_p.batch_size = Po.batch_size;
            GPUPI_RETURN_IF(p_mul(ctx, "P", n1, Pe_p, Po, Pn));
            GPUPI_RETURN_IF(compact(ctx, Pn, "P", n1));
        } else {
            Pn.length = 0;  // the top level: P is not used again
        }
        GPUPI_RETURN_IF(out_view(t.bQn, half, Q.length * 2, Qn));
        GPUPI_RETURN_IF(p_mul(ctx, "Q", n1, Qe, Qo, Qn));
        GPUPI_RETURN_IF(compact(ctx, Qn, "Q", n1));
        GPUPI_RETURN_IF(out_view(t.bR1, half, R.length + Q.length, R1));
        GPUPI_RETURN_IF(p_mul(ctx, "R*Q", n1, Re, Qo, R1));
        GPUPI_RETURN_IF(out_view(t.bR2, half, R.length + P.length, R2));
        GPUPI_RETURN_IF(p_mul(ctx, "P*R", n1, Pe, Ro, R2));
    } else {
        // P = P_e P_o (the last group's P is never used once rows are
        // whole groups) and P*R = P_e R_o share P_e's transform
        GPUPI_RETURN_IF(out_view(t.bPn, half - ignore_last, P.length * 2, Pn));
        GPUPI_RETURN_IF(out_view(t.bR2, half, R.length + P.length, R2));
        if (Po.batch_size > 0) {
            GPUPI_RETURN_IF(p_mul_pair(ctx, "P,P*R", "P", "P*R", n1, true, Po, Ro, Pe, Pn, R2));
        } else {
            GPUPI_RETURN_IF(p_mul(ctx, "P*R", n1, Pe, Ro, R2));
        }
        // Q = Q_e Q_o and R*Q = R_e Q_o share Q_o's transform
        GPUPI_RETURN_IF(out_view(t.bQn, half, Q.length * 2, Qn));
        GPUPI_RETURN_IF(out_view(t.bR1, half, R.length + Q.length, R1));
        GPUPI_RETURN_IF(p_mul_pair(ctx, "Q,R*Q", "Q", "R*Q", n1, false, Qe, Re, Qo, Qn, R1));
    }
    // R(a,c) = R(a,b) Q(b,c) 2^(15 terms(b,c)) +- P(a,b) R(b,c)
    const uint32_t shifted_len = R1.length + (shift >> 5) + ((shift & 31u) != 0);
    const uint32_t rlen = std::max(shifted_len, R2.length) + (subtract ? 0u : 1u);
    GPUPI_RETURN_IF(out_view(t.bRn, half, rlen, Rn));
    GPUPI_RETURN_IF(p_shift_addsub(ctx, R1, R2, shift, subtract, n1, Rn));
    if (mode == Lengths::kPerOp) {
        GPUPI_RETURN_IF(compact(ctx, Rn, "R", n1));
    } else if (mode == Lengths::kShared) {
        // the level's three compactions in one round trip (R is formed from
        // the products' bound lengths, as it would be after compacting P, Q)
        GPUPI_RETURN_IF(compact_many(ctx, {&Pn, &Qn, &Rn}, "P,Q,R", n1));
    }
    t.bP.swap(t.bPn);
    t.bQ.swap(t.bQn);
    t.bR.swap(t.bRn);
    P = Pn;
    Q = Qn;
    R = Rn;
    return hipSuccess;
}

// The head term 17g+16 of each group appended to its 16-term body, then the
// block's common factor divided out (pi_bs_gpu.cu:725-779).
hipError_t head_merge(BatchMPContext *ctx, TreeBufs &t, uint32_t n1, Lengths mode, const BatchMPArray &P0,
                      const BatchMPArray &Q0, const BatchMPArray &R0, BatchMPArray &P, BatchMPArray &Q,
                      BatchMPArray &R) {
    const uint32_t div_lo = 830297u, div_hi = 156590819u;  // 13^2 17^3, 7^6 11^3 (pi_bs_gpu.cu:151-152)
    const uint32_t half = P.batch_size;
    const bool compacting = mode != Lengths::kBound;
    BatchMPArray Ph, Qh, Rh, Rp1, Rp2;
    GPUPI_RETURN_IF(t.bPn.view(half, P.length + P0.length, Ph));
    GPUPI_RETURN_IF(p_mul(ctx, "P", n1, P, P0, Ph));
    if (compacting) GPUPI_RETURN_IF(compact(ctx, Ph, "P", n1));
    GPUPI_RETURN_IF(t.bQn.view(half, Q.length + Q0.length, Qh));
    GPUPI_RETURN_IF(p_mul(ctx, "Q", n1, Q, Q0, Qh));
    if (compacting) GPUPI_RETURN_IF(compact(ctx, Qh, "Q", n1));
    GPUPI_RETURN_IF(t.bR1.view(half, R.length + Q0.length, Rp1));
    GPUPI_RETURN_IF(p_mul(ctx, "R*Q", n1, R, Q0, Rp1));
    GPUPI_RETURN_IF(t.bR2.view(half, P.length + R0.length, Rp2));
    GPUPI_RETURN_IF(p_mul(ctx, "P*R", n1, P, R0, Rp2));
    const uint32_t hl = std::max(Rp1.length + 1, Rp2.length) + 1;
    GPUPI_RETURN_IF(t.bRn.view(half, hl, Rh));
    GPUPI_RETURN_IF(p_shift_addsub(ctx, Rp1, Rp2, 15u, false, n1, Rh));
    if (compacting) GPUPI_RETURN_IF(compact(ctx, Rh, "R", n1));
    GPUPI_RETURN_IF(p_exactdiv(ctx, "P", n1, Ph, div_lo));
    GPUPI_RETURN_IF(p_exactdiv(ctx, "P", n1, Ph, div_hi));
    if (compacting) GPUPI_RETURN_IF(compact(ctx, Ph, "P", n1));
    GPUPI_RETURN_IF(p_exactdiv(ctx, "Q", n1, Qh, div_lo));
    GPUPI_RETURN_IF(p_exactdiv(ctx, "Q", n1, Qh, div_hi));
    if (compacting) GPUPI_RETURN_IF(compact(ctx, Qh, "Q", n1));
    GPUPI_RETURN_IF(p_exactdiv(ctx, "R", n1, Rh, div_lo));
    GPUPI_RETURN_IF(p_exactdiv(ctx, "R", n1, Rh, div_hi));
    if (compacting) GPUPI_RETURN_IF(compact(ctx, Rh, "R", n1));
    t.bP.swap(t.bPn);
    t.bQ.swap(t.bQn);
    t.bR.swap(t.bRn);
    P = Ph;
    Q = Qh;
    R = Rh;
    return hipSuccess;
}

// The 17-term group stage (pi_bs_gpu.cu:555-779): leaf terms, the four
// pairwise levels of each 16-term body and the head merge; one row of P, Q,
// R per group.  kPerOp is the reference's form (a synchronizing compaction
// after each product it compacts); kBound keeps the allocation bounds as
// lengths (at most 51, 68 and 83 limbs at the end) and launches everything
// without a host round trip.
hipError_t leaf_stage(BatchMPContext *ctx, TreeBufs &t, uint32_t N, Lengths mode, BatchMPArray &P,
                      BatchMPArray &Q, BatchMPArray &R) {
    const uint32_t n = N * 16u;
    BatchMPArray P0, Q0, R0;
    GPUPI_RETURN_IF(t.bP0.view(N, 3, P0));
    GPUPI_RETURN_IF(t.bQ0.view(N, 4, Q0));
    GPUPI_RETURN_IF(t.bR0.view(N, 6, R0));
    GPUPI_RETURN_IF(t.bP.view(n, 3, P));
    GPUPI_RETURN_IF(t.bQ.view(n, 4, Q));
    GPUPI_RETURN_IF(t.bR.view(n, 6, R));
    // capacities as the reference pre-allocates (pi_bs_gpu.cu:611-617)
    GPUPI_RETURN_IF(t.bPn.reserve((size_t)(n / 2) * 6));
    GPUPI_RETURN_IF(t.bQn.reserve((size_t)(n / 2) * 10));
    GPUPI_RETURN_IF(t.bRn.reserve((size_t)(n / 2) * 12));
    GPUPI_RETURN_IF(t.bR1.reserve((size_t)(n / 2) * 11));
    GPUPI_RETURN_IF(t.bR2.reserve((size_t)(n / 2) * 9));
    GPUPI_RETURN_IF(stage(ctx, "leaf", "construct_leaf", N * 17u, N, P0.length, 0, false,
                          [&] { return launch_leaf(ctx, 1, P0.data, Q0.data, R0.data, N); }));
    GPUPI_RETURN_IF(stage(ctx, "leaf", "construct_leaf", N * 17u, n, P.length, 0, false,
                          [&] { return launch_leaf(ctx, 0, P.data, Q.data, R.data, n); }));
    for (uint32_t n1 = n; n1 >= 2 * N; n1 >>= 1) GPUPI_RETURN_IF(tree_level(ctx, t, N, n1, mode, P, Q, R));
    return head_merge(ctx, t, 2 * N, mode, P0, Q0, R0, P, Q, R);
}

// The same stage as one fused kernel (rows of the bound lengths 51, 68, 83,
// exactly the batched form's before compaction).
hipError_t leaf_stage_fused(BatchMPContext *ctx, TreeBufs &t, uint32_t N, BatchMPArray &P, BatchMPArray &Q,
                            BatchMPArray &R) {
    GPUPI_RETURN_IF(t.bP.view(N, 51, P));
    GPUPI_RETURN_IF(t.bQ.view(N, 68, Q));
    GPUPI_RETURN_IF(t.bR.view(N, 83, R));
    // the tree's next levels write into the partner buffers
    GPUPI_RETURN_IF(t.bPn.reserve((size_t)N * 51));
    GPUPI_RETURN_IF(t.bQn.reserve((size_t)N * 68));
    GPUPI_RETURN_IF(t.bRn.reserve((size_t)N * 83));
    return stage(ctx, "leaf", "fused_group_stage", N * 17u, N, 51, 0, false,
                 [&] { return gpupi::launch_group_stage(N, P.data, Q.data, R.data, ctx->stream); });
}

}  // namespace

// ------------------------------------------------ 17-term group stage
// Stage object for the three forms of the group stage.
struct GpupiLeafStage {
    BatchMPContext *ctx = nullptr;
    uint32_t N = 0;
    int mode = GPUPI_LEAF_PLAIN;
    TreeBufs t;
    BatchMPArray P{}, Q{}, R{};  // GRAPH: the captured outputs (bound lengths)
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint32_t *ws = nullptr;  // GRAPH: scratch the captured carry passes use
    size_t ws_bytes = 0;
};

namespace {

// Runs `fn` with the context's stream and workspace replaced by the stage's
// own, so that what a graph captures never points at memory the context may
// later reallocate.
template <typename F>
hipError_t with_stage_context(GpupiLeafStage *s, F &&fn) {
    BatchMPContext *ctx = s->ctx;
    const hipStream_t old_stream = ctx->stream;
    uint32_t *const old_ws = ctx->workspace;
    const size_t old_bytes = ctx->workspace_size_bytes;
    ctx->stream = s->stream;
    ctx->workspace = s->ws;
    ctx->workspace_size_bytes = s->ws_bytes;
    const hipError_t err = fn();
    s->ws = ctx->workspace;
    s->ws_bytes = ctx->workspace_size_bytes;
    ctx->stream = old_stream;
    ctx->workspace = old_ws;
    ctx->workspace_size_bytes = old_bytes;
    return err;
}

hipError_t leaf_stage_capture(GpupiLeafStage *s) {
    BatchMPArray P, Q, R;
    // two eager runs: every ping-pong buffer reaches the size of both roles
    for (int i = 0; i < 2; ++i) GPUPI_RETURN_IF(leaf_stage(s->ctx, s->t, s->N, Lengths::kBound, P, Q, R));
    GPUPI_RETURN_IF(hipStreamSynchronize(s->stream));
    s->t.freeze(true);
    GPUPI_RETURN_IF(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    const hipError_t body = leaf_stage(s->ctx, s->t, s->N, Lengths::kBound, P, Q, R);
    hipGraph_t graph = nullptr;
    const hipError_t end = hipStreamEndCapture(s->stream, &graph);
    s->t.freeze(false);
    s->graph = graph;
    GPUPI_RETURN_IF(body);
    GPUPI_RETURN_IF(end);
    GPUPI_RETURN_IF(hipGraphInstantiate(&s->exec, graph, nullptr, nullptr, 0));
    s->P = P;
    s->Q = Q;
    s->R = R;
    return hipStreamSynchronize(s->stream);
}

}  // namespace

// ------------------------------------------- tree levels as a hipGraph
// The bottom levels of the binary-splitting tree (rows up to one-tile
// products: schoolbook and single-tile NTT, where launch latency and the
// per-level compaction round trip cost as much as the work) run with
// lengths fixed on the host and no compaction (the values are the same
// integers: compaction only drops zero top limbs), captured once as a
// hipGraph on a private stream and replayed.  The graph belongs to the
// context and is keyed by the group count, the compacted lengths of the
// group stage's rows and the schoolbook/NTT crossover; its buffers and
// workspace are its own (hipMalloc'd, never from the block cache), so
// nothing it points at moves between replays.
struct TreeGraph {
    uint32_t N = 0, LP = 0, LQ = 0, LR = 0, levels = 0;
    int ntt_min = 0;
    Buf in_P, in_Q, in_R;  // the group stage's rows: the graph's input
    TreeBufs t;            // the captured levels' buffers
    uint32_t *ws = nullptr;
    size_t ws_bytes = 0;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    BatchMPArray P{}, Q{}, R{};  // outputs of the last captured level
    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        t.freeze(false);
    }
    ~TreeGraph() {
        drop_graph();
        if (ws) (void)hipFree(ws);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

void tree_graph_free(TreeGraph *g) { delete g; }

namespace {

// Largest L_a + L_b a captured level may multiply: one 4096-element tile.
constexpr uint32_t kGraphMaxProduct = 4096;

// Levels (from n1 = N down) whose products all fit one tile when lengths
// grow by the rules of tree_level without compaction, starting from the
// compacted group-stage lengths; never the top level (n1 = 2).
uint32_t graph_levels(uint32_t N, uint32_t LP, uint32_t LQ, uint32_t LR) {
    const uint32_t n = N * 16u;
    uint32_t levels = 0;
    for (uint32_t n1 = N; n1 > 2; n1 >>= 1) {
        const uint64_t pp = 2ull * LP, qq = 2ull * LQ, rq = (uint64_t)LR + LQ, pr = (uint64_t)LP + LR;
        if (std::max(std::max(pp, qq), std::max(rq, pr)) > kGraphMaxProduct) break;
        const bool subtract = (n1 == n) || (n1 == (n >> 4));
        const uint32_t shift = 15u * ((n / n1) + ((n / n1) >> 4));
        const uint64_t shifted = rq + (shift >> 5) + ((shift & 31u) != 0);
        LR = (uint32_t)(std::max<uint64_t>(shifted, pr) + (subtract ? 0u : 1u));
        LP = (uint32_t)pp;
        LQ = (uint32_t)qq;
        ++levels;
    }
    return levels;
}

// Runs `fn` on the graph's stream with its workspace as the context's.
template <typename F>
hipError_t with_graph_context(BatchMPContext *ctx, TreeGraph *g, F &&fn) {
    const hipStream_t old_stream = ctx->stream;
    uint32_t *const old_ws = ctx->workspace;
    const size_t old_bytes = ctx->workspace_size_bytes;
    ctx->stream = g->stream;
    ctx->workspace = g->ws;
    ctx->workspace_size_bytes = g->ws_bytes;
    const hipError_t err = fn();
    g->ws = ctx->workspace;
    g->ws_bytes = ctx->workspace_size_bytes;
    ctx->stream = old_stream;
    ctx->workspace = old_ws;
    ctx->workspace_size_bytes = old_bytes;
    return err;
}

// The captured levels over the graph's input rows (views P, Q, R of the
// compacted group stage); leaves the last level's rows in P, Q, R.
hipError_t graph_body(BatchMPContext *ctx, TreeGraph *g, BatchMPArray &P, BatchMPArray &Q, BatchMPArray &R) {
    uint32_t n1 = g->N;
    for (uint32_t l = 0; l < g->levels; ++l, n1 >>= 1)
        GPUPI_RETURN_IF(tree_level(ctx, g->t, g->N, n1, Lengths::kBoundShared, P, Q, R));
    return hipSuccess;
}

// Captures the levels of g (two eager runs first: every ping-pong buffer
// and the workspace reach the sizes of both roles), then instantiates.
hipError_t graph_capture(BatchMPContext *ctx, TreeGraph *g, const BatchMPArray &P0, const BatchMPArray &Q0,
                         const BatchMPArray &R0) {
    if (g->stream == nullptr) GPUPI_RETURN_IF(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
    // the captured work follows the context stream's earlier work (the
    // group stage that wrote the input rows)
    hipEvent_t ready = nullptr;
    GPUPI_RETURN_IF(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
    hipError_t err = hipEventRecord(ready, ctx->stream);
    if (err == hipSuccess) err = hipStreamWaitEvent(g->stream, ready, 0);
    (void)hipEventDestroy(ready);
    GPUPI_RETURN_IF(err);
    return with_graph_context(ctx, g, [&]() -> hipError_t {
        BatchMPArray P, Q, R;
        for (int i = 0; i < 2; ++i) {
            P = P0;
            Q = Q0;
            R = R0;
            GPUPI_RETURN_IF(graph_body(ctx, g, P, Q, R));
        }
        GPUPI_RETURN_IF(hipStreamSynchronize(g->stream));
        g->t.freeze(true);
        gpupi::ntt_stats_suspend(true);
        P = P0;
        Q = Q0;
        R = R0;
        hipError_t body = hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal);
        if (body == hipSuccess) {
            body = graph_body(ctx, g, P, Q, R);
            hipGraph_t graph = nullptr;
            const hipError_t end = hipStreamEndCapture(g->stream, &graph);
            g->graph = graph;
            if (body == hipSuccess) body = end;
        }
        gpupi::ntt_stats_suspend(false);
        if (body != hipSuccess) {
            g->drop_graph();
            return body;
        }
        GPUPI_RETURN_IF(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
        g->P = P;
        g->Q = Q;
        g->R = R;
        return hipStreamSynchronize(g->stream);
    });
}

}  // namespace

extern "C" {

hipError_t gpupi_construct_leaf_17(BatchMPContext *ctx, int is_head, uint32_t *P, uint32_t *Q, uint32_t *R,
                                   uint32_t count) {
    if (ctx == nullptr || (count > 0 && (!P || !Q || !R))) return hipErrorInvalidValue;
    return launch_leaf(ctx, is_head, P, Q, R, count);
}

// pi_bs_gpu.cu:503-792
hipError_t gpupi_binary_split(BatchMPContext *ctx, uint32_t N, BatchMPArray *Q_out, BatchMPArray *R_out) {
    if (ctx == nullptr || Q_out == nullptr || R_out == nullptr || N == 0 || (N & (N - 1)) != 0) return hipErrorInvalidValue;
    TreeBufs t(ctx);
    BatchMPArray P, Q, R;
    if (ctx->pipeline_form == GPUPI_FORM_PLAIN) {
        // the reference's form: the group stage op by op, every level's
        // compactions one by one
        GPUPI_RETURN_IF(leaf_stage(ctx, t, N, Lengths::kPerOp, P, Q, R));
        for (uint32_t n1 = N; n1 > 1; n1 >>= 1) GPUPI_RETURN_IF(tree_level(ctx, t, N, n1, Lengths::kPerOp, P, Q, R));
    } else if (ctx->cache->pool() != nullptr || (ctx->profiler != nullptr && ctx->profiler->enabled)) {
        // the batched form, every level launched and compacted in turn (a
        // context on a memory pool keeps no buffers between runs; the stage
        // profiler times each operation)
        GPUPI_RETURN_IF(leaf_stage_fused(ctx, t, N, P, Q, R));
        GPUPI_RETURN_IF(compact_many(ctx, {&P, &Q, &R}, "P,Q,R", 2 * N));
        for (uint32_t n1 = N; n1 > 1; n1 >>= 1) GPUPI_RETURN_IF(tree_level(ctx, t, N, n1, Lengths::kShared, P, Q, R));
    } else {
        // the 17-term group stage with bound lengths (no synchronization
        // inside) into the tree graph's input rows, one compaction of each
        // array, the bottom levels replayed from the graph, one compaction,
        // the remaining levels compacted level by level
        if (ctx->tree_graph == nullptr) {
            ctx->tree_graph = new (std::nothrow) TreeGraph();
            if (ctx->tree_graph == nullptr) return hipErrorOutOfMemory;
        }
        TreeGraph *g = ctx->tree_graph;
        if (g->N != N) {
            g->drop_graph();
            g->N = N;
        }
        GPUPI_RETURN_IF(g->in_P.view(N, 51, P));
        GPUPI_RETURN_IF(g->in_Q.view(N, 68, Q));
        GPUPI_RETURN_IF(g->in_R.view(N, 83, R));
        GPUPI_RETURN_IF(stage(ctx, "leaf", "fused_group_stage", N * 17u, N, 51, 0, false,
                              [&] { return gpupi::launch_group_stage(N, P.data, Q.data, R.data, ctx->stream); }));
        GPUPI_RETURN_IF(compact_many(ctx, {&P, &Q, &R}, "P,Q,R", 2 * N));
        const uint32_t levels = graph_levels(N, P.length, Q.length, R.length);
        uint32_t n1 = N;
        if (levels > 0) {
            if (g->exec == nullptr || g->LP != P.length || g->LQ != Q.length || g->LR != R.length ||
                g->levels != levels || g->ntt_min != ctx->ntt_min_limbs) {
                g->drop_graph();
                g->LP = P.length;
                g->LQ = Q.length;
                g->LR = R.length;
                g->levels = levels;
                g->ntt_min = ctx->ntt_min_limbs;
                GPUPI_RETURN_IF(graph_capture(ctx, g, P, Q, R));
            }
            GPUPI_RETURN_IF(hipGraphLaunch(g->exec, ctx->stream));
            P = g->P;
            Q = g->Q;
            R = g->R;
            GPUPI_RETURN_IF(compact_many(ctx, {&P, &Q, &R}, "P,Q,R", n1 >> (levels - 1)));
            n1 >>= levels;
        }
        for (; n1 > 1; n1 >>= 1) GPUPI_RETURN_IF(tree_level(ctx, t, N, n1, Lengths::kShared, P, Q, R));
        if (t.bQ.ptr == nullptr) {
            // no level ran on this call's buffers (a single group): the rows
            // are the graph's input rows; copy them out for the caller
            BatchMPArray Qc, Rc;
            GPUPI_RETURN_IF(t.bQ.view(1, Q.length, Qc));
            GPUPI_RETURN_IF(t.bR.view(1, R.length, Rc));
            GPUPI_RETURN_IF(hipMemcpyAsync(Qc.data, Q.data, (size_t)Q.length * 4, hipMemcpyDeviceToDevice, ctx->stream));
            GPUPI_RETURN_IF(hipMemcpyAsync(Rc.data, R.data, (size_t)R.length * 4, hipMemcpyDeviceToDevice, ctx->stream));
            Q = Qc;
            R = Rc;
        }
    }
    // Q and R now hold one row each; hand their buffers to the caller.
    *Q_out = view_of(t.bQ.release(), Q.length, 1, Q.stride);
    *R_out = view_of(t.bR.release(), R.length, 1, R.stride);
    return hipSuccess;
}

// Limbs that hold p_j (or q_j) after j Pell doublings from (4001, 40):
// p_j + q_j sqrt(10005) = x^(2^j) with x = 4001 + 40 sqrt(10005) and
// p_j^2 - 10005 q_j^2 = 1, so p_j < x^(2^j) and q_j < x^(2^j) / sqrt(10005);
// one limb of margin covers the rounding of the double-precision logarithm.
static uint32_t sqrt_limbs_bound(uint32_t doublings, bool q) {
    const double log2x = std::log2(4001.0 + 40.0 * std::sqrt(10005.0));
    double bits = std::ldexp(log2x, (int)doublings);
    if (q) bits -= 0.5 * std::log2(10005.0);
    return (uint32_t)(std::ceil(bits / 32.0)) + 1u;
}

// pi_bs_gpu.cu:1346-1522
hipError_t gpupi_sqrt_10005(BatchMPContext *ctx, int target_digits, BatchMPArray *P_out, BatchMPArray *Q_out) {
    if (ctx == nullptr || P_out == nullptr || Q_out == nullptr) return hipErrorInvalidValue;
    Buf bP(ctx), bQ(ctx), bPQ(ctx), bPP(ctx), bQQ(ctx), bQQ5(ctx), bPn(ctx), bQn(ctx);
    BatchMPArray P, Q;
    GPUPI_RETURN_IF(bP.view(1, 1, P));
    GPUPI_RETURN_IF(bQ.view(1, 1, Q));
    const uint32_t init[2] = {4001u, 40u};
    GPUPI_RETURN_IF(hipMemcpyAsync(P.data, &init[0], 4, hipMemcpyHostToDevice, ctx->stream));
    GPUPI_RETURN_IF(hipMemcpyAsync(Q.data, &init[1], 4, hipMemcpyHostToDevice, ctx->stream));
    GPUPI_RETURN_IF(hipStreamSynchronize(ctx->stream));
    const bool plain = ctx->pipeline_form == GPUPI_FORM_PLAIN;
    double valid_digits = 7.5;
    uint32_t doublings = 0;
    while (valid_digits < target_digits) {
        valid_digits *= 2;
        ++doublings;
        const uint32_t sd = (uint32_t)valid_digits;
        BatchMPArray PQ, Qn, PP, QQ, QQ5, Pn;
        // batched form: every result keeps a length known on the host, so
        // nothing synchronizes: products and sums their allocation lengths,
        // the new P and Q the bound sqrt_limbs_bound (the values are the same
        // integers; the top limbs may be zero).  Plain form: every result
        // compacted as the reference does (pi_bs_gpu.cu:1443-1515).
        GPUPI_RETURN_IF(bPQ.view(1, P.length + Q.length, PQ));
        GPUPI_RETURN_IF(p_mul(ctx, "P*Q", sd, P, Q, PQ));
        if (plain) GPUPI_RETURN_IF(compact(ctx, PQ, "P*Q", sd));
        GPUPI_RETURN_IF(bQn.view(1, PQ.length + 1, Qn));
        GPUPI_RETURN_IF(p_shift_bits(ctx, "2*P*Q", sd, PQ, Qn, 1));
        if (plain) GPUPI_RETURN_IF(compact(ctx, Qn, "Q", sd));
        GPUPI_RETURN_IF(bPP.view(1, P.length * 2, PP));
        GPUPI_RETURN_IF(p_mul(ctx, "P*P", sd, P, P, PP));
        if (plain) GPUPI_RETURN_IF(compact(ctx, PP, "P*P", sd));
        GPUPI_RETURN_IF(bQQ.view(1, Q.length * 2, QQ));
        GPUPI_RETURN_IF(p_mul(ctx, "Q*Q", sd, Q, Q, QQ));
        if (plain) GPUPI_RETURN_IF(compact(ctx, QQ, "Q*Q", sd));
        GPUPI_RETURN_IF(bQQ5.view(1, QQ.length + 1, QQ5));
        GPUPI_RETURN_IF(p_mul_small(ctx, "10005*Q*Q", sd, QQ, 10005u, QQ5));
        if (plain) GPUPI_RETURN_IF(compact(ctx, QQ5, "10005*Q*Q", sd));
        GPUPI_RETURN_IF(bPn.view(1, std::max(PP.length, QQ5.length) + 1, Pn));
        GPUPI_RETURN_IF(p_add(ctx, "P*P+10005*Q*Q", sd, PP, QQ5, Pn));
        if (plain) {
            GPUPI_RETURN_IF(compact(ctx, Pn, "P", sd));
        } else {
            Pn.length = std::min(Pn.length, sqrt_limbs_bound(doublings, false));
            Qn.length = std::min(Qn.length, sqrt_limbs_bound(doublings, true));
        }
        bP.swap(bPn);
        bQ.swap(bQn);
        P = Pn;
        Q = Qn;
    }
    // the reference's results are compacted (one round trip here for both)
    if (!plain) GPUPI_RETURN_IF(compact_many(ctx, {&P, &Q}, "P,Q", (uint32_t)valid_digits));
    *P_out = view_of(bP.release(), P.length, 1, P.stride);
    *Q_out = view_of(bQ.release(), Q.length, 1, Q.stride);
    return hipSuccess;
}

// pi_bs_gpu.cu:1143-1344: y ~ 2^L / x by precision doubling from a host base
// case.  One flat workspace holds res | res2 | y2 exactly as the reference
// lays it out, so `res << m` is a view that extends downwards.
hipError_t gpupi_recursive_inverse(BatchMPContext *ctx, BatchMPArray x, uint64_t L, BatchMPArray *y_out) {
    if (ctx == nullptr || y_out == nullptr || x.batch_size != 1 || x.data == nullptr) return hipErrorInvalidValue;
    const uint64_t cpu_limit = 32u * 1024u;
    std::vector<uint64_t> levels{L};
    while (levels.back() > cpu_limit) {
        const uint64_t m = ((levels.back() - 2u) / 4u) & ~31ull;
        levels.push_back(levels.back() - m * 2u);
    }
    const size_t L_res = (size_t)((L + 31) / 32 + 2), L_res2 = (size_t)((L + 31) / 32 + 1),
                 L_y2 = (size_t)x.length + 1;
    Buf by(ctx), bws(ctx);
    BatchMPArray y;
    GPUPI_RETURN_IF(by.view(1, x.length, y));
    GPUPI_RETURN_IF(gpupi::launch_fill_u32(y.data, x.length, 0u, ctx->stream));
    GPUPI_RETURN_IF(bws.reserve(L_res + L_res2 + L_y2));
    uint32_t *res_base = bws.ptr, *res2_base = bws.ptr + L_res, *y2_base = res2_base + L_res2;
    GPUPI_RETURN_IF(gpupi::launch_fill_u32(res_base, L_res, 0u, ctx->stream));
    GPUPI_RETURN_IF(gpupi::launch_fill_u32(res2_base, L_res2 + L_y2, 0xA5A5A5A5u, ctx->stream));

    BatchMPArray res;
    {
        const uint64_t L0 = levels.back();
        const uint32_t start = (uint32_t)((L - L0) >> 6);
        const BatchMPArray x0 = make_subview(x, start, x.length - start);
        BatchMPArray y0 = make_subview(y, start, y.length - start);
        res = view_of(res_base + (size_t)start * 2, x0.length + 1, 1, x0.length + 1);
        if (!inside(res, bws.ptr, bws.cap)) return hipErrorInvalidValue;
        Limbs xv, num((size_t)(L0 / 32) + 1, 0), qv, rv;
        GPUPI_RETURN_IF(download(ctx, x0, xv));
        num[L0 / 32] = 1u << (L0 % 32);
        trim(xv);
        if (xv.empty()) return hipErrorInvalidValue;
        divmod(num, xv, qv, rv);
        GPUPI_RETURN_IF(upload(ctx, qv, y0));
        GPUPI_RETURN_IF(upload(ctx, rv, res));
    }
    GPUPI_RETURN_IF(compact(ctx, res, "res", (uint32_t)(levels.back() >> 5)));

    for (size_t i = levels.size() - 1; i > 0; --i) {
        const uint64_t L1 = levels[i], L2 = levels[i - 1];
        const uint64_t m = (L2 - L1) >> 1;
        const uint32_t mw = (uint32_t)(m >> 5);
        const BatchMPArray x2 = make_subview(x, (uint32_t)((L - L2) >> 6), mw);
        const BatchMPArray y1 = make_subview(y, (uint32_t)((L - L1) >> 6), y.length - (uint32_t)((L - L1) >> 6));
        const BatchMPArray x2y1 = view_of(res2_base, x2.length + y1.length, 1, x2.length + y1.length);
        BatchMPArray res1_m = view_of(res.data - mw, res.length + mw + 1, 1, res.length + mw + 1);
        if (!inside(x2y1, bws.ptr, bws.cap) || !inside(res1_m, bws.ptr, bws.cap)) return hipErrorInvalidValue;
        const uint32_t d5 = (uint32_t)(L2 >> 5);
        GPUPI_RETURN_IF(p_mul(ctx, "x2*y", d5, x2, y1, x2y1));
        GPUPI_RETURN_IF(p_sub(ctx, "res1_m", d5, res1_m, x2y1, res1_m));
        uint32_t top = 0;
        GPUPI_RETURN_IF(hipMemcpyAsync(ctx->h_scalar, res1_m.data + res1_m.length - 1, 4, hipMemcpyDeviceToHost,
                                       ctx->stream));
        GPUPI_RETURN_IF(hipStreamSynchronize(ctx->stream));
        top = *ctx->h_scalar;
        const bool negative = top >= 0x80000000u;
        if (negative) {
            GPUPI_RETURN_IF(p_sub(ctx, "res1_m_neg", d5, view_of(res1_m.data, 0, 1, 0), res1_m, res1_m));
        }
        const BatchMPArray res2 = view_of(res2_base, res1_m.length + y1.length, 1, res1_m.length + y1.length);
        GPUPI_RETURN_IF(p_mul(ctx, "res2*y", d5, res1_m, y1, res2));
        const uint32_t drop = (uint32_t)((L2 - m * 2) >> 5);
        const BatchMPArray y2 = view_of(y2_base, res2.length - drop, 1, res2.length - drop);
        if (res2.length < drop || !inside(res2, bws.ptr, bws.cap) || !inside(y2, bws.ptr, bws.cap)) return hipErrorInvalidValue;
        GPUPI_RETURN_IF(shift_right_large(ctx, res2, L2 - m * 2u, y2, "y2", d5));
        if (negative) GPUPI_RETURN_IF(p_add_small(ctx, "y2_inc", d5, y2, 1u, y2));
        const BatchMPArray y_new = make_subview(y, (uint32_t)((L - L2) >> 6), y.length - (uint32_t)((L - L2) >> 6));
        if (!negative) GPUPI_RETURN_IF(p_add(ctx, "(y<<m)+y2", d5, y_new, y2, y_new));
        else GPUPI_RETURN_IF(p_sub(ctx, "(y<<m)-y2", d5, y_new, y2, y_new));
        if (L2 != L) {
            const BatchMPArray x_next = make_subview(x, (uint32_t)((L - L2) >> 6), x.length - (uint32_t)((L - L2) >> 6));
            res = view_of(res1_m.data - mw, res1_m.length + mw, 1, res1_m.length + mw);
            const BatchMPArray y2x = view_of(res2_base, y2.length + x_next.length, 1, y2.length + x_next.length);
            if (!inside(res, bws.ptr, bws.cap) || !inside(y2x, bws.ptr, bws.cap)) return hipErrorInvalidValue;
            GPUPI_RETURN_IF(p_mul(ctx, "y2*x", d5, y2, x_next, y2x));
            if (!negative) GPUPI_RETURN_IF(p_sub(ctx, "(res<<m)-y2x", d5, res, y2x, res));
            else GPUPI_RETURN_IF(p_sub(ctx, "y2x-(res<<m)", d5, y2x, res, res));
            GPUPI_RETURN_IF(compact(ctx, res, "res", d5));
        }
    }
    GPUPI_RETURN_IF(hipStreamSynchronize(ctx->stream));
    *y_out = view_of(by.release(), y.length, 1, y.stride);
    return hipSuccess;
}

}  // extern "C"

namespace {

// Groups of the binary-splitting tree for target_digits (pi_bs_gpu.cu:1537-1540:
// the smallest power of two with 17 N 14.3 >= target_digits).
uint32_t pi_groups(int target_digits) {
    uint32_t N = 1;
    while (N * 17 * 14.3 < target_digits) N *= 2;
    return N;
}

// pi_fractional after the binary splitting (pi_bs_gpu.cu:1557-1685): Q_bs,
// R_bs are the root rows of the tree over N groups, handed over (blocks of
// ctx's cache, or of its memory pool).
hipError_t pi_fractional_from_tree(BatchMPContext *ctx, int target_digits, int N, BatchMPArray Q_bs,
                                   BatchMPArray R_bs, BatchMPArray *ret_out, uint32_t *groups_out,
                                   uint64_t *inverse_bits_out) {
    const int prec_limbs = (int)(((long)(target_digits * 3.322) + 31) / 32 + 1);
    Buf bQbs(ctx), bRbs(ctx);
    bQbs.adopt(Q_bs.data, Q_bs.stride);
    bRbs.adopt(R_bs.data, R_bs.stride);
    const unsigned long long r_total_shift = 15ull * (unsigned long long)N * 17ull - 8ull;
    const int rs_bits = (int)(r_total_shift % 32), rs_limbs = (int)(r_total_shift / 32);
    if ((uint32_t)rs_limbs >= R_bs.length) return hipErrorInvalidValue;
    const BatchMPArray R_pre = view_of(R_bs.data + rs_limbs, R_bs.length - (uint32_t)rs_limbs, 1,
                                       R_bs.length - (uint32_t)rs_limbs);
    Buf bRsh(ctx);
    BatchMPArray R_sh;
    GPUPI_RETURN_IF(bRsh.view(1, R_pre.length + 1, R_sh));
    const uint32_t td = (uint32_t)target_digits;
    GPUPI_RETURN_IF(p_shift_bits(ctx, "R_shift", td, R_pre, R_sh, -rs_bits));
    GPUPI_RETURN_IF(compact(ctx, R_sh, "R", td));

    long extra = std::max(0l, (long)std::min(Q_bs.length, R_sh.length) - prec_limbs);
    const BatchMPArray Q_tr = view_of(Q_bs.data + extra, Q_bs.length - (uint32_t)extra, 1, Q_bs.length - (uint32_t)extra);
    const BatchMPArray R_tr = view_of(R_sh.data + extra, R_sh.length - (uint32_t)extra, 1, R_sh.length - (uint32_t)extra);

    BatchMPArray P15{}, Q15{};
    GPUPI_RETURN_IF(gpupi_sqrt_10005(ctx, target_digits, &P15, &Q15));
    Buf bP15(ctx), bQ15(ctx);
    bP15.adopt(P15.data, P15.stride);
    bQ15.adopt(Q15.data, Q15.stride);
    Buf bPf(ctx), bQf(ctx);
    BatchMPArray P_final, Q_final;
    GPUPI_RETURN_IF(bPf.view(1, P15.length + Q_tr.length, P_final));
    GPUPI_RETURN_IF(bQf.view(1, Q15.length + R_tr.length, Q_final));
    GPUPI_RETURN_IF(p_mul(ctx, "P15*Q", td, P15, Q_tr, P_final));
    GPUPI_RETURN_IF(p_mul(ctx, "R*Q15", td, R_tr, Q15, Q_final));
    GPUPI_RETURN_IF(compact(ctx, P_final, "P", td));
    GPUPI_RETURN_IF(compact(ctx, Q_final, "Q", td));

    extra = std::max(0l, (long)std::min(P_final.length, Q_final.length) - prec_limbs);
    const BatchMPArray P_tf = make_subview(P_final, (uint32_t)extra, P_final.length - (uint32_t)extra);
    BatchMPArray Q_tf = make_subview(Q_final, (uint32_t)extra, Q_final.length - (uint32_t)extra);
    GPUPI_RETURN_IF(compact(ctx, Q_tf, "Q", td));
    uint32_t q_bits = 0;
    if (Q_tf.length > 0) {
        GPUPI_RETURN_IF(hipMemcpyAsync(ctx->h_scalar, Q_tf.data + Q_tf.length - 1, 4, hipMemcpyDeviceToHost, ctx->stream));
        GPUPI_RETURN_IF(hipStreamSynchronize(ctx->stream));
        q_bits = (Q_tf.length - 1) * 32u + (32u - (uint32_t)__builtin_clz(*ctx->h_scalar));
    }
    const uint64_t inverse_bits = (uint64_t)q_bits * 2u - 1u;

    BatchMPArray inv{};
    GPUPI_RETURN_IF(gpupi_recursive_inverse(ctx, Q_tf, inverse_bits, &inv));
    Buf binv(ctx);
    binv.adopt(inv.data, inv.stride);
    Buf bprod(ctx);
    BatchMPArray prod;
    GPUPI_RETURN_IF(bprod.view(1, P_tf.length + inv.length, prod));
    GPUPI_RETURN_IF(p_mul(ctx, "P*invQ", td, P_tf, inv, prod));
    GPUPI_RETURN_IF(compact(ctx, prod, "P*invQ", td));

    Buf bret(ctx);
    BatchMPArray ret;
    GPUPI_RETURN_IF(bret.view(1, prod.length - (uint32_t)(inverse_bits / 32u) + (uint32_t)prec_limbs, ret));
    const uint64_t final_shift = inverse_bits - (uint64_t)prec_limbs * 32u;
    GPUPI_RETURN_IF(shift_right_large(ctx, prod, final_shift, ret, "RET_shift", td));
    GPUPI_RETURN_IF(compact(ctx, ret, "RET", td));
    ret.data = bret.release();
    *ret_out = ret;
    if (groups_out) *groups_out = (uint32_t)N;
    if (inverse_bits_out) *inverse_bits_out = inverse_bits;
    return hipSuccess;
}

}  // namespace

extern "C" {

// pi_bs_gpu.cu:1524-1685
hipError_t gpupi_pi_fractional(BatchMPContext *ctx, int target_digits, BatchMPArray *ret_out, uint32_t *groups_out,
                               uint64_t *inverse_bits_out) {
    if (ctx == nullptr || ret_out == nullptr) return hipErrorInvalidValue;
    const uint32_t N = pi_groups(target_digits);
    BatchMPArray Q_bs{}, R_bs{};
    GPUPI_RETURN_IF(gpupi_binary_split(ctx, N, &Q_bs, &R_bs));
    return pi_fractional_from_tree(ctx, target_digits, (int)N, Q_bs, R_bs, ret_out, groups_out, inverse_bits_out);
}

// pi_bs_gpu.cu:924-939
uint32_t gpupi_round_up_decimal_digits(uint32_t requested, uint32_t *leaf_digits) {
    if (requested == 0) {
        if (leaf_digits) *leaf_digits = 477u;
        return 0;
    }
    uint32_t digits = requested, levels = 0;
    while (digits >= 32u * 18u) {
        digits >>= 1;
        ++levels;
    }
    if ((digits << levels) < requested) ++digits;
    if (leaf_digits) *leaf_digits = digits;
    return digits << levels;
}

// pi_bs_gpu.cu:1537-1540
uint32_t gpupi_pi_groups_for_digits(uint32_t prec_digits) { return pi_groups((int)prec_digits); }

// pi_bs_gpu.cu:941-944
uint32_t gpupi_prec_limbs_for_digits(uint32_t digits) {
    const uint32_t prec_bits = (uint32_t)((double)digits * 3.322) + 31u;
    return prec_bits / 32u + 1u;
}

hipError_t gpupi_device_info(int *compute_units, int *clock_khz, char *arch, int arch_len) {
    if (compute_units == nullptr || clock_khz == nullptr) return hipErrorInvalidValue;
    int dev = 0;
    GPUPI_RETURN_IF(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    GPUPI_RETURN_IF(hipGetDeviceProperties(&prop, dev));
    *compute_units = prop.multiProcessorCount;
    *clock_khz = prop.clockRate;
    if (arch != nullptr && arch_len > 0) std::snprintf(arch, (size_t)arch_len, "%s", prop.gcnArchName);
    return hipSuccess;
}

}  // extern "C"

// ========================================================= radix conversion
namespace {

uint32_t bits_for_digits(uint32_t digits) {  // convert_decimal.cu:36-39
    return (uint32_t)(std::ceil((double)digits * 3.32192809488736234787) + 34.0);
}

uint32_t levels_for_shape(uint32_t total, uint32_t leaf) {  // convert_decimal.cu:41-52
    if (leaf == 0 || total == 0 || total % leaf) return UINT32_MAX;
    uint32_t ratio = total / leaf, lv = 0;
    while (ratio > 1 && (ratio & 1u) == 0) {
        ratio >>= 1;
        ++lv;
    }
    return ratio == 1 ? lv : UINT32_MAX;
}

}  // namespace

extern "C" {

}  // extern "C"

namespace {

// 5^leaf_digits and its repeated squares: powers[l] = 5^(leaf_digits 2^l)
// for l < levels, row l in the block of bufs[l] (convert_decimal.cu:198-269).
// With host_lengths, each power keeps a length fixed on the host (at least
// its limbs: 5^e has floor(e log2 5) + 1 bits, one limb of margin for the
// rounding of the logarithm) instead of a synchronizing compaction; the
// conversion reads bits at positions fixed by the digit counts, so a power's
// zero top limbs change nothing.
hipError_t fill_powers(BatchMPContext *ctx, uint32_t leaf_digits, uint32_t levels, BatchMPArray *powers, Buf *bufs,
                       bool host_lengths) {
    // 5^leaf_digits on the host (at most ~1400 bits), then squarings.
    Limbs five(1, 1);
    for (uint32_t k = 0; k < leaf_digits; ++k) {
        uint64_t carry = 0;
        for (auto &w : five) {
            const uint64_t t = (uint64_t)w * 5u + carry;
            w = (uint32_t)t;
            carry = t >> 32;
        }
        if (carry) five.push_back((uint32_t)carry);
    }
    GPUPI_RETURN_IF(bufs[0].view(1, (uint32_t)five.size(), powers[0]));
    GPUPI_RETURN_IF(hipMemcpyAsync(powers[0].data, five.data(), five.size() * 4, hipMemcpyHostToDevice, ctx->stream));
    GPUPI_RETURN_IF(hipStreamSynchronize(ctx->stream));
    for (uint32_t l = 1; l < levels; ++l) {
        GPUPI_RETURN_IF(bufs[l].view(1, powers[l - 1].length * 2, powers[l]));
        GPUPI_RETURN_IF(batch_mp_mul(ctx, powers[l - 1], powers[l - 1], powers[l]));
        if (host_lengths) {
            const double bits = std::ldexp((double)leaf_digits * std::log2(5.0), (int)l) + 1.0;
            powers[l].length = std::min(powers[l].length, (uint32_t)std::ceil(bits / 32.0) + 1u);
        } else {
            GPUPI_RETURN_IF(batch_mp_array_compact(ctx, &powers[l]));
        }
    }
    return hipSuccess;
}

// The power table of one pi run, its blocks from the context's cache.
struct CachedPowers {
    std::unique_ptr<Buf[]> bufs;
    std::vector<BatchMPArray> arrays;
    ConvertDecimalPowerTable table{};
    hipError_t fill(BatchMPContext *ctx, uint32_t leaf_digits, uint32_t levels) {
        table.leaf_digits = leaf_digits;
        if (levels == 0) return hipSuccess;
        bufs.reset(new (std::nothrow) Buf[levels]);
        if (!bufs) return hipErrorOutOfMemory;
        for (uint32_t l = 0; l < levels; ++l) bufs[l].ctx = ctx;
        arrays.assign(levels, BatchMPArray{});
        GPUPI_RETURN_IF(fill_powers(ctx, leaf_digits, levels, arrays.data(), bufs.get(),
                                    ctx->pipeline_form != GPUPI_FORM_PLAIN));
        table.powers = arrays.data();
        table.levels = levels;
        return hipSuccess;
    }
};

}  // namespace

extern "C" {

// convert_decimal.cu:198-269
hipError_t convert_decimal_precompute_powers(BatchMPContext *ctx, ConvertDecimalPowerTable *table,
                                             uint32_t leaf_digits, uint32_t levels) {
    if (ctx == nullptr || table == nullptr || leaf_digits == 0) return hipErrorInvalidValue;
    table->powers = nullptr;
    table->levels = 0;
    table->leaf_digits = leaf_digits;
    if (levels == 0) return hipSuccess;
    // the caller releases each row with batch_mp_array_release: plain blocks
    std::unique_ptr<Buf[]> bufs(new (std::nothrow) Buf[levels]);
    BatchMPArray *powers = new (std::nothrow) BatchMPArray[levels];
    if (!bufs || powers == nullptr) {
        delete[] powers;
        return hipErrorOutOfMemory;
    }
    for (uint32_t i = 0; i < levels; ++i) powers[i] = BatchMPArray{};
    const hipError_t err = fill_powers(ctx, leaf_digits, levels, powers, bufs.get(), false);
    if (err != hipSuccess) {
        delete[] powers;
        return err;
    }
    for (uint32_t i = 0; i < levels; ++i) powers[i].data = bufs[i].release();
    table->powers = powers;
    table->levels = levels;
    return hipSuccess;
}

void convert_decimal_release_powers(ConvertDecimalPowerTable *table) {
    if (table == nullptr) return;
    if (table->powers) {
        for (uint32_t i = 0; i < table->levels; ++i) batch_mp_array_release(table->powers[i]);
        delete[] table->powers;
    }
    table->powers = nullptr;
    table->levels = 0;
}

// convert_decimal.cu:285-474.  The leaf digits go straight to their natural
// positions (the reference writes reversed chunks and reorders them).
hipError_t convert_decimal_equal_split(BatchMPContext *ctx, BatchMPArray input, uint32_t input_bits,
                                       uint32_t total_digits, uint32_t leaf_digits,
                                       const ConvertDecimalPowerTable *table, char *output_digits) {
    if (ctx == nullptr || output_digits == nullptr || input.batch_size != 1 || input.data == nullptr)
        return hipErrorInvalidValue;
    const uint32_t levels = levels_for_shape(total_digits, leaf_digits);
    if (levels == UINT32_MAX || table == nullptr || table->leaf_digits != leaf_digits || table->levels != levels)
        return hipErrorInvalidValue;
    std::vector<uint32_t> dig(levels + 1), bits(levels + 1), limbs(levels + 1);
    dig[0] = total_digits;
    bits[0] = input_bits;
    limbs[0] = input.length;
    for (uint32_t l = 1; l <= levels; ++l) {
        dig[l] = dig[l - 1] >> 1;
        bits[l] = bits_for_digits(dig[l]);
        limbs[l] = ceil_div(bits[l], 32);
    }
    Buf bcur(ctx), bnext(ctx), bmant(ctx), bprod(ctx);
    BatchMPArray cur = input;
    if (levels == 0) {
        const uint32_t need = bits_for_digits(total_digits);
        if (input_bits < need) return hipErrorInvalidValue;
        BatchMPArray t;
        GPUPI_RETURN_IF(bcur.view(1, ceil_div(need, 32), t));
        GPUPI_RETURN_IF(gpupi::launch_extract_shifted(input, t, 1, input_bits - need, 0, 2, ctx->stream));
        cur = t;
    } else {
        for (uint32_t l = 0; l < levels; ++l) {
            const uint32_t child_digits = dig[l + 1], child_bits = bits[l + 1];
            if (cur.length < limbs[l]) return hipErrorInvalidValue;
            if (bits[l] <= child_digits || bits[l] < child_bits) return hipErrorInvalidValue;
            BatchMPArray next, mant, prod;
            GPUPI_RETURN_IF(bnext.view(cur.batch_size * 2u, limbs[l + 1], next));
            GPUPI_RETURN_IF(gpupi::launch_extract_shifted(cur, next, cur.batch_size, bits[l] - child_bits, 0, 2,
                                                          ctx->stream));
            const uint32_t mant_bits = bits[l] - child_digits;
            if (l == 0) {
                // the caller's rows: a copy of their low bits
                GPUPI_RETURN_IF(bmant.view(cur.batch_size, ceil_div(mant_bits, 32), mant));
                GPUPI_RETURN_IF(gpupi::launch_extract_low(cur, mant, cur.batch_size, mant_bits, ctx->stream));
            } else {
                // this conversion's own rows, not needed again once their
                // left halves are out: their low bits kept in place
                mant = cur;
                GPUPI_RETURN_IF(gpupi::launch_keep_low_bits(mant, cur.batch_size, mant_bits, ctx->stream));
            }
            const BatchMPArray &pw = table->powers[levels - 1 - l];
            const BatchMPArray power = view_of(pw.data, pw.length, cur.batch_size, 0);
            GPUPI_RETURN_IF(bprod.view_aligned(cur.batch_size, mant.length + power.length, prod));
            GPUPI_RETURN_IF(batch_mp_mul(ctx, mant, power, prod));
            GPUPI_RETURN_IF(gpupi::launch_extract_shifted(prod, next, cur.batch_size, mant_bits - child_bits, 1, 2,
                                                          ctx->stream));
            bcur.swap(bnext);
            cur = next;
        }
    }
    Buf baligned(ctx);
    BatchMPArray aligned;
    GPUPI_RETURN_IF(baligned.view(cur.batch_size, cur.length, aligned));
    const uint32_t leaf_bits = levels == 0 ? bits_for_digits(total_digits) : bits[levels];
    GPUPI_RETURN_IF(batch_mp_shift_bits(ctx, cur, aligned, (int32_t)(cur.length * 32u - leaf_bits)));
    GPUPI_RETURN_IF(gpupi::launch_decimal_leaf(aligned.data, output_digits, aligned.batch_size, aligned.length,
                                               aligned.stride, leaf_digits, leaf_digits, true, ctx->stream));
    return hipStreamSynchronize(ctx->stream);
}

}  // extern "C"

namespace {

// One pi job's sizes (pi_bs_gpu.cu:2096-2136): digits converted, leaf digits
// of the conversion, precision in digits and limbs, conversion levels, and
// the binary-splitting tree's groups.
struct PiPlan {
    uint32_t requested, leaf, convert_digits, prec_digits, prec_limbs, levels, groups;
};
PiPlan pi_plan(uint32_t requested_digits, int prec_convention) {
    PiPlan pl{};
    pl.requested = requested_digits;
    pl.leaf = 477;
    pl.convert_digits = gpupi_round_up_decimal_digits(requested_digits, &pl.leaf);
    pl.prec_digits = prec_convention == GPUPI_PREC_ROUNDED ? pl.convert_digits : requested_digits;
    pl.prec_limbs = gpupi_prec_limbs_for_digits(pl.prec_digits);
    pl.levels = pl.convert_digits ? (uint32_t)__builtin_ctz(pl.convert_digits / pl.leaf) : 0u;
    pl.groups = pi_groups((int)pl.prec_digits);
    return pl;
}

// Everything of a pi job after its binary splitting: the rest of
// pi_fractional, the power table and the radix conversion
// (pi_bs_gpu.cu:2137-2193).  Q_bs, R_bs: the root rows of the tree over
// pl.groups groups, handed over.  elapsed_ms runs from `start`.
hipError_t pi_finish(BatchMPContext *ctx, const PiPlan &pl, BatchMPArray Q_bs, BatchMPArray R_bs, hipEvent_t start,
                     GpupiPiResult *result) {
    hipEvent_t stop = nullptr;
    if (const hipError_t e = hipEventCreate(&stop); e != hipSuccess) {
        gpupi::ctx_free(ctx, Q_bs.data);
        gpupi::ctx_free(ctx, R_bs.data);
        return e;
    }
    struct EventGuard {
        hipEvent_t e;
        ~EventGuard() { (void)hipEventDestroy(e); }
    } guard{stop};
    BatchMPArray ret{};
    uint32_t groups = 0;
    uint64_t inverse_bits = 0;
    GPUPI_RETURN_IF(pi_fractional_from_tree(ctx, (int)pl.prec_digits, (int)pl.groups, Q_bs, R_bs, &ret, &groups,
                                            &inverse_bits));
    CachedPowers powers;
    hipError_t err = stage(ctx, "decimal_powers", "precompute_5pow", pl.convert_digits, pl.levels, pl.leaf, 0, false,
                           [&] { return powers.fill(ctx, pl.leaf, pl.levels); });
    char *digits = nullptr;
    if (err == hipSuccess)
        err = gpupi::ctx_alloc(ctx, reinterpret_cast<void **>(&digits), std::max<uint32_t>(pl.convert_digits, 1));
    if (err == hipSuccess && pl.requested > 0) {
        const BatchMPArray frac = make_subview(ret, 0, std::min(pl.prec_limbs, ret.length));
        err = stage(ctx, "decimal_convert", "equal_split", pl.convert_digits, 1, frac.length, 0, false, [&] {
            return convert_decimal_equal_split(ctx, frac, pl.prec_limbs * 32u, pl.convert_digits, pl.leaf,
                                               &powers.table, digits);
        });
    }
    if (err == hipSuccess) err = hipEventRecord(stop, ctx->stream);
    if (err == hipSuccess) err = hipEventSynchronize(stop);
    float ms = 0.0f;
    if (err == hipSuccess) err = hipEventElapsedTime(&ms, start, stop);
    if (err != hipSuccess) {
        gpupi::ctx_free(ctx, ret.data);
        gpupi::ctx_free(ctx, digits);
        return err;
    }
    result->requested_digits = pl.requested;
    result->convert_digits = pl.convert_digits;
    result->leaf_digits = pl.leaf;
    result->prec_limbs = pl.prec_limbs;
    result->groups = groups;
    result->ret_length = ret.length;
    result->ret = ret.data;
    result->digits = digits;
    result->elapsed_ms = ms;
    result->inverse_bits = inverse_bits;
    return hipSuccess;
}

struct StartEvent {
    hipEvent_t e = nullptr;
    ~StartEvent() {
        if (e) (void)hipEventDestroy(e);
    }
};

}  // namespace

extern "C" {

// pi_bs_gpu.cu:2073-2193: the reference's timed region.
hipError_t gpupi_pi_run(BatchMPContext *ctx, uint32_t requested_digits, int prec_convention, GpupiPiResult *result) {
    if (ctx == nullptr || result == nullptr) return hipErrorInvalidValue;
    std::memset(result, 0, sizeof(*result));
    const PiPlan pl = pi_plan(requested_digits, prec_convention);
    StartEvent start;
    GPUPI_RETURN_IF(hipEventCreate(&start.e));
    GPUPI_RETURN_IF(hipEventRecord(start.e, ctx->stream));
    BatchMPArray Q_bs{}, R_bs{};
    GPUPI_RETURN_IF(gpupi_binary_split(ctx, pl.groups, &Q_bs, &R_bs));
    return pi_finish(ctx, pl, Q_bs, R_bs, start.e, result);
}

// Several pi jobs with one binary-splitting tree (no counterpart in the
// reference, which runs one job per process).  The tree over N groups is the
// tree over every smaller power of two of groups continued: the leaf terms
// depend only on the term index, and a level's sign and shift only on how
// many groups its nodes cover, so the root of a job over N_i groups is row
// 0 of the level whose nodes cover N_i groups in the tree over the largest
// N.  That tree runs once; each job's root is copied out (and compacted to
// its own length) when its level is reached, and the job's remaining stages
// run at once, smallest jobs first.  Results equal gpupi_pi_run's for each
// job.  In the plain form the jobs run one after the other.
hipError_t gpupi_pi_run_batch(BatchMPContext *ctx, const uint32_t *requested_digits, int count, int prec_convention,
                              GpupiPiResult *results) {
    if (ctx == nullptr || requested_digits == nullptr || results == nullptr || count <= 0) return hipErrorInvalidValue;
    for (int i = 0; i < count; ++i) std::memset(&results[i], 0, sizeof(results[i]));
    const auto fail = [&](hipError_t e) {
        for (int i = 0; i < count; ++i) {
            gpupi::ctx_free(ctx, results[i].ret);
            gpupi::ctx_free(ctx, results[i].digits);
            std::memset(&results[i], 0, sizeof(results[i]));
        }
        return e;
    };
    if (ctx->pipeline_form == GPUPI_FORM_PLAIN) {
        for (int i = 0; i < count; ++i) {
            const hipError_t e = gpupi_pi_run(ctx, requested_digits[i], prec_convention, &results[i]);
            if (e != hipSuccess) return fail(e);
        }
        return hipSuccess;
    }
    std::vector<PiPlan> plans(count);
    uint32_t n_max = 1;
    for (int i = 0; i < count; ++i) {
        plans[i] = pi_plan(requested_digits[i], prec_convention);
        n_max = std::max(n_max, plans[i].groups);
    }
    StartEvent start;
    hipError_t err = hipEventCreate(&start.e);
    if (err == hipSuccess) err = hipEventRecord(start.e, ctx->stream);
    if (err != hipSuccess) return fail(err);
    TreeBufs t(ctx);
    BatchMPArray P, Q, R;
    // the jobs whose tree covers `groups` groups: root = row 0 of Q and R,
    // copied out for each (at the top level the last one takes the tree's
    // own buffers, and the tree's other buffers go back before the jobs'
    // remaining stages run)
    const auto finish_level = [&](uint32_t groups, bool top) -> hipError_t {
        int last = -1;
        for (int i = 0; i < count; ++i)
            if (plans[i].groups == groups) last = i;
        if (last < 0) return hipSuccess;
        struct Root {
            int job;
            BatchMPArray Q, R;
            std::unique_ptr<Buf> bq, br;  // own the rows until the job takes them
        };
        std::vector<Root> roots;
        for (int i = 0; i <= last; ++i) {
            if (plans[i].groups != groups) continue;
            Root r{i, {}, {}, std::unique_ptr<Buf>(new Buf(ctx)), std::unique_ptr<Buf>(new Buf(ctx))};
            if (top && i == last) {
                r.Q = view_of(t.bQ.ptr, Q.length, 1, Q.stride);
                r.R = view_of(t.bR.ptr, R.length, 1, R.stride);
                r.bq->swap(t.bQ);
                r.br->swap(t.bR);
            } else {
                GPUPI_RETURN_IF(r.bq->view(1, Q.length, r.Q));
                GPUPI_RETURN_IF(r.br->view(1, R.length, r.R));
                GPUPI_RETURN_IF(hipMemcpyAsync(r.Q.data, Q.data, (size_t)Q.length * 4, hipMemcpyDeviceToDevice,
                                               ctx->stream));
                GPUPI_RETURN_IF(hipMemcpyAsync(r.R.data, R.data, (size_t)R.length * 4, hipMemcpyDeviceToDevice,
                                               ctx->stream));
                GPUPI_RETURN_IF(compact_many(ctx, {&r.Q, &r.R}, "Q,R", groups));
            }
            roots.push_back(std::move(r));
        }
        if (top) t.drop();
        for (Root &r : roots) {
            r.bq->release();
            r.br->release();
            GPUPI_RETURN_IF(pi_finish(ctx, plans[r.job], r.Q, r.R, start.e, &results[r.job]));
        }
        return hipSuccess;
    };
    err = leaf_stage_fused(ctx, t, n_max, P, Q, R);
    if (err == hipSuccess) err = compact_many(ctx, {&P, &Q, &R}, "P,Q,R", 2 * n_max);
    if (err == hipSuccess) err = finish_level(1, n_max == 1);
    for (uint32_t n1 = n_max; err == hipSuccess && n1 > 1; n1 >>= 1) {
        err = tree_level(ctx, t, n_max, n1, Lengths::kShared, P, Q, R);
        if (err == hipSuccess) err = finish_level(2 * n_max / n1, n1 == 2);
    }
    return err == hipSuccess ? hipSuccess : fail(err);
}

hipError_t gpupi_set_memory_pool(BatchMPContext *ctx, hipMemPool_t pool) {
    if (ctx == nullptr) return hipErrorInvalidValue;
    if (pool != ctx->cache->pool()) {
        GPUPI_RETURN_IF(ctx->cache->trim());
        if (ctx->workspace != nullptr) {
            GPUPI_RETURN_IF(hipStreamSynchronize(ctx->stream));
            if (ctx->cache->pool() != nullptr) GPUPI_RETURN_IF(hipFreeAsync(ctx->workspace, ctx->stream));
            else GPUPI_RETURN_IF(hipFree(ctx->workspace));
            GPUPI_RETURN_IF(hipStreamSynchronize(ctx->stream));
            ctx->workspace = nullptr;
            ctx->workspace_size_bytes = 0;
        }
    }
    ctx->cache->set_pool(pool);
    return hipSuccess;
}

hipError_t gpupi_set_pipeline_form(BatchMPContext *ctx, int form) {
    if (ctx == nullptr || (form != GPUPI_FORM_BATCHED && form != GPUPI_FORM_PLAIN)) return hipErrorInvalidValue;
    ctx->pipeline_form = form;
    return hipSuccess;
}

hipError_t gpupi_stage_profile_enable(BatchMPContext *ctx, int enable) {
    if (ctx == nullptr) return hipErrorInvalidValue;
    if (ctx->profiler == nullptr) {
        if (!enable) return hipSuccess;
        StageProfiler *p = new (std::nothrow) StageProfiler();
        if (p == nullptr) return hipErrorOutOfMemory;
        if (hipEventCreate(&p->start) != hipSuccess || hipEventCreate(&p->stop) != hipSuccess) {
            stage_profiler_free(p);
            return hipErrorOutOfMemory;
        }
        ctx->profiler = p;
    }
    ctx->profiler->enabled = enable != 0;
    if (enable) ctx->profiler->entries.clear();
    return hipSuccess;
}

int gpupi_stage_profile_query(const BatchMPContext *ctx, GpupiStageEntry *out, int max_entries) {
    if (ctx == nullptr) return -1;
    if (ctx->profiler == nullptr) return 0;
    const int n = (int)ctx->profiler->entries.size();
    for (int i = 0; i < n && i < max_entries && out != nullptr; ++i) out[i] = ctx->profiler->entries[i];
    return n;
}

hipError_t gpupi_leaf_stage_create(BatchMPContext *ctx, uint32_t groups, int mode, GpupiLeafStage **out) {
    if (ctx == nullptr || out == nullptr || groups == 0 || (groups & (groups - 1)) != 0 ||
        groups > (1u << 26) || mode < GPUPI_LEAF_PLAIN || mode > GPUPI_LEAF_FUSED)
        return hipErrorInvalidValue;
    *out = nullptr;
    GpupiLeafStage *s = new (std::nothrow) GpupiLeafStage();
    if (s == nullptr) return hipErrorOutOfMemory;
    s->ctx = ctx;
    s->N = groups;
    s->mode = mode;
    if (mode == GPUPI_LEAF_GRAPH) {
        hipError_t err = hipStreamCreate(&s->stream);
        if (err == hipSuccess) err = with_stage_context(s, [&] { return leaf_stage_capture(s); });
        if (err != hipSuccess) {
            gpupi_leaf_stage_destroy(s);
            return err;
        }
    }
    *out = s;
    return hipSuccess;
}

hipError_t gpupi_leaf_stage_run(GpupiLeafStage *s, BatchMPArray *P, BatchMPArray *Q, BatchMPArray *R) {
    if (s == nullptr || P == nullptr || Q == nullptr || R == nullptr) return hipErrorInvalidValue;
    BatchMPContext *ctx = s->ctx;
    BatchMPArray p, q, r;
    if (s->mode == GPUPI_LEAF_GRAPH) {
        p = s->P;
        q = s->Q;
        r = s->R;
        GPUPI_RETURN_IF(with_stage_context(s, [&]() -> hipError_t {
            GPUPI_RETURN_IF(hipGraphLaunch(s->exec, s->stream));
            GPUPI_RETURN_IF(compact(ctx, p, "P", 2 * s->N));
            GPUPI_RETURN_IF(compact(ctx, q, "Q", 2 * s->N));
            return compact(ctx, r, "R", 2 * s->N);
        }));
    } else {
        const bool plain = s->mode == GPUPI_LEAF_PLAIN;
        if (s->mode == GPUPI_LEAF_FUSED) GPUPI_RETURN_IF(leaf_stage_fused(ctx, s->t, s->N, p, q, r));
        else GPUPI_RETURN_IF(leaf_stage(ctx, s->t, s->N, plain ? Lengths::kPerOp : Lengths::kBound, p, q, r));
        if (!plain) {
            GPUPI_RETURN_IF(compact(ctx, p, "P", 2 * s->N));
            GPUPI_RETURN_IF(compact(ctx, q, "Q", 2 * s->N));
            GPUPI_RETURN_IF(compact(ctx, r, "R", 2 * s->N));
        }
        GPUPI_RETURN_IF(hipStreamSynchronize(ctx->stream));
    }
    *P = p;
    *Q = q;
    *R = r;
    return hipSuccess;
}

void gpupi_leaf_stage_destroy(GpupiLeafStage *s) {
    if (s == nullptr) return;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    if (s->exec) (void)hipGraphExecDestroy(s->exec);
    if (s->graph) (void)hipGraphDestroy(s->graph);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    if (s->ws) (void)hipFree(s->ws);
    delete s;
}

hipError_t gpupi_cache_query(const BatchMPContext *ctx, size_t *cached_bytes, uint64_t *allocations) {
    if (ctx == nullptr) return hipErrorInvalidValue;
    if (cached_bytes) *cached_bytes = ctx->cache->cached_bytes();
    if (allocations) *allocations = ctx->cache->allocations();
    return hipSuccess;
}

hipError_t gpupi_cache_trim(BatchMPContext *ctx) {
    if (ctx == nullptr) return hipErrorInvalidValue;
    return ctx->cache->trim();
}

void gpupi_pi_free(GpupiPiResult *result) {
    if (result == nullptr) return;
    if (result->ret) (void)hipFree(result->ret);
    if (result->digits) (void)hipFree(result->digits);
    result->ret = nullptr;
    result->digits = nullptr;
}

}  // extern "C"
