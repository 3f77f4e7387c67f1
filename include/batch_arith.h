/*
 * gpupi-mi355x: batched multi-precision arithmetic, C ABI.
 *
 * Drop-in for the reference's batch_arith.h (fanhqme2/gpupi
 * batch_arith.h:8-152).  Same names, argument order and meaning; the entry
 * points have C linkage and return hipError_t where the reference returns
 * cudaError_t (hipErrorInvalidValue <-> cudaErrorInvalidValue,
 * hipErrorOutOfMemory <-> cudaErrorMemoryAllocation).
 *
 * Conventions (reference batch_arith.h:16-24, batch_arith.cu:89-100):
 *   - data is a device pointer; limbs are little-endian uint32;
 *   - row i starts at data + i * stride; stride >= length, or stride == 0 to
 *     broadcast one row to every batch item;
 *   - results are truncated modulo 2^(32 * C.length), high words zero-filled.
 * Work is issued on the context's stream (the null stream unless
 * batch_mp_set_stream was called).  The *_max / compact calls synchronize.
 */
#ifndef GPUPI_BATCH_ARITH_H
#define GPUPI_BATCH_ARITH_H

#include <stddef.h>
#include <stdint.h>

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
#define GPUPI_DEFAULT(v) = v
extern "C" {
#else
#define GPUPI_DEFAULT(v)
#endif

typedef struct BatchMPContext BatchMPContext;

/* reference batch_arith.h:10-17 */
typedef struct BatchMPArray {
    uint32_t *data;
    uint32_t length;
    uint32_t batch_size;
    uint32_t stride;
#ifdef __cplusplus
    /* reference batch_arith.cu:415-424: length <- max non-zero limb count */
    inline hipError_t compact(BatchMPContext *ctx);
#endif
} BatchMPArray;

/* reference batch_arith.h:19-28: stride 0 means stride = length; on failure
 * the returned array has data == NULL. */
BatchMPArray batch_mp_array_create(uint32_t batch_size, uint32_t length,
                                   uint32_t stride GPUPI_DEFAULT(0));
/* reference batch_arith.h:30-31 */
void batch_mp_array_release(BatchMPArray array);

/* reference batch_arith.h:33-35: NULL on failure */
BatchMPContext *batch_mp_init(void);
/* reference batch_arith.h:37-38 */
void batch_mp_destroy(BatchMPContext *ctx);
/* reference batch_arith.h:40-41 */
size_t batch_mp_workspace_size(const BatchMPContext *ctx);
/* reference batch_arith.h:43 */
hipError_t batch_mp_ensure_workspace(BatchMPContext *ctx, size_t required_bytes);

/* reference batch_arith.h:45-53: C = A * B; C.length must equal
 * A.length + B.length.  Schoolbook or NTT is chosen internally. */
hipError_t batch_mp_mul(BatchMPContext *ctx, BatchMPArray A, BatchMPArray B, BatchMPArray C);
/* Ragged batched product (no counterpart in the reference, whose batch API
 * has one L_a/L_b per call): items[i].C = items[i].A * items[i].B with
 * L_a + L_b limbs written, for pairs of any lengths.  Items are grouped by
 * transform size and each group runs as one launch sequence; an unbalanced
 * pair may run as pieces of its longer operand sharing the shorter one's
 * transform, and items passing the same B (pointer and length) share its
 * transform.  Descriptors are read from host memory before the call
 * returns.  No item's C may overlap any item's operands, and distinct
 * items' C must not overlap. */
typedef struct GpupiMulItem {
    const uint32_t *A;
    const uint32_t *B;
    uint32_t *C;
    uint32_t L_a;
    uint32_t L_b;
} GpupiMulItem;
hipError_t gpupi_mul_ragged(BatchMPContext *ctx, const GpupiMulItem *items, uint32_t count);
/* reference batch_arith.h:55-63: C = A * B (single host limb); C != A */
hipError_t batch_mp_mul_small(BatchMPContext *ctx, BatchMPArray A, uint32_t B, BatchMPArray C);
/* reference batch_arith.h:65-72: A /= B in place; B odd, A.length <= 128 */
hipError_t batch_mp_exactdiv_small(BatchMPContext *ctx, BatchMPArray A, uint32_t B);
/* reference batch_arith.h:74-82: in place allowed (C == A or C == B) */
hipError_t batch_mp_add(BatchMPContext *ctx, BatchMPArray A, BatchMPArray B, BatchMPArray C);
/* reference batch_arith.h:84-92 */
hipError_t batch_mp_add_small(BatchMPContext *ctx, BatchMPArray A, uint32_t B, BatchMPArray C);
/* reference batch_arith.h:94-104: C = (A << shift) + B; C aliases neither */
hipError_t batch_mp_shift_add(BatchMPContext *ctx, BatchMPArray A, BatchMPArray B,
                              uint32_t shift, BatchMPArray C);
/* reference batch_arith.h:106-116: C = (A << shift) - B */
hipError_t batch_mp_shift_sub(BatchMPContext *ctx, BatchMPArray A, BatchMPArray B,
                              uint32_t shift, BatchMPArray C);
/* reference batch_arith.h:118-125: C = A - B mod 2^(32 C.length) */
hipError_t batch_mp_sub(BatchMPContext *ctx, BatchMPArray A, BatchMPArray B, BatchMPArray C);
/* reference batch_arith.h:127-134 */
hipError_t batch_mp_sub_small(BatchMPContext *ctx, BatchMPArray A, uint32_t B, BatchMPArray C);
/* reference batch_arith.h:136-143: B = A << s (s > 0) or A >> -s */
hipError_t batch_mp_shift_bits(BatchMPContext *ctx, BatchMPArray A, BatchMPArray B,
                               int32_t shift_bits);
/* reference batch_arith.h:145-146 (synchronizes) */
hipError_t batch_mp_bitlength_max(BatchMPContext *ctx, BatchMPArray A, uint32_t *result);
/* reference batch_arith.h:148-152 (synchronizes) */
hipError_t batch_mp_limblength_max(BatchMPContext *ctx, BatchMPArray A, uint32_t *result);

/* C spelling of BatchMPArray::compact (reference batch_arith.h:16) */
hipError_t batch_mp_array_compact(BatchMPContext *ctx, BatchMPArray *array);

/* Extension (no reference counterpart: the reference always uses the legacy
 * default stream).  Subsequent work of this context goes to `stream`. */
hipError_t batch_mp_set_stream(BatchMPContext *ctx, hipStream_t stream);
hipStream_t batch_mp_get_stream(const BatchMPContext *ctx);

/* Extension: products with A.length + B.length >= min_total_limbs use the
 * NTT, shorter ones the schoolbook kernels (the reference fixes 1024,
 * batch_arith.cu:182). */
hipError_t batch_mp_set_ntt_threshold(BatchMPContext *ctx, uint32_t min_total_limbs);

/* Extension (no reference counterpart): two products with a common operand,
 * C1 = A1 * B and C2 = A2 * B, row for row (C1.length = A1.length +
 * B.length, likewise C2).  A1/C1 and A2/C2 have equal batch sizes, which may
 * differ; B has at least the larger batch size, or is one row at stride 0.
 * Results equal two batch_mp_mul calls; when both products take the NTT
 * path with one transform size, B's transform is computed once (the
 * binary-splitting tree multiplies P_e by P_o and R_o, and Q_o by Q_e and
 * R_e, pi_bs_gpu.cu:637-711).  Outputs must not alias inputs or each other. */
hipError_t batch_mp_mul_shared(BatchMPContext *ctx, BatchMPArray A1, BatchMPArray A2, BatchMPArray B,
                               BatchMPArray C1, BatchMPArray C2);

/* Bytes of the static per-device scratch of batch_mp_mul_noalloc. */
#define GPUPI_NOALLOC_SCRATCH_BYTES ((size_t)128 << 20)

/* Extension with no counterpart in the reference, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops): batch_mp_mul with the same arguments, validation and results, but
 * no hipMalloc, hipFree or allocating library call.  NTT products use a
 * static per-device scratch array of GPUPI_NOALLOC_SCRATCH_BYTES (looked up
 * once per device by batch_mp_init or the first call) instead of the
 * context's workspace, running their rows in chunks that fit it; a product
 * whose single row needs more (two distinct operands: L_a + L_b above 2^21
 * limbs) is rejected with hipErrorInvalidValue.  Every call on
 * one device shares that device's scratch: calls that overlap on one device
 * (two streams, two threads) race and must be serialized by the caller.
 * Kernel statistics are not recorded for these calls. */
hipError_t batch_mp_mul_noalloc(BatchMPContext *ctx, BatchMPArray A, BatchMPArray B, BatchMPArray C);

#ifdef __cplusplus
}

inline hipError_t BatchMPArray::compact(BatchMPContext *ctx) {
    return batch_mp_array_compact(ctx, this);
}
#endif

#undef GPUPI_DEFAULT

#endif
