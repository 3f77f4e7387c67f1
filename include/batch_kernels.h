/*
 * gpupi-mi355x: per-op batched kernels (the reference's L1 layer), C ABI.
 *
 * Each entry point keeps the name, argument order and meaning of the
 * reference's per-op header cited beside it.  As in the reference these
 * functions do not validate beyond cheap host checks and return void; call
 * hipGetLastError() (or use the validated batch_mp_* layer in batch_arith.h).
 * All pointers are device pointers.  Work goes to the null stream.
 */
#ifndef GPUPI_BATCH_KERNELS_H
#define GPUPI_BATCH_KERNELS_H

#include <stddef.h>
#include <stdint.h>

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* reference batch_add.h:6-40: C = A + B, truncated/zero-filled to L_c;
 * in place allowed when the aliased operand has C's length and stride. */
void batch_add_simple(uint32_t *A, uint32_t *B, uint32_t *C, uint32_t *workspace,
                      uint32_t N, uint32_t L_a, uint32_t L_b, uint32_t L_c,
                      uint32_t stride_A, uint32_t stride_B, uint32_t stride_C);
size_t batch_add_simple_workspace_size(uint32_t N, uint32_t L_a, uint32_t L_b, uint32_t L_c);

/* reference batch_sub.h:6-36: C = A - B mod 2^(32 L_c) */
void batch_sub_simple(uint32_t *A, uint32_t *B, uint32_t *C, uint32_t *workspace,
                      uint32_t N, uint32_t L_a, uint32_t L_b, uint32_t L_c,
                      uint32_t stride_A, uint32_t stride_B, uint32_t stride_C);
size_t batch_sub_simple_workspace_size(uint32_t N, uint32_t L_a, uint32_t L_b, uint32_t L_c);

/* reference batch_add_small.h:6-35: C = A + B (host scalar) */
void batch_add_small(const uint32_t *A, uint32_t B, uint32_t *C, uint32_t *workspace,
                     uint32_t N, uint32_t L_a, uint32_t L_c, uint32_t stride_A, uint32_t stride_C);
size_t batch_add_small_workspace_size(uint32_t N, uint32_t L_a, uint32_t L_c);

/* reference batch_sub_small.h:6-33: C = A - B mod 2^(32 L_c) */
void batch_sub_small(const uint32_t *A, uint32_t B, uint32_t *C, uint32_t *workspace,
                     uint32_t N, uint32_t L_a, uint32_t L_c, uint32_t stride_A, uint32_t stride_C);
size_t batch_sub_small_workspace_size(uint32_t N, uint32_t L_a, uint32_t L_c);

/* reference batch_mul_small.h:6-35: C = A * B over L_a + 1 limbs; C != A */
void batch_mul_small(const uint32_t *A, uint32_t B, uint32_t *C, uint32_t *workspace,
                     uint32_t N, uint32_t L_a, uint32_t L_c, uint32_t stride_A, uint32_t stride_C);
size_t batch_mul_small_workspace_size(uint32_t N, uint32_t L_a, uint32_t L_c);

/* reference batch_shift_add.h:6-43: C = (A << shift) + B; C aliases neither */
void batch_shift_add_simple(const uint32_t *A, const uint32_t *B, uint32_t *C, uint32_t *workspace,
                            uint32_t shift, uint32_t N, uint32_t L_a, uint32_t L_b, uint32_t L_c,
                            uint32_t stride_A, uint32_t stride_B, uint32_t stride_C);
size_t batch_shift_add_simple_workspace_size(uint32_t N, uint32_t L_a, uint32_t L_b,
                                             uint32_t L_c, uint32_t shift);

/* reference batch_shift_sub.h:6-40: C = (A << shift) - B mod 2^(32 L_c) */
void batch_shift_sub_simple(const uint32_t *A, const uint32_t *B, uint32_t *C, uint32_t *workspace,
                            uint32_t shift, uint32_t N, uint32_t L_a, uint32_t L_b, uint32_t L_c,
                            uint32_t stride_A, uint32_t stride_B, uint32_t stride_C);
size_t batch_shift_sub_simple_workspace_size(uint32_t N, uint32_t L_a, uint32_t L_b,
                                             uint32_t L_c, uint32_t shift);

/* reference batch_shift.h:5-47: B = A << s (s > 0) or A >> -s */
void batch_shift_bits(const uint32_t *A, uint32_t *B, uint32_t N, uint32_t L_in, uint32_t L_out,
                      uint32_t stride_in, uint32_t stride_out, int32_t shift_bits);

/* reference batch_exactdiv_small.h:5-16: A /= B in place, B odd, L <= 128 */
void batch_exactdiv_small(uint32_t *A, uint32_t B, uint32_t N, uint32_t L, uint32_t stride_A);

/* reference batch_bitlength.h:5-35 (synchronous) */
uint32_t batch_bitlength_max(const uint32_t *A, uint32_t *workspace, uint32_t N, uint32_t L,
                             uint32_t stride_A);
uint32_t batch_limblength_max(const uint32_t *A, uint32_t *workspace, uint32_t N, uint32_t L,
                              uint32_t stride_A);
size_t batch_bitlength_workspace_size(uint32_t N, uint32_t L);

/* reference batch_mul_naive.h:5-25: schoolbook product, L_a + L_b limbs.
 * The length limit here is this library's own (the reference's is 1024). */
#define BATCH_MUL_NAIVE_L_MAX 16384
void batch_mul_naive(uint32_t *A, uint32_t *B, uint32_t *ret, int N, int L_a, int L_b,
                     int stride_A, int stride_B, int stride_ret);

/* reference batch_mul_ntt.h:5-11.  The reference's tables are uint3 powers of
 * a root mod 2^96-2^32+1; this library's NTT works over two primes below 2^50
 * in binary64 arithmetic, so the struct carries its own tables. */
typedef struct NTTPrecomputedTables {
    double *roots;         /* device: root power tables, both primes */
    uint32_t max_log2;     /* largest supported transform length, log2 */
} NTTPrecomputedTables;

/* reference batch_mul_ntt.h:13: allocates and fills the tables (device) */
void init_ntt_precomputed_tables(NTTPrecomputedTables *tables);
void release_ntt_precomputed_tables(NTTPrecomputedTables *tables);

/* reference batch_mul_ntt.h:15-37: exact product of N pairs by NTT */
void batch_mul_ntt(uint32_t *A, uint32_t *B, uint32_t *ret, uint32_t *workspace,
                   NTTPrecomputedTables tables, uint32_t N, uint32_t L_a, uint32_t L_b,
                   uint32_t stride_A, uint32_t stride_B, uint32_t stride_ret);
/* reference batch_mul_ntt.h:39-40: bytes */
size_t batch_mul_ntt_workspace_size(uint32_t N, uint32_t L_a, uint32_t L_b);

/* reference batch_decimal_naive.h:3-31: N fractions X / 2^(32 L_limb) to
 * L_dec ASCII digits each, least significant first. */
#define BATCH_DECIMAL_NAIVE_L_LIMB_MAX 2047
void batch_decimal_naive(const uint32_t *A, char *B, uint32_t N, int L_limb, uint32_t stride_A,
                         int L_dec, uint32_t stride_B);

#ifdef __cplusplus
}
#endif

#endif
