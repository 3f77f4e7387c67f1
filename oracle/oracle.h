/*
 * CPU oracle for the gpupi pi pipeline — TEST INFRASTRUCTURE ONLY.
 *
 * A restatement, on GMP, of the reference's algorithm (fanhqme2/gpupi), op for
 * op, used to check the HIP library.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it.  Nothing in the product links it.
 *
 * Arrays follow the reference's BatchMPArray convention (batch_arith.h:10-17):
 * row i of a batch starts at data + i * stride; limbs are little-endian
 * uint32; stride == 0 broadcasts one row.  Here data points to host memory.
 */
#ifndef GPUPI_ORACLE_H
#define GPUPI_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    uint32_t *data;
    uint32_t length;
    uint32_t batch_size;
    uint32_t stride;
} OArray;

/* Batched big-integer ops; semantics of the reference's batch_mp_* calls.
 * Return 0 on success, -1 on invalid arguments (the reference's
 * cudaErrorInvalidValue cases), -2 if an exact division was not exact. */
int o_mul(OArray A, OArray B, OArray C);
/* threads for the rows of o_mul (default 1; CPU-baseline runs only) */
void o_set_threads(int threads);
int o_mul_small(OArray A, uint32_t b, OArray C);
int o_exactdiv_small(OArray A, uint32_t b);
int o_add(OArray A, OArray B, OArray C);
int o_add_small(OArray A, uint32_t b, OArray C);
int o_sub(OArray A, OArray B, OArray C);
int o_sub_small(OArray A, uint32_t b, OArray C);
int o_shift_add(OArray A, OArray B, uint32_t shift, OArray C);
int o_shift_sub(OArray A, OArray B, uint32_t shift, OArray C);
int o_shift_bits(OArray A, OArray B, int32_t shift_bits);
uint32_t o_limblength_max(OArray A);
uint32_t o_bitlength_max(OArray A);

/* Fixed-point fraction rows -> ASCII digits, least significant first
 * (batch_decimal_naive.h:7-31). */
void o_decimal_naive(const uint32_t *A, char *B, uint32_t N, int L_limb,
                     uint32_t stride_A, int L_dec, uint32_t stride_B);

/* Pipeline stages. */
void o_leaf17(int is_head, uint32_t *P, uint32_t *Q, uint32_t *R, uint32_t count);
/* Returns malloc'd Q and R rows (batch 1) of the whole 17*N-term series. */
int o_binary_split(uint32_t N, OArray *Q_out, OArray *R_out);
int o_leaf_stage(uint32_t N, OArray *P_out, OArray *Q_out, OArray *R_out);
int o_sqrt_10005(int target_digits, OArray *P_out, OArray *Q_out);
/* y = approximate 2^L / x, length x.length (pi_bs_gpu.cu:1179-1344). */
int o_recursive_inverse(OArray x, uint64_t L, OArray *y_out);

typedef struct {
    uint32_t target_digits;   /* digits the series/precision are sized for */
    uint32_t groups;          /* N (17-term groups) */
    uint32_t prec_limbs;
    uint64_t inverse_bits;
} OPiInfo;
int o_pi_fractional(int target_digits, OArray *ret_out, OPiInfo *info);

uint32_t o_round_up_decimal_digits(uint32_t requested, uint32_t *leaf_digits);
uint32_t o_prec_limbs_for_digits(uint32_t digits);
uint32_t o_bits_for_digits(uint32_t digits);

/* 5^(leaf << i) for i < levels; rows malloc'd, lengths compacted. */
int o_convert_precompute_powers(uint32_t leaf_digits, uint32_t levels, OArray *powers);
int o_convert_equal_split(OArray input, uint32_t input_bits, uint32_t total_digits,
                          uint32_t leaf_digits, const OArray *powers, uint32_t levels,
                          char *output_digits);

/* Whole program (pi_bs_gpu.cu:2073-2195): RET limbs and the first
 * `requested` decimal digits after the "3".  readme_prec != 0 sizes the
 * precision from the rounded-up digit count (the convention of the
 * reference README's published RET length and hash). */
int o_pi_digits(uint32_t requested, int readme_prec, uint32_t **ret_limbs,
                uint32_t *ret_len, char **digits, uint32_t *convert_digits_out);

uint32_t o_mod_hash(const uint32_t *limbs, uint32_t length);
uint32_t o_decimal_hash(const char *digits, uint32_t length);

void o_free(void *p);

#ifdef __cplusplus
}
#endif

#endif
