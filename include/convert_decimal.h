/*
 * gpupi-mi355x: binary fraction -> decimal digits by equal splits, C ABI.
 * Drop-in for the reference's convert_decimal.h:7-41.
 */
#ifndef GPUPI_CONVERT_DECIMAL_H
#define GPUPI_CONVERT_DECIMAL_H

#include <stdint.h>

#include "batch_arith.h"

#ifdef __cplusplus
extern "C" {
#endif

/* reference convert_decimal.h:7-11 */
typedef struct ConvertDecimalPowerTable {
    BatchMPArray *powers;
    uint32_t levels;
    uint32_t leaf_digits;
} ConvertDecimalPowerTable;

/* reference convert_decimal.h:13-21: powers 5^(leaf_digits << level) for
 * level < levels, one right-aligned row each (broadcast with stride 0). */
hipError_t convert_decimal_precompute_powers(BatchMPContext *ctx, ConvertDecimalPowerTable *table,
                                             uint32_t leaf_digits, uint32_t levels);

/* reference convert_decimal.h:23 */
void convert_decimal_release_powers(ConvertDecimalPowerTable *table);

/* reference convert_decimal.h:25-41: input (batch 1) holds X; the value
 * X / 2^input_bits becomes total_digits = leaf_digits << levels ASCII digits
 * at output_digits (device), natural order (output_digits[0] is 10^-1). */
hipError_t convert_decimal_equal_split(BatchMPContext *ctx, BatchMPArray input, uint32_t input_bits,
                                       uint32_t total_digits, uint32_t leaf_digits,
                                       const ConvertDecimalPowerTable *table, char *output_digits);

#ifdef __cplusplus
}
#endif

#endif
