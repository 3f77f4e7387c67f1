/*
 * gpupi-mi355x: the pi pipeline as a library call, C ABI.
 *
 * The reference exposes the pipeline only through its CLI
 * (pi_bs_gpu.cu:2073-2431, `pi_bs_gpu <digits> [--benchmark|--profile-stages|
 * --full-print]`).  These entry points expose the same computation to other
 * hosts (the CLI in this library, tests, benchmarks):
 *   gpupi_pi_run   = main_fractional's timed region (pi_bs_gpu.cu:2137-2193):
 *                    pi_fractional + convert_decimal_precompute_powers +
 *                    convert_decimal_equal_split;
 *   the stage entry points mirror pi_bs_gpu.cu's binary_split_batched (503),
 *   sqrt_10005 (1346), recursive_inverse (1179) and pi_fractional (1524).
 */
#ifndef GPUPI_PI_H
#define GPUPI_PI_H

#include <stdint.h>

#include "batch_arith.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Precision convention.  GPUPI_PREC_REQUESTED sizes the precision from the
 * requested digit count (the reference's current code, pi_bs_gpu.cu:2081);
 * GPUPI_PREC_ROUNDED sizes it from the rounded-up conversion digit count,
 * which reproduces the reference README's published RET length and hash. */
enum { GPUPI_PREC_REQUESTED = 0, GPUPI_PREC_ROUNDED = 1 };

typedef struct GpupiPiResult {
    uint32_t requested_digits;
    uint32_t convert_digits;   /* leaf_digits << levels */
    uint32_t leaf_digits;
    uint32_t prec_limbs;
    uint32_t groups;           /* 17-term groups N */
    uint32_t ret_length;       /* limbs of RET */
    uint32_t *ret;             /* device, RET limbs (owned; gpupi_pi_free) */
    char *digits;              /* device, convert_digits ASCII digits after "3." */
    float elapsed_ms;          /* HIP-event time of the timed region */
    uint64_t inverse_bits;
} GpupiPiResult;

/* Runs the pipeline for `requested_digits`.  Device outputs stay resident. */
hipError_t gpupi_pi_run(BatchMPContext *ctx, uint32_t requested_digits, int prec_convention,
                        GpupiPiResult *result);
void gpupi_pi_free(GpupiPiResult *result);
/* Several pi jobs with one binary-splitting tree (no reference counterpart;
 * its CLI runs one job per process).  The tree over the largest job's groups
 * runs once: a job over N groups takes as its root row 0 of the level whose
 * nodes cover N groups, as soon as that level is reached, and runs its
 * remaining stages (sqrt, inverse, division, conversion) at once, smallest
 * jobs first.  results[i] equals gpupi_pi_run(requested_digits[i]) (free
 * each with gpupi_pi_free); elapsed_ms runs from the call's start to that
 * job's end.  On an error every result is released and zeroed.  In the
 * plain pipeline form the jobs run one after the other. */
hipError_t gpupi_pi_run_batch(BatchMPContext *ctx, const uint32_t *requested_digits, int count, int prec_convention,
                              GpupiPiResult *results);

/* How the pipeline issues its work (no reference counterpart).
 *   GPUPI_FORM_BATCHED (default): the 17-term group stage as one fused
 *     kernel; products sharing an operand transform it once; the P, Q, R
 *     compactions of a tree level share one host round trip; buffers come
 *     from the context's block cache.
 *   GPUPI_FORM_PLAIN: the reference's form (pi_bs_gpu.cu:503-792,
 *     :1346-1522, convert_decimal.cu:361-412): one launch sequence per
 *     operation, a synchronizing compaction after every product and sum the
 *     reference compacts, the group stage op by op, and buffers allocated
 *     with hipMalloc and freed with hipFree.
 * Both forms compute the same integers (RET limbs and digits are
 * identical); only the time differs. */
enum { GPUPI_FORM_BATCHED = 0, GPUPI_FORM_PLAIN = 1 };
hipError_t gpupi_set_pipeline_form(BatchMPContext *ctx, int form);

/* Stage entry points (batch-1 outputs are created arrays; release them). */
hipError_t gpupi_binary_split(BatchMPContext *ctx, uint32_t groups, BatchMPArray *Q, BatchMPArray *R);
hipError_t gpupi_sqrt_10005(BatchMPContext *ctx, int target_digits, BatchMPArray *P, BatchMPArray *Q);
hipError_t gpupi_recursive_inverse(BatchMPContext *ctx, BatchMPArray x, uint64_t L, BatchMPArray *y);
hipError_t gpupi_pi_fractional(BatchMPContext *ctx, int target_digits, BatchMPArray *ret,
                               uint32_t *groups, uint64_t *inverse_bits);
/* Leaf terms (pi_bs_gpu.cu:36-131): head rows 17t+16, body rows t + t/16. */
hipError_t gpupi_construct_leaf_17(BatchMPContext *ctx, int is_head, uint32_t *P, uint32_t *Q,
                                   uint32_t *R, uint32_t count);

/* Instrumentation (no reference counterpart).  While enabled, every NTT
 * kernel launch is bracketed by HIP events on its stream; query aggregates
 * per kernel: launches, summed event time and summed algorithmic bytes.  A
 * launch's algorithmic bytes count each residue pair it must read or write
 * once (16 B) and each input or output limb once (4 B): the minimum traffic
 * of that pass.  query synchronizes on the recorded events and returns the
 * number of entries written. */
typedef struct GpupiKernelStat {
    char name[48];
    uint64_t launches;
    double total_ms;
    double total_bytes;
} GpupiKernelStat;
void gpupi_kernel_stats_reset(int enable);
int gpupi_kernel_stats_query(GpupiKernelStat *out, int max_entries);
/* The same records split by transform size: one entry per (kernel, log2 K)
 * pair, named "<kernel>@2^<log2 K>". */
int gpupi_kernel_stats_query_sizes(GpupiKernelStat *out, int max_entries);

/* NTT plan tuning (no reference counterpart; the reference hard-codes its
 * pass schedules, batch_mul_ntt.cu:1810-1915).  A transform of 2^k points
 * runs as one local level (k <= 12), a column level and a contiguous level
 * (13 <= k <= 21), or two column levels and a contiguous level (k >= 22).
 * set_split picks the column sizes for one k: log2_m0 for the first level
 * and, with three levels, log2_m1 for the second (0 = the rest after a
 * 4096-point contiguous level).  Column levels take 2^5..2^9 points, the
 * contiguous level 2^6..2^12.  log2_m0 = 0 restores the built-in choice.
 * Returns hipErrorInvalidValue for a split outside those limits.  Results
 * are bit-identical for every split; only speed changes.  The workspace
 * size does not depend on the split.  get_split writes the level sizes
 * (log2) in use for k to log2_m[0..3) and returns the number of levels
 * (0 for an unsupported k). */
hipError_t gpupi_ntt_set_split(int log2_k, int log2_m0, int log2_m1);
int gpupi_ntt_get_split(int log2_k, int *log2_m);
/* NTT launch shape (no reference counterpart; the reference fixes its block
 * shapes, batch_mul_ntt.cu:1810-1915), process-wide, taking effect from the
 * next multiply on:
 *   pair_mask       column level kinds that run two neighbouring 4096-element
 *                   tiles per 512-thread block (levels of 2^7 points or more):
 *                   bit 0 first forward level, bit 1 second forward level,
 *                   bit 2 second inverse level, bit 3 first inverse level
 *                   (built-in 5);
 *   xcd_chunk_log2  consecutive tiles each of the 8 XCDs takes in turn, as a
 *                   log2 in 0..8 (0: the plain round-robin order; built-in 5).
 * hipErrorInvalidValue outside those ranges.  Results are bit-identical for
 * every shape; only speed changes.  gpupi_amd.tune --launch times them. */
hipError_t gpupi_ntt_set_launch(uint32_t pair_mask, int xcd_chunk_log2);
void gpupi_ntt_get_launch(uint32_t *pair_mask, int *xcd_chunk_log2);

/* The 17-term group stage on its own (pi_bs_gpu.cu:555-779) for `groups`
 * groups (a power of two), in three forms that leave the same P, Q, R rows
 * (one per group, compacted):
 *   GPUPI_LEAF_PLAIN    the reference's per-op form: every product is
 *                       followed by a synchronizing compaction;
 *   GPUPI_LEAF_BATCHED  lengths fixed in advance (allocation bounds), all
 *                       launches queued without a host round trip, one
 *                       compaction per array at the end;
 *   GPUPI_LEAF_GRAPH    the batched form captured once as a hipGraph on the
 *                       stage's own stream (with its own scratch) and
 *                       replayed by every run;
 *   GPUPI_LEAF_FUSED    one kernel: each wave builds one group in LDS
 *                       (leaf terms, levels, head merge, exact division).
 * No counterpart in the reference (its tree always runs the plain form).
 * run returns views of P, Q, R that stay valid until the next run or
 * destroy.  A GRAPH stage must not outlive its context. */
enum { GPUPI_LEAF_PLAIN = 0, GPUPI_LEAF_BATCHED = 1, GPUPI_LEAF_GRAPH = 2, GPUPI_LEAF_FUSED = 3 };
typedef struct GpupiLeafStage GpupiLeafStage;
hipError_t gpupi_leaf_stage_create(BatchMPContext *ctx, uint32_t groups, int mode, GpupiLeafStage **out);
hipError_t gpupi_leaf_stage_run(GpupiLeafStage *stage, BatchMPArray *P, BatchMPArray *Q, BatchMPArray *R);
void gpupi_leaf_stage_destroy(GpupiLeafStage *stage);
/* Block shape of the fused group-stage kernel (no reference counterpart),
 * process-wide, from the next launch on: groups per block, one wave each
 * (1, 2, 4, 8 or 16; built-in 4; else hipErrorInvalidValue).  Rows are
 * identical for every shape.  gpupi_amd.tune --launch times them. */
hipError_t gpupi_group_stage_set_shape(uint32_t groups_per_block);
uint32_t gpupi_group_stage_get_shape(void);

/* Per-stage profile of the pipeline (the reference's StageProfiler,
 * pi_bs_gpu.cu:133-489, behind --profile-stages).  When enabled, every
 * multiply, add/sub, shift, compaction, exact division, leaf construction
 * and decimal stage run through the context is bracketed by HIP events on
 * the context stream and synchronized, and one entry is kept per stage;
 * `kind` is one of leaf, mul, mul_small, shift, add, sub, compact,
 * exactdiv, decimal_powers, decimal_convert.  enable(ctx, 1) clears the
 * entries; query copies up to max_entries of them and returns how many are
 * held (or -1 for a null context). */
typedef struct GpupiStageEntry {
    char kind[16];
    char label[24];
    uint32_t n1;
    uint32_t batch_size;
    uint32_t lhs_length;
    uint32_t rhs_length;
    float elapsed_ms;
    int uses_ntt;
} GpupiStageEntry;
hipError_t gpupi_stage_profile_enable(BatchMPContext *ctx, int enable);
int gpupi_stage_profile_query(const BatchMPContext *ctx, GpupiStageEntry *out, int max_entries);

/* Device blocks the context keeps for the pipeline (no reference
 * counterpart; the reference allocates and frees every buffer, and each
 * free waits for the whole device).  The pipeline's buffers take blocks from
 * the context's cache and hand them back, so a repeated run allocates
 * nothing beyond its two result buffers.  query writes the bytes cached now
 * and the number of device allocations the cache has made; trim frees the
 * cached blocks (batch_mp_destroy does too).  Arrays returned by the stage
 * entry points are ordinary device allocations: release them as before. */
hipError_t gpupi_cache_query(const BatchMPContext *ctx, size_t *cached_bytes, uint64_t *allocations);
hipError_t gpupi_cache_trim(BatchMPContext *ctx);
/* Attaches a stream-ordered memory pool (hipMemPoolCreate) to the context:
 * from then on its pipeline buffers, NTT workspace and pi results come from
 * that pool with hipMallocFromPoolAsync on the context's stream and go back
 * with hipFreeAsync (no hipMalloc / hipFree, which wait for the whole
 * device); results of gpupi_pi_run are then freed with hipFreeAsync rather
 * than gpupi_pi_free.  Attach before the first allocation; null detaches. */
hipError_t gpupi_set_memory_pool(BatchMPContext *ctx, hipMemPool_t pool);

/* Sizing helpers (pi_bs_gpu.cu:924-944) */
uint32_t gpupi_round_up_decimal_digits(uint32_t requested, uint32_t *leaf_digits);
uint32_t gpupi_prec_limbs_for_digits(uint32_t digits);
/* 17-term groups of the binary-splitting tree for a precision of
 * `prec_digits` digits (pi_bs_gpu.cu:1537-1540). */
uint32_t gpupi_pi_groups_for_digits(uint32_t prec_digits);

/* Measurement helper (no reference counterpart): the current device's
 * compute units, peak engine clock (kHz) and architecture name, for the
 * integer-VALU roofline of the benchmark. */
hipError_t gpupi_device_info(int *compute_units, int *clock_khz, char *arch, int arch_len);

/* Calibration (no reference counterpart): the measured peak of back-to-back
 * v_mad_u64_u32 (32x32->64 multiply-adds) over the whole device, the
 * integer roofline of the schoolbook and group-stage kernels; and a plain
 * 16-byte copy kernel of `bytes` (a multiple of 16) whose HBM counters scale
 * the NTT kernels' FETCH_SIZE / WRITE_SIZE to bytes. */
hipError_t gpupi_calib_mac_peak(double *macs_per_second);
hipError_t gpupi_calib_copy(void *dst, const void *src, size_t bytes, hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif
