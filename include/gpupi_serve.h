/* Serving loop: pi-digit jobs and raw big-integer multiply jobs of mixed
 * sizes, admitted against an HBM budget as they arrive and run with their
 * work overlapped on one GPU.  No counterpart in the reference (its CLI runs
 * one pi job per process).
 *
 *   - Admission is first come, first served against the budget: a job whose
 *     estimated device footprint does not fit waits, and so does everything
 *     behind it (no starvation of large jobs); a pi job that fits but finds
 *     every pi worker busy lets later jobs go ahead.  A job larger than the
 *     whole budget is refused at submission.
 *   - Pi jobs run on `pi_workers` host threads, each with its own context and
 *     stream, so several pipelines are in flight on the device at once.  A
 *     worker starting a pi job also takes pi jobs not yet admitted because
 *     every worker was busy whose binary-splitting tree has the same number
 *     of groups: in arrival order up to the first job that does not fit the
 *     budget, at most its share of them (their count over pi_workers) and
 *     pi_batch_max in all.  They share one tree (gpupi_pi_run_batch) and one
 *     worker, and report batch_size.
 *   - Multiply jobs are batched continuously: the multiply lane takes every
 *     admitted multiply job (up to mul_batch_max) into one shared ragged
 *     product launch; jobs arriving meanwhile form the next batch.
 *   - Results are bit-identical to running each job alone: every job's
 *     arithmetic is exact and independent of what it shares launches with.
 *
 * Host operand and output buffers passed to submit must stay valid until the
 * job's result has been returned.  All functions are thread-safe. */
#ifndef GPUPI_SERVE_H
#define GPUPI_SERVE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { GPUPI_JOB_PI = 0, GPUPI_JOB_MUL = 1 };

typedef struct GpupiServeConfig {
    uint64_t hbm_budget_bytes; /* admission budget; 0 = 200 GB */
    uint32_t pi_workers;       /* pi jobs in flight at once; 0 = 4 */
    uint32_t mul_batch_max;    /* multiply jobs per shared launch; 0 = 512 */
    uint32_t pi_batch_max;     /* pi jobs of one tree size sharing a tree; 0 = 8, 1 = none */
} GpupiServeConfig;

typedef struct GpupiJobResult {
    int64_t id;
    int32_t kind;
    int32_t status;          /* hipError_t of the job */
    double submit_ms;        /* server clock: ms since creation */
    double start_ms;         /* admitted and started */
    double finish_ms;        /* result in the caller's buffer */
    uint64_t reserved_bytes; /* admission estimate it held */
    uint32_t batch_size;     /* jobs in its shared launch (multiply) or tree (pi) */
    uint32_t ret_length;     /* pi jobs: RET limbs */
    uint32_t ret_hash;       /* pi jobs: RET mod 1e9+7 (pi_bs_gpu.cu:852-858) */
    uint32_t dec_hash;       /* pi jobs: hash of all digits (pi_bs_gpu.cu:860-866) */
} GpupiJobResult;

typedef struct GpupiServer GpupiServer;

/* Null on failure (no GPU, or contexts could not be set up). */
GpupiServer *gpupi_server_create(const GpupiServeConfig *cfg);
/* Returns the job id (>= 0) or -1 if the job can never be admitted or the
 * arguments are invalid.  digits_out: optional host buffer of `digits` bytes
 * (the digits after "3."); C_out: host buffer of L_a + L_b limbs. */
int64_t gpupi_server_submit_pi(GpupiServer *s, uint32_t digits, char *digits_out);
int64_t gpupi_server_submit_mul(GpupiServer *s, const uint32_t *A, uint32_t L_a, const uint32_t *B, uint32_t L_b,
                                uint32_t *C_out);
/* Moves up to max results of finished jobs into out; with wait != 0 blocks
 * until at least one is available or nothing is pending.  Returns the count. */
int gpupi_server_results(GpupiServer *s, GpupiJobResult *out, int max, int wait);
/* Blocks until every submitted job has finished. */
void gpupi_server_drain(GpupiServer *s);
void gpupi_server_destroy(GpupiServer *s);
/* Peak device bytes the server's memory pool has handed out (used) and held
 * (reserved) since it was created.  Every worker takes its device memory
 * (pipeline buffers, NTT workspace, operand arenas) from this one pool in
 * stream order; memory above the HBM budget goes back to the device. */
hipError_t gpupi_server_memory(GpupiServer *s, uint64_t *used_high, uint64_t *reserved_high);

/* Device footprint estimate used for admission (bytes). */
uint64_t gpupi_serve_job_bytes(int kind, uint32_t digits_or_la, uint32_t lb);

/* The admission rule on its own, for a queue of jobs in arrival order:
 * bytes[i], kind[i]; `in_use` bytes already reserved and `free_pi_workers`
 * idle pi workers.  Writes the indices admitted now (in order) to out and
 * returns their count.  Host-only: no device work. */
int gpupi_serve_admit(const uint64_t *bytes, const int32_t *kind, int n, uint64_t budget, uint64_t in_use,
                      int free_pi_workers, int *out);

#ifdef __cplusplus
}
#endif

#endif
