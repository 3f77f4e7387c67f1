// gpupi-mi355x internal header: device helpers shared by the HIP sources.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <map>

#include "batch_arith.h"
#include "batch_kernels.h"
#include "ragged.hpp"

namespace gpupi {

constexpr int kWave = 64;  // CDNA wavefront width

#define GPUPI_RETURN_IF(err)                     \
    do {                                         \
        hipError_t _e = (err);                   \
        if (_e != hipSuccess) return _e;         \
    } while (0)

inline uint32_t ceil_div(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }
inline uint64_t ceil_div64(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint64_t ballot64(bool p) { return (uint64_t)__ballot(p); }

// Resolve a 64-lane carry (or borrow) chain.  Bit l of G: lane l produces a
// carry with no carry-in; bit l of P: lane l passes an incoming carry on
// (G and P never share a bit).  Treating (G|P) and G as two 64-bit numbers,
// the carry into lane l of their sum is exactly the carry lane l receives.
// Returns the carry-in mask; *cout receives the carry out of lane 63.
__device__ __forceinline__ uint64_t resolve_carries(uint64_t G, uint64_t P, uint32_t cin,
                                                    uint32_t *cout) {
    const uint64_t X = G | P;
    const uint64_t s1 = X + G;
    const uint32_t c1 = s1 < X;
    const uint64_t s2 = s1 + (uint64_t)cin;
    const uint32_t c2 = s2 < s1;
    *cout = c1 | c2;
    return s2 ^ X ^ G;
}

// Internal stream-aware launchers (defined across the .hip sources).
// `flags` is scratch of at least carry_flags_bytes(N, L_c) bytes.
size_t carry_flags_bytes(uint32_t N, uint32_t L_c);

enum class AddSubKind : int { Add, Sub, AddSmall, SubSmall, MulSmall, ShiftAdd, ShiftSub, SpillAdd };

struct AddSubArgs {
    AddSubKind kind;
    const uint32_t *A;
    const uint32_t *B;
    uint32_t *C;
    uint32_t scalar;   // AddSmall/SubSmall/MulSmall operand
    uint32_t shift;    // ShiftAdd/ShiftSub bit shift of A
    uint32_t N, L_a, L_b, L_c;
    uint32_t stride_A, stride_B, stride_C;
};

hipError_t launch_addsub(const AddSubArgs &args, uint8_t *flags, hipStream_t stream);

// C (limbs) += spill words: for each run r (run_len limbs), the 64-bit
// value spill[r] is added at limb (r + 1) * run_len of the same row
// (run_len a power of two).
hipError_t launch_add_spills(uint32_t *C, const uint64_t *spill, uint32_t N, uint32_t L_c,
                             uint32_t stride_C, uint32_t run_len, uint32_t runs_per_row,
                             uint8_t *flags, hipStream_t stream);

hipError_t launch_shift_bits(const uint32_t *A, uint32_t *B, uint32_t N, uint32_t L_in,
                             uint32_t L_out, uint32_t stride_in, uint32_t stride_out,
                             int64_t shift_bits, hipStream_t stream);

hipError_t launch_exactdiv_small(uint32_t *A, uint32_t B, uint32_t N, uint32_t L,
                                 uint32_t stride_A, hipStream_t stream);

// Writes max(limb length) (or bit length) over the batch into *d_result,
// which the caller zeroed.
hipError_t launch_length_max(const uint32_t *A, uint32_t N, uint32_t L, uint32_t stride_A,
                             bool bits, uint32_t *d_result, hipStream_t stream);

hipError_t launch_mul_naive(const uint32_t *A, const uint32_t *B, uint32_t *C, uint32_t N,
                            uint32_t L_a, uint32_t L_b, uint32_t stride_A, uint32_t stride_B,
                            uint32_t stride_C, hipStream_t stream);
bool mul_naive_supports(uint32_t L_a, uint32_t L_b);

hipError_t launch_mul_ntt(const uint32_t *A, const uint32_t *B, uint32_t *C, void *workspace,
                          size_t workspace_bytes, const NTTPrecomputedTables &tables, uint32_t N,
                          uint32_t L_a, uint32_t L_b, uint32_t stride_A, uint32_t stride_B,
                          uint32_t stride_C, hipStream_t stream);
size_t ntt_workspace_bytes(uint32_t N, uint32_t L_a, uint32_t L_b);
// The same products with at most `workspace_bytes` of workspace: rows run in
// chunks that fit it (hipErrorInvalidValue when one row does not); records
// no kernel statistics (creates no events).
hipError_t launch_mul_ntt_bounded(const uint32_t *A, const uint32_t *B, uint32_t *C, void *workspace,
                                  size_t workspace_bytes, const NTTPrecomputedTables &tables, uint32_t N,
                                  uint32_t L_a, uint32_t L_b, uint32_t stride_A, uint32_t stride_B,
                                  uint32_t stride_C, hipStream_t stream);
int ntt_transform_log2(uint32_t L_a, uint32_t L_b);
// The fused 17-term group stage: N rows each of P (51 limbs), Q (68), R (83).
hipError_t launch_group_stage(uint32_t N, uint32_t *P, uint32_t *Q, uint32_t *R, hipStream_t stream);
// Rows a ragged launch is filled to (one launch chunk of 2^log2K-point
// rows); the pieces of one product are never split across launches, so a
// launch holding one product's pieces may have more.
uint32_t ntt_ragged_max_rows(int log2K);
// C = L * S in m >= 2 pieces of s limbs (piece_row; s >= LS when m > 2,
// every piece's product above one tile), assembled into C.  The rows are
// generated on the device; the workspace holds them and the NTT scratch.
size_t ntt_pieces_workspace_bytes(uint32_t LL, uint32_t LS, uint32_t m, uint32_t s);
hipError_t launch_mul_ntt_pieces(const uint32_t *L, uint32_t LL, const uint32_t *S, uint32_t LS, uint32_t *C,
                                 uint32_t m, uint32_t s, void *workspace, size_t workspace_bytes,
                                 const NTTPrecomputedTables &tables, hipStream_t stream);
// The assemblies of a ragged launch from its staged rows (K limbs each):
// n assemblies, `chunks` chunks in all; `flags` holds one byte per chunk.
hipError_t launch_ragged_assemble(const uint32_t *staging, uint64_t K, const RaggedAssembly *as, uint32_t n,
                                  uint32_t chunks, uint8_t *flags, hipStream_t stream);
size_t ntt_ragged_workspace_bytes(uint32_t N, uint32_t nb, uint32_t chunks, int log2K);
// N rows over nb B operands and n_asm assemblies of `chunks` chunks in all,
// as device arrays with host copies: every row's L_a + L_b <= 2^log2K and
// b < nb; rows without C (assembled) only above one tile (2^log2K > 4096).
hipError_t launch_mul_ntt_ragged(const RaggedRow *rows, const RaggedRow *host_rows, uint32_t N, const RaggedB *bset,
                                 const RaggedB *host_bset, uint32_t nb, const RaggedAssembly *asms,
                                 const RaggedAssembly *host_asms, uint32_t n_asm, uint32_t chunks, int log2K,
                                 void *workspace, size_t workspace_bytes, const NTTPrecomputedTables &tables,
                                 hipStream_t stream);
bool ntt_supports(uint32_t L_a, uint32_t L_b);
// Two NTT products sharing operand B (C1 = A1 B over N1 rows, C2 = A2 B over
// N2 rows; B has max(N1, N2) rows, or one row at stride 0): B is
// transformed once.  Supported when both products have the same transform
// size above one tile.
bool ntt_shared_supports(uint32_t L1, uint32_t L2, uint32_t Lb);
size_t ntt_shared_workspace_bytes(uint32_t L1, uint32_t N1, uint32_t L2, uint32_t N2, uint32_t Lb, uint32_t sB);
hipError_t launch_mul_ntt_shared(const uint32_t *A1, uint32_t L1, uint32_t s1, uint32_t N1, uint32_t *C1,
                                 uint32_t sC1, const uint32_t *A2, uint32_t L2, uint32_t s2, uint32_t N2,
                                 uint32_t *C2, uint32_t sC2, const uint32_t *B, uint32_t Lb, uint32_t sB,
                                 void *workspace, size_t workspace_bytes, const NTTPrecomputedTables &tables,
                                 hipStream_t stream);
hipError_t ntt_tables_init(NTTPrecomputedTables *tables);
void ntt_tables_release(NTTPrecomputedTables *tables);

hipError_t launch_decimal_leaf(const uint32_t *A, char *B, uint32_t N, uint32_t L_limb,
                               uint32_t stride_A, uint32_t L_dec, uint32_t stride_B,
                               bool natural_order, hipStream_t stream);

hipError_t launch_fill_u32(uint32_t *p, uint64_t count, uint32_t value, hipStream_t stream);

}  // namespace gpupi

namespace gpupi {

// Device blocks kept by a context for reuse by the pi pipeline's buffers.
// Every block serves one stream at a time, so a block handed back can serve
// the next request at once: stream order runs the new user's kernels after
// the old user's.  A request from another stream first waits for the device.
// Only trim() frees blocks (hipFree waits for the whole device, which would
// stall the host between stages).
//
// With a memory pool attached (set_pool: the serving loop's one pool for all
// its workers) blocks come from and go back to that pool in stream order
// (hipMallocFromPoolAsync / hipFreeAsync): nothing is kept here and nothing
// waits for the device; the pool reuses freed memory across streams.
class BlockCache {
  public:
    // A block of at least `bytes` for work on `stream`; *got is its size.
    hipError_t acquire(hipStream_t stream, size_t bytes, void **ptr, size_t *got);
    // Hands a block back after the work queued on `stream`; `bytes` may be
    // less than its size, never more.
    void give_back(void *ptr, size_t bytes, hipStream_t stream);
    hipError_t trim();
    void set_pool(hipMemPool_t pool) { pool_ = pool; }
    hipMemPool_t pool() const { return pool_; }
    size_t cached_bytes() const { return cached_; }
    uint64_t allocations() const { return allocations_; }
    ~BlockCache() { (void)trim(); }

  private:
    std::multimap<size_t, void *> free_;  // size -> block
    hipStream_t stream_ = nullptr;
    bool bound_ = false;
    size_t cached_ = 0;
    uint64_t allocations_ = 0;
    hipMemPool_t pool_ = nullptr;
};

// Device memory of a context: from its memory pool in stream order when one
// is attached, else hipMalloc / hipFree.
hipError_t ctx_alloc(BatchMPContext *ctx, void **ptr, size_t bytes);
// While set on this thread, NTT launches record no kernel statistics (no
// events: work being captured into a graph).
void ntt_stats_suspend(bool suspend);
void ctx_free(BatchMPContext *ctx, void *ptr);

constexpr int kScalarWords = 4;  // device words of BatchMPContext::d_scalar

// Compacts up to kScalarWords arrays with one host round trip: each
// array's length becomes its batch's limb length (batch_mp_array_compact).
hipError_t compact_arrays(BatchMPContext *ctx, BatchMPArray *const *arrays, int n);

}  // namespace gpupi

// Context internals shared between arith.hip and the pipeline sources.
// Per-stage event timings of the pi pipeline (defined with the pipeline).
struct StageProfiler;
void stage_profiler_free(StageProfiler *p);
// The bottom levels of the binary-splitting tree as a hipGraph (defined with
// the pipeline).
struct TreeGraph;
void tree_graph_free(TreeGraph *g);

struct BatchMPContext {
    NTTPrecomputedTables ntt_tables;
    uint32_t *workspace;
    size_t workspace_size_bytes;
    uint32_t *d_scalar;       // kScalarWords device words for length reductions
    uint32_t *h_scalar;       // pinned host mirror
    hipStream_t stream;
    int ntt_min_limbs;        // schoolbook -> NTT crossover on L_a + L_b
    StageProfiler *profiler;  // null unless stage profiling is on
    gpupi::BlockCache *cache; // blocks of the pipeline's buffers
    int pipeline_form;        // GPUPI_FORM_BATCHED or GPUPI_FORM_PLAIN
    TreeGraph *tree_graph;    // captured tree levels (null until first used)
};
