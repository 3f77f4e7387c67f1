// Ragged NTT launches: the row, B-set and assembly descriptors and the
// index rules of products run in pieces, as plain host/device code so that
// the assembly can be checked on the CPU.
#pragma once

#include <stddef.h>
#include <stdint.h>

#ifndef GPUPI_HD
#if defined(__HIPCC__)
#define GPUPI_HD __host__ __device__ __forceinline__
#else
#define GPUPI_HD inline
#endif
#endif

namespace gpupi {

// One product row of a ragged NTT launch: A (L_a limbs) times row b of the
// launch's set of B operands (B, L_b limbs), written to C as L_a + L_b
// limbs, or left staged for a RaggedAssembly when C is null.  Rows sharing
// a B operand (the pieces of one long product) have it transformed once.
struct RaggedRow {
    const uint32_t *A;
    const uint32_t *B;
    uint32_t *C;
    uint32_t L_a, L_b, b, pad;
};
struct RaggedB {
    const uint32_t *B;
    uint64_t L;
};
// A product L * S (c = |L| + |S| limbs) run in m pieces of s limbs of L (the
// last one shorter) against S: its pieces are the launch's rows row0 ..
// row0 + m - 1, and C word x is the sum of piece floor(x / s)'s product
// word x - floor(x / s) s and, below it, piece floor(x / s) - 1's word
// x - (floor(x / s) - 1) s (the last piece takes every x from (m - 1) s
// on), with carries.  Assembled by 1024-limb chunks, numbered in the launch
// from chunk0.
struct RaggedAssembly {
    uint32_t *C;
    uint32_t c, s, S, m, row0, chunk0;
};
constexpr uint32_t kAssembleChunk = 1024;
// Row i of a product L * S (LL and LS limbs) run in m pieces of s limbs of
// L (the last one LL - (m - 1) s limbs), all against S, row b of the
// launch's B set; its product stays staged for the assembly.
GPUPI_HD RaggedRow piece_row(uint32_t i, const uint32_t *L, uint32_t LL, const uint32_t *S, uint32_t LS, uint32_t m,
                             uint32_t s, uint32_t b) {
    const uint32_t len = i + 1 < m ? s : LL - (m - 1) * s;
    return RaggedRow{L + (size_t)i * s, S, nullptr, len, LS, b, 0};
}
// The staged words that limb x of an assembly sums: piece *i's word *off
// and, when *below, piece *i - 1's word *off + s.
GPUPI_HD void assembly_source(const RaggedAssembly &a, uint32_t x, uint32_t *i, uint32_t *off, bool *below) {
    const uint32_t q = x / a.s;
    *i = q < a.m - 1 ? q : a.m - 1;
    *off = x - *i * a.s;
    *below = *i > 0 && *off < a.S;
}
// Chunks of an assembly of c limbs.
GPUPI_HD uint32_t assembly_chunks(uint32_t c) { return (c + kAssembleChunk - 1) / kAssembleChunk; }

}  // namespace gpupi
