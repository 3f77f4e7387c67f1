// Leaf terms of the Chudnovsky series (pi_bs_gpu.cu:17-131), shared by the
// leaf kernel and the fused 17-term group kernel.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gpupi {

__device__ __forceinline__ void reduce_pair(uint32_t &a, uint32_t &b, uint32_t d) {
    if (a % d == 0u && b % d == 0u) {
        a /= d;
        b /= d;
    }
}

__device__ __forceinline__ void reduce_3_5_23_29(uint32_t &a, uint32_t &b) {
    reduce_pair(a, b, 3u);
    reduce_pair(a, b, 5u);
    reduce_pair(a, b, 23u);
    reduce_pair(a, b, 29u);
}

template <int W>
__device__ __forceinline__ void mul_word(uint32_t (&v)[W], uint32_t m) {
    uint32_t carry = 0;
#pragma unroll
    for (int j = 0; j < W; ++j) {
        const uint64_t t = (uint64_t)v[j] * m + carry;
        v[j] = (uint32_t)t;
        carry = (uint32_t)(t >> 32);
    }
}

// Term i: p = (6i+1)(2i+1)(6i+5), q = (i+1)^3 * 10005^2 * 3335 (= (i+1)^3
// 640320^3 / 24 / 2^15), r = (13591409 + 545140134 i) * q / 3335, with the
// common factors of p and q cancelled as in pi_bs_gpu.cu:17-74.  Writes
// P (3 limbs), Q (4) and R (6).
__device__ __forceinline__ void leaf_term(int64_t i, uint32_t *po, uint32_t *qo, uint32_t *ro) {
    uint32_t p1 = (uint32_t)(6 * i + 1), p2 = (uint32_t)(2 * i + 1), p3 = (uint32_t)(6 * i + 5);
    uint32_t q1 = (uint32_t)(i + 1), q2 = q1, q3 = q1;
    uint32_t q4 = 10005u, q5 = 10005u;
    const uint32_t q6 = 3335u;
    const uint64_t r1 = (uint64_t)(545140134ll * i + 13591409ll);
    reduce_pair(p1, q1, 5u);
    reduce_pair(p1, q2, 5u);
    reduce_pair(p1, q3, 5u);
    reduce_3_5_23_29(p1, q4);
    reduce_3_5_23_29(p1, q5);
    reduce_3_5_23_29(p2, q4);
    reduce_3_5_23_29(p2, q5);
    reduce_3_5_23_29(p3, q4);
    reduce_3_5_23_29(p3, q5);
    uint32_t pv[3] = {p1, 0, 0};
    mul_word(pv, p2);
    mul_word(pv, p3);
    uint32_t qv[4] = {q1, 0, 0, 0};
    mul_word(qv, q2);
    mul_word(qv, q3);
    mul_word(qv, q4);
    mul_word(qv, q5);
    mul_word(qv, q6);
    uint32_t rv[6] = {(uint32_t)r1, (uint32_t)(r1 >> 32), 0, 0, 0, 0};
    mul_word(rv, q1);
    mul_word(rv, q2);
    mul_word(rv, q3);
    mul_word(rv, q4);
    mul_word(rv, q5);
#pragma unroll
    for (int j = 0; j < 3; ++j) po[j] = pv[j];
#pragma unroll
    for (int j = 0; j < 4; ++j) qo[j] = qv[j];
#pragma unroll
    for (int j = 0; j < 6; ++j) ro[j] = rv[j];
}

}  // namespace gpupi
