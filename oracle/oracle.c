/*
 * CPU oracle for the gpupi pi pipeline — TEST INFRASTRUCTURE ONLY.
 *
 * Restates, on GMP (libgmp.so.10, declared by hand in gmp_decl.h), the
 * reference algorithm of fanhqme2/gpupi op for op so that every value the HIP
 * library produces can be checked bit for bit.  Each function names the
 * reference lines it follows.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load this file's library; the product never
 * does.
 *
 * Parity pins (see tests/test_oracle_*.py):
 *   - batch ops: Python big-integer closed forms on random and edge inputs;
 *   - o_leaf17: exact rational identities P/Q = p/q, R/Q = a/3335 per term;
 *   - o_binary_split: exact divisibility of each 17-term block, and the
 *     series value against mpmath;
 *   - o_sqrt_10005: Pell identity p^2 - 10005 q^2 = 1 and the closed form;
 *   - o_recursive_inverse: |y - 2^L/x| bound against exact division;
 *   - o_pi_digits: digits against mpmath's pi, RET's top limbs against pi's
 *     binary expansion;
 *   - convert: the reference's own known-answer cases (convert_decimal_test.cu
 *     generator + exact decimal expansion).
 * The exact low limbs of RET (truncation and inverse rounding) are pinned
 * only through the 1e9 README hash, which the GPU run checks.
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "gmp_decl.h"

#if !defined(__x86_64__) && !defined(__aarch64__)
#error "oracle assumes a little-endian 64-bit host"
#endif

/* ------------------------------------------------------------------ */
/* helpers                                                              */
/* ------------------------------------------------------------------ */

static inline uint32_t *row_ptr(OArray a, uint32_t i) {
    return a.data + (size_t)i * a.stride;
}

/* batch_arith.cu:89-100 */
static int valid_array(OArray a) {
    if (a.batch_size == 0) return a.length <= a.stride;
    if (a.data == NULL) return 0;
    if (a.stride == 0) return 1;
    return a.length <= a.stride;
}

static void *xmalloc(size_t bytes) {
    void *p = malloc(bytes ? bytes : 8);
    if (!p) abort();
    return p;
}

static void *xcalloc(size_t words) {
    void *p = calloc(words ? words : 2, 4);
    if (!p) abort();
    return p;
}

/* Copy `len` 32-bit words into a zero-extended buffer of n64 64-bit limbs. */
static void load_words(mp_limb_t *dst, size_t n64, const uint32_t *src, uint32_t len) {
    memset(dst, 0, n64 * sizeof(mp_limb_t));
    size_t w = len;
    if (w > 2 * n64) w = 2 * n64;
    if (w) memcpy(dst, src, w * 4);
}

/* Store the low `len` words of a limb buffer (zero-fill above it). */
static void store_words(uint32_t *dst, uint32_t len, const mp_limb_t *src, size_t n64) {
    size_t w = len;
    if (w > 2 * n64) w = 2 * n64;
    if (w) memmove(dst, src, w * 4);
    for (size_t i = w; i < len; ++i) dst[i] = 0;
}

/* dst (n words) = (src << shift) mod 2^(32 n) */
static void shl_words(uint32_t *dst, size_t n, const uint32_t *src, uint32_t len, uint64_t shift) {
    memset(dst, 0, n * 4);
    const uint64_t ws = shift >> 5;
    const uint32_t bs = (uint32_t)(shift & 31u);
    for (uint64_t i = 0; i < len; ++i) {
        const uint64_t j = i + ws;
        if (j >= n) break;
        const uint64_t v = (uint64_t)src[i] << bs;
        dst[j] |= (uint32_t)v;
        if (j + 1 < n) dst[j + 1] |= (uint32_t)(v >> 32);
    }
}

/* ------------------------------------------------------------------ */
/* batch ops (batch_arith.cu:170-424 and the per-op headers)           */
/* ------------------------------------------------------------------ */

/* Worker threads for the batched products (CPU-baseline runs only; the
 * tests keep one thread).  Rows are independent, so the result does not
 * depend on the thread count. */
static int g_threads = 1;

void o_set_threads(int threads) { g_threads = threads < 1 ? 1 : threads; }

static void mul_rows(OArray A, OArray B, OArray C, uint32_t r0, uint32_t r1) {
    const size_t na = (A.length + 1) / 2, nb = (B.length + 1) / 2;
    mp_limb_t *a = xmalloc((na + 1) * 8), *b = xmalloc((nb + 1) * 8);
    mp_limb_t *r = xmalloc((na + nb + 1) * 8);
    for (uint32_t i = r0; i < r1; ++i) {
        uint32_t *c = row_ptr(C, i);
        if (na == 0 || nb == 0) {
            memset(c, 0, (size_t)C.length * 4);
            continue;
        }
        load_words(a, na, row_ptr(A, i), A.length);
        load_words(b, nb, row_ptr(B, i), B.length);
        if (na >= nb) mpn_mul(r, a, (mp_size_t)na, b, (mp_size_t)nb);
        else mpn_mul(r, b, (mp_size_t)nb, a, (mp_size_t)na);
        store_words(c, C.length, r, na + nb);
    }
    free(a); free(b); free(r);
}

typedef struct {
    OArray A, B, C;
    uint32_t r0, r1;
} MulTask;

static void *mul_worker(void *p) {
    MulTask *t = p;
    mul_rows(t->A, t->B, t->C, t->r0, t->r1);
    return NULL;
}

/* batch_arith.cu:170-202: C = A * B, C.length == A.length + B.length */
int o_mul(OArray A, OArray B, OArray C) {
    if (!valid_array(A) || !valid_array(B) || !valid_array(C)) return -1;
    if (A.batch_size != B.batch_size || A.batch_size != C.batch_size) return -1;
    if (C.length != A.length + B.length) return -1;
    uint32_t T = (uint32_t)g_threads;
    if (T > A.batch_size) T = A.batch_size;
    if (T <= 1) {
        mul_rows(A, B, C, 0, A.batch_size);
        return 0;
    }
    pthread_t th[256];
    MulTask task[256];
    if (T > 256) T = 256;
    for (uint32_t t = 0; t < T; ++t) {
        task[t] = (MulTask){A, B, C, (uint32_t)((uint64_t)A.batch_size * t / T),
                            (uint32_t)((uint64_t)A.batch_size * (t + 1) / T)};
        if (t + 1 < T && pthread_create(&th[t], NULL, mul_worker, &task[t]) != 0) abort();
    }
    mul_rows(A, B, C, task[T - 1].r0, task[T - 1].r1);
    for (uint32_t t = 0; t + 1 < T; ++t) pthread_join(th[t], NULL);
    return 0;
}

/* batch_mul_small.h:7-15: C = A * b over L_a + 1 limbs, truncated/zero-filled */
int o_mul_small(OArray A, uint32_t bval, OArray C) {
    if (!valid_array(A) || !valid_array(C)) return -1;
    if (A.batch_size != C.batch_size) return -1;
    const size_t n = A.length / 2 + 1;
    mp_limb_t *a = xmalloc(n * 8), *r = xmalloc(n * 8);
    for (uint32_t i = 0; i < A.batch_size; ++i) {
        load_words(a, n, row_ptr(A, i), A.length);
        mpn_mul_1(r, a, (mp_size_t)n, (mp_limb_t)bval);
        store_words(row_ptr(C, i), C.length, r, n);
    }
    free(a); free(r);
    return 0;
}

/* batch_arith.cu:225-235 + batch_exactdiv_small.h: in-place exact A /= b */
int o_exactdiv_small(OArray A, uint32_t bval) {
    if (!valid_array(A)) return -1;
    if (bval == 0 || (bval & 1u) == 0 || A.length > 128u) return -1;
    const size_t n = (A.length + 1) / 2;
    if (n == 0) return 0;
    mp_limb_t *a = xmalloc(n * 8), *q = xmalloc(n * 8);
    int status = 0;
    for (uint32_t i = 0; i < A.batch_size; ++i) {
        uint32_t *row = row_ptr(A, i);
        load_words(a, n, row, A.length);
        if (mpn_mod_1(a, (mp_size_t)n, (mp_limb_t)bval) != 0) status = -2;
        mpn_divexact_1(q, a, (mp_size_t)n, (mp_limb_t)bval);
        store_words(row, A.length, q, n);
    }
    free(a); free(q);
    return status;
}

/* Generic (X op Y) mod 2^(32 L_c) where X = A << shift, Y = B (or a scalar). */
static int addsub_generic(OArray A, const uint32_t *bscalar, OArray B, uint64_t shift,
                          int subtract, OArray C) {
    uint64_t wa = A.length + (shift >> 5) + 1;
    uint64_t lb = bscalar ? 1u : B.length;
    uint64_t w = wa;
    if (lb > w) w = lb;
    if (C.length > w) w = C.length;
    w += 2;
    const size_t n64 = (size_t)((w + 1) / 2);
    uint32_t *xs = xcalloc(2 * n64);
    mp_limb_t *x = xmalloc(n64 * 8), *y = xmalloc(n64 * 8), *r = xmalloc(n64 * 8);
    for (uint32_t i = 0; i < A.batch_size; ++i) {
        shl_words(xs, 2 * n64, row_ptr(A, i), A.length, shift);
        memcpy(x, xs, n64 * 8);
        if (bscalar) load_words(y, n64, bscalar, 1);
        else load_words(y, n64, row_ptr(B, i), B.length);
        if (subtract) mpn_sub(r, x, (mp_size_t)n64, y, (mp_size_t)n64);
        else mpn_add(r, x, (mp_size_t)n64, y, (mp_size_t)n64);
        store_words(row_ptr(C, i), C.length, r, n64);
    }
    free(xs); free(x); free(y); free(r);
    return 0;
}

/* batch_add.h:6-20; in place allowed */
int o_add(OArray A, OArray B, OArray C) {
    if (!valid_array(A) || !valid_array(B) || !valid_array(C)) return -1;
    if (A.batch_size != B.batch_size || A.batch_size != C.batch_size) return -1;
    return addsub_generic(A, NULL, B, 0, 0, C);
}

/* batch_sub.h:6-16 */
int o_sub(OArray A, OArray B, OArray C) {
    if (!valid_array(A) || !valid_array(B) || !valid_array(C)) return -1;
    if (A.batch_size != B.batch_size || A.batch_size != C.batch_size) return -1;
    return addsub_generic(A, NULL, B, 0, 1, C);
}

/* batch_add_small.h:6-18 */
int o_add_small(OArray A, uint32_t bval, OArray C) {
    if (!valid_array(A) || !valid_array(C)) return -1;
    if (A.batch_size != C.batch_size) return -1;
    OArray none = {0};
    return addsub_generic(A, &bval, none, 0, 0, C);
}

/* batch_sub_small.h:6-16 */
int o_sub_small(OArray A, uint32_t bval, OArray C) {
    if (!valid_array(A) || !valid_array(C)) return -1;
    if (A.batch_size != C.batch_size) return -1;
    OArray none = {0};
    return addsub_generic(A, &bval, none, 0, 1, C);
}

/* batch_arith.cu:279-329; batch_shift_add.h:7-20: C must not alias A or B */
int o_shift_add(OArray A, OArray B, uint32_t shift, OArray C) {
    if (!valid_array(A) || !valid_array(B) || !valid_array(C)) return -1;
    if (A.batch_size != B.batch_size || A.batch_size != C.batch_size) return -1;
    if (C.data == A.data || C.data == B.data) return -1;
    return addsub_generic(A, NULL, B, shift, 0, C);
}

int o_shift_sub(OArray A, OArray B, uint32_t shift, OArray C) {
    if (!valid_array(A) || !valid_array(B) || !valid_array(C)) return -1;
    if (A.batch_size != B.batch_size || A.batch_size != C.batch_size) return -1;
    if (C.data == A.data || C.data == B.data) return -1;
    return addsub_generic(A, NULL, B, shift, 1, C);
}

/* batch_shift.h:5-15: B = A << s (s > 0) or A >> -s, truncated/zero-filled */
int o_shift_bits(OArray A, OArray B, int32_t sbits) {
    if (!valid_array(A) || !valid_array(B)) return -1;
    if (A.batch_size != B.batch_size) return -1;
    uint32_t *tmp = xcalloc((size_t)B.length + 1);
    for (uint32_t i = 0; i < A.batch_size; ++i) {
        const uint32_t *a = row_ptr(A, i);
        if (sbits >= 0) {
            shl_words(tmp, B.length, a, A.length, (uint64_t)sbits);
        } else {
            const uint64_t s = (uint64_t)(-(int64_t)sbits);
            const uint64_t ws = s >> 5;
            const uint32_t bs = (uint32_t)(s & 31u);
            for (uint64_t j = 0; j < B.length; ++j) {
                const uint64_t k = j + ws;
                uint64_t v = 0;
                if (k < A.length) v = a[k];
                if (k + 1 < A.length) v |= (uint64_t)a[k + 1] << 32;
                tmp[j] = (uint32_t)(v >> bs);
            }
        }
        memcpy(row_ptr(B, i), tmp, (size_t)B.length * 4);
    }
    free(tmp);
    return 0;
}

/* batch_bitlength.h:19-25 */
uint32_t o_limblength_max(OArray A) {
    uint32_t best = 0;
    for (uint32_t i = 0; i < A.batch_size; ++i) {
        const uint32_t *a = row_ptr(A, i);
        uint32_t l = A.length;
        while (l > best && a[l - 1] == 0) --l;
        if (l > best) best = l;
    }
    return best;
}

/* batch_bitlength.h:9-16 */
uint32_t o_bitlength_max(OArray A) {
    uint32_t best = 0;
    for (uint32_t i = 0; i < A.batch_size; ++i) {
        const uint32_t *a = row_ptr(A, i);
        uint32_t l = A.length;
        while (l > 0 && a[l - 1] == 0) --l;
        if (l == 0) continue;
        const uint32_t bits = (l - 1) * 32u + (32u - (uint32_t)__builtin_clz(a[l - 1]));
        if (bits > best) best = bits;
    }
    return best;
}

/* batch_decimal_naive.h:7-31: X / 2^(32 L_limb) -> L_dec digits, written
 * least significant first (B[L_dec-1] is the 10^-1 digit).  The digits are
 * those of floor(X * 10^L_dec / 2^(32 L_limb)). */
void o_decimal_naive(const uint32_t *A, char *B, uint32_t N, int L_limb,
                     uint32_t stride_A, int L_dec, uint32_t stride_B) {
    if (N == 0 || L_limb <= 0 || L_dec <= 0) return;
    uint32_t *w = xcalloc((size_t)L_limb);
    for (uint32_t r = 0; r < N; ++r) {
        memcpy(w, A + (size_t)r * stride_A, (size_t)L_limb * 4);
        char *out = B + (size_t)r * stride_B;
        for (int start = 0; start < L_dec; start += 9) {
            uint64_t carry = 0;
            for (int i = 0; i < L_limb; ++i) {
                const uint64_t t = (uint64_t)w[i] * 1000000000ull + carry;
                w[i] = (uint32_t)t;
                carry = t >> 32;
            }
            char buf[10];
            uint64_t v = carry;
            for (int k = 8; k >= 0; --k) {
                buf[k] = (char)('0' + v % 10);
                v /= 10;
            }
            for (int k = 0; k < 9 && start + k < L_dec; ++k) {
                out[L_dec - 1 - (start + k)] = buf[k];
            }
        }
    }
    free(w);
}

/* ------------------------------------------------------------------ */
/* leaf terms (pi_bs_gpu.cu:17-131)                                     */
/* ------------------------------------------------------------------ */

static void reduce_if_both(uint32_t *a, uint32_t *b, uint32_t d) {
    if (*a % d == 0 && *b % d == 0) {
        *a /= d;
        *b /= d;
    }
}

/* pi_bs_gpu.cu:17-34 */
static void reduce_small_primes(uint32_t *a, uint32_t *b) {
    reduce_if_both(a, b, 3);
    reduce_if_both(a, b, 5);
    reduce_if_both(a, b, 23);
    reduce_if_both(a, b, 29);
}

/* multiply a little-endian word vector (width words, truncating) by m */
static void words_mul_small(uint32_t *v, int width, uint32_t m) {
    uint64_t carry = 0;
    for (int j = 0; j < width; ++j) {
        const uint64_t t = (uint64_t)v[j] * m + carry;
        v[j] = (uint32_t)t;
        carry = t >> 32;
    }
}

/* pi_bs_gpu.cu:36-131.  Term index: head t -> 17 t + 16; body t -> t + t/16. */
void o_leaf17(int is_head, uint32_t *Pout, uint32_t *Qout, uint32_t *Rout, uint32_t count) {
    for (uint32_t t = 0; t < count; ++t) {
        const int64_t i = is_head ? (int64_t)t * 17 + 16 : (int64_t)t + (int64_t)(t >> 4);
        uint32_t p1 = (uint32_t)(6 * i + 1), p2 = (uint32_t)(2 * i + 1), p3 = (uint32_t)(6 * i + 5);
        uint32_t q1 = (uint32_t)(i + 1), q2 = q1, q3 = q1;
        uint32_t q4 = 640320u >> 6, q5 = 40020u >> 2, q6 = 426880u >> 7;
        const uint64_t r1 = (uint64_t)(545140134ll * i + 13591409ll);
        reduce_if_both(&p1, &q1, 5);
        reduce_if_both(&p1, &q2, 5);
        reduce_if_both(&p1, &q3, 5);
        reduce_small_primes(&p1, &q4);
        reduce_small_primes(&p1, &q5);
        reduce_small_primes(&p2, &q4);
        reduce_small_primes(&p2, &q5);
        reduce_small_primes(&p3, &q4);
        reduce_small_primes(&p3, &q5);

        uint32_t P[3] = {p1, 0, 0};
        words_mul_small(P, 3, p2);
        words_mul_small(P, 3, p3);
        uint32_t Q[4] = {q1, 0, 0, 0};
        const uint32_t qf[5] = {q2, q3, q4, q5, q6};
        for (int k = 0; k < 5; ++k) words_mul_small(Q, 4, qf[k]);
        uint32_t R[6] = {(uint32_t)r1, (uint32_t)(r1 >> 32), 0, 0, 0, 0};
        const uint32_t rf[5] = {q1, q2, q3, q4, q5};
        for (int k = 0; k < 5; ++k) words_mul_small(R, 6, rf[k]);
        memcpy(Pout + (size_t)t * 3, P, 12);
        memcpy(Qout + (size_t)t * 4, Q, 16);
        memcpy(Rout + (size_t)t * 6, R, 24);
    }
}

/* ------------------------------------------------------------------ */
/* binary splitting tree (pi_bs_gpu.cu:503-792)                         */
/* ------------------------------------------------------------------ */

static OArray new_array(uint32_t batch, uint32_t length) {
    OArray a;
    a.batch_size = batch;
    a.length = length;
    a.stride = length;
    a.data = xcalloc((size_t)batch * length);
    return a;
}

static void compact(OArray *a) {
    a->length = o_limblength_max(*a);
}

static int mul_into(OArray A, OArray B, OArray *C, uint32_t batch) {
    *C = new_array(batch, A.length + B.length);
    return o_mul(A, B, *C);
}

#define CHECK(x) do { int _s = (x); if (_s != 0) return _s; } while (0)

/* Merge of the whole 17N-term series.  Level loop: pi_bs_gpu.cu:625-785. */
/* The tree; with leaf_only the levels stop after the head merge and the
 * block's exact division (the 17-term group stage), leaving N rows each of
 * P, Q and R. */
static int binary_split_impl(uint32_t N, int leaf_only, OArray *P_out, OArray *Q_out, OArray *R_out) {
    if (N == 0 || (N & (N - 1)) != 0) return -1;
    const uint32_t n = N * 16u;
    const uint32_t divisor_lo = 830297u, divisor_hi = 156590819u; /* pi_bs_gpu.cu:151-152 */

    OArray P0 = new_array(N, 3), Q0 = new_array(N, 4), R0 = new_array(N, 6);
    o_leaf17(1, P0.data, Q0.data, R0.data, N);
    OArray P = new_array(n, 3), Q = new_array(n, 4), R = new_array(n, 6);
    o_leaf17(0, P.data, Q.data, R.data, n);

    for (uint32_t n1 = n; n1 > 1; n1 >>= 1) {
        const int subtract = (n1 == n) || (n1 == (n >> 4));
        const uint32_t ignore_last_P = (n1 <= N) ? 1u : 0u;
        const uint32_t half = n1 / 2;

        /* P = P_even * P_odd, the last row skipped once rows are whole groups */
        OArray Pn = new_array(half - ignore_last_P, P.length * 2);
        OArray Pe = {P.data, P.length, half - ignore_last_P, P.stride * 2};
        OArray Po = {P.data + P.stride, P.length, half - ignore_last_P, P.stride * 2};
        if (Po.batch_size > 0) {
            CHECK(o_mul(Pe, Po, Pn));
            compact(&Pn);
        }
        OArray Qn = new_array(half, Q.length * 2);
        OArray Qe = {Q.data, Q.length, half, Q.stride * 2};
        OArray Qo = {Q.data + Q.stride, Q.length, half, Q.stride * 2};
        CHECK(o_mul(Qe, Qo, Qn));
        compact(&Qn);

        OArray Re = {R.data, R.length, half, R.stride * 2};
        OArray Ro = {R.data + R.stride, R.length, half, R.stride * 2};
        OArray R1;
        CHECK(mul_into(Re, Qo, &R1, half));
        const uint32_t terms_right = (n / n1) + ((n / n1) >> 4);
        const uint32_t shift = 15u * terms_right;
        Pe.batch_size = half;
        OArray R2;
        CHECK(mul_into(Pe, Ro, &R2, half));
        const uint32_t shifted_len = R1.length + (shift >> 5) + ((shift & 31u) != 0);
        uint32_t rlen = (shifted_len > R2.length ? shifted_len : R2.length) + (subtract ? 0u : 1u);
        OArray Rn = new_array(half, rlen);
        if (subtract) CHECK(o_shift_sub(R1, R2, shift, Rn));
        else CHECK(o_shift_add(R1, R2, shift, Rn));
        compact(&Rn);
        free(R1.data); free(R2.data);
        free(P.data); free(Q.data); free(R.data);
        P = Pn; Q = Qn; R = Rn;

        if (n1 == (n >> 3)) {
            /* append the head term 17g+16 to each 16-term body block
             * (pi_bs_gpu.cu:725-758), then divide out 13^2 17^3 7^6 11^3
             * (pi_bs_gpu.cu:760-779) */
            OArray Ph, Qh, Rp1, Rp2;
            CHECK(mul_into(P, P0, &Ph, half));
            compact(&Ph);
            CHECK(mul_into(Q, Q0, &Qh, half));
            compact(&Qh);
            CHECK(mul_into(R, Q0, &Rp1, half));
            CHECK(mul_into(P, R0, &Rp2, half));
            uint32_t l1 = Rp1.length + 1;
            uint32_t hl = (l1 > Rp2.length ? l1 : Rp2.length) + 1;
            OArray Rh = new_array(half, hl);
            CHECK(o_shift_add(Rp1, Rp2, 15u, Rh));
            compact(&Rh);
            CHECK(o_exactdiv_small(Ph, divisor_lo));
            CHECK(o_exactdiv_small(Ph, divisor_hi));
            compact(&Ph);
            CHECK(o_exactdiv_small(Qh, divisor_lo));
            CHECK(o_exactdiv_small(Qh, divisor_hi));
            compact(&Qh);
            CHECK(o_exactdiv_small(Rh, divisor_lo));
            CHECK(o_exactdiv_small(Rh, divisor_hi));
            compact(&Rh);
            free(Rp1.data); free(Rp2.data);
            free(P.data); free(Q.data); free(R.data);
            P = Ph; Q = Qh; R = Rh;
            if (leaf_only) break;
        }
    }
    free(P0.data); free(Q0.data); free(R0.data);
    if (P_out) *P_out = P;
    else free(P.data);
    *Q_out = Q;
    *R_out = R;
    return 0;
}

int o_binary_split(uint32_t N, OArray *Q_out, OArray *R_out) {
    return binary_split_impl(N, 0, NULL, Q_out, R_out);
}

/* pi_bs_gpu.cu:555-779: leaf terms, the four pairwise levels of each
 * 16-term body, the head-term merge and the exact division by
 * 13^2 17^3 7^6 11^3: one row of P, Q, R per 17-term group. */
int o_leaf_stage(uint32_t N, OArray *P_out, OArray *Q_out, OArray *R_out) {
    if (P_out == NULL) return -1;
    return binary_split_impl(N, 1, P_out, Q_out, R_out);
}

/* ------------------------------------------------------------------ */
/* sqrt(10005) by Pell doubling (pi_bs_gpu.cu:1346-1522)                */
/* ------------------------------------------------------------------ */

int o_sqrt_10005(int target_digits, OArray *P_out, OArray *Q_out) {
    double valid_digits = 7.5;
    OArray P = new_array(1, 1), Q = new_array(1, 1);
    P.data[0] = 4001;
    Q.data[0] = 40;
    while (valid_digits < target_digits) {
        valid_digits *= 2;
        OArray PQ;
        CHECK(mul_into(P, Q, &PQ, 1));
        compact(&PQ);
        OArray Qn = new_array(1, PQ.length + 1);
        CHECK(o_shift_bits(PQ, Qn, 1));
        compact(&Qn);
        OArray PP;
        CHECK(mul_into(P, P, &PP, 1));
        compact(&PP);
        OArray QQ;
        CHECK(mul_into(Q, Q, &QQ, 1));
        compact(&QQ);
        OArray QQ5 = new_array(1, QQ.length + 1);
        CHECK(o_mul_small(QQ, 10005u, QQ5));
        compact(&QQ5);
        const uint32_t l = (PP.length > QQ5.length ? PP.length : QQ5.length) + 1;
        OArray Pn = new_array(1, l);
        CHECK(o_add(PP, QQ5, Pn));
        compact(&Pn);
        free(PQ.data); free(PP.data); free(QQ.data); free(QQ5.data);
        free(P.data); free(Q.data);
        P = Pn;
        Q = Qn;
    }
    *P_out = P;
    *Q_out = Q;
    return 0;
}

/* ------------------------------------------------------------------ */
/* inverse (pi_bs_gpu.cu:1068-1110, 1143-1344)                          */
/* ------------------------------------------------------------------ */

/* pi_bs_gpu.cu:1095-1110 */
static OArray make_subview(OArray a, uint32_t start, uint32_t length) {
    if (start > a.length) start = a.length;
    const uint32_t max_len = a.length - start;
    if (length > max_len) length = max_len;
    OArray v;
    v.data = a.data + start;
    v.length = length;
    v.batch_size = a.batch_size;
    v.stride = (a.stride >= start) ? (a.stride - start) : 0u;
    return v;
}

/* pi_bs_gpu.cu:1068-1093 */
static int shift_right_large(OArray A, uint64_t shift_bits, OArray B) {
    const uint64_t limb_shift = shift_bits / 32u;
    const uint32_t bit_shift = (uint32_t)(shift_bits % 32u);
    if (limb_shift >= A.length) return -1;
    OArray pre;
    pre.data = A.data + limb_shift;
    pre.length = A.length - (uint32_t)limb_shift;
    pre.batch_size = A.batch_size;
    pre.stride = A.stride - (uint32_t)limb_shift;
    return o_shift_bits(pre, B, -(int32_t)bit_shift);
}

/* pi_bs_gpu.cu:1112-1126 */
static void load_mpz(mpz_t v, OArray a) {
    mpz_import(v, a.length, -1, 4, 0, 0, a.data);
}

/* pi_bs_gpu.cu:1128-1141: writes a.stride words; returns used length */
static int store_mpz(const mpz_t v, OArray *a) {
    size_t cap = a->stride ? a->stride : 1;
    uint32_t *host = xcalloc(cap + ((mpz_sizeinbase(v, 2) + 31) / 32) + 1);
    size_t written = 0;
    mpz_export(host, &written, -1, 4, 0, 0, v);
    uint32_t used = written ? (uint32_t)written : 1u;
    if (used > a->stride) {
        free(host);
        return -1;
    }
    memcpy(a->data, host, (size_t)a->stride * 4);
    a->length = used;
    free(host);
    return 0;
}

#define POISON_WORD 0xA5A5A5A5u

int o_recursive_inverse(OArray x, uint64_t L, OArray *y_out) {
    const uint64_t cpu_limit = 32u * 1024u; /* pi_bs_gpu.cu:1143 */
    uint64_t levels[128];
    int nlev = 0;
    levels[nlev++] = L;
    while (levels[nlev - 1] > cpu_limit) {
        const uint64_t m = ((levels[nlev - 1] - 2u) / 4u) & ~31ull;
        levels[nlev] = levels[nlev - 1] - m * 2u;
        ++nlev;
    }
    const size_t L_res = (size_t)((L + 31) / 32 + 2);
    const size_t L_res2 = (size_t)((L + 31) / 32 + 1);
    const size_t L_y2 = (size_t)x.length + 1;

    OArray y = new_array(1, x.length); /* zeroed, as the reference memsets */
    uint32_t *ws = xmalloc((L_res + L_res2 + L_y2) * 4);
    for (size_t i = 0; i < L_res + L_res2 + L_y2; ++i) ws[i] = POISON_WORD;
    uint32_t *res_base = ws;
    uint32_t *res2_base = ws + L_res;
    uint32_t *y2_base = res2_base + L_res2;
    memset(ws, 0, L_res * 4);

    OArray res;
    {
        const uint64_t L0 = levels[nlev - 1];
        const uint32_t start_limb = (uint32_t)((L - L0) >> 6);
        OArray x0 = make_subview(x, start_limb, x.length - start_limb);
        OArray y0 = make_subview(y, start_limb, y.length - start_limb);
        res.data = res_base + (size_t)start_limb * 2;
        res.length = x0.length + 1;
        res.batch_size = 1;
        res.stride = x0.length + 1;
        mpz_t total, xv, yv, rv;
        mpz_init(total); mpz_init(xv); mpz_init(yv); mpz_init(rv);
        load_mpz(xv, x0);
        mpz_set_ui(total, 1);
        mpz_mul_2exp(total, total, L0);
        mpz_fdiv_qr(yv, rv, total, xv);
        int st = store_mpz(yv, &y0);
        if (st == 0) st = store_mpz(rv, &res);
        mpz_clear(total); mpz_clear(xv); mpz_clear(yv); mpz_clear(rv);
        if (st != 0) { free(ws); free(y.data); return st; }
    }
    compact(&res);

    for (int i = nlev - 1; i > 0; --i) {
        const uint64_t L1 = levels[i], L2 = levels[i - 1];
        const uint64_t m = (L2 - L1) >> 1;
        const uint32_t mw = (uint32_t)(m >> 5);
        OArray x2 = make_subview(x, (uint32_t)((L - L2) >> 6), mw);
        OArray y1 = make_subview(y, (uint32_t)((L - L1) >> 6), y.length - (uint32_t)((L - L1) >> 6));
        OArray x2y1 = {res2_base, x2.length + y1.length, 1, x2.length + y1.length};
        CHECK(o_mul(x2, y1, x2y1));
        OArray res1_m = {res.data - mw, res.length + mw + 1, 1, res.length + mw + 1};
        CHECK(o_sub(res1_m, x2y1, res1_m));
        const uint32_t top = res1_m.data[res1_m.length - 1];
        const int negative = top >= 0x80000000u;
        if (negative) {
            OArray zero = {res1_m.data, 0, 1, 0};
            CHECK(o_sub(zero, res1_m, res1_m));
        }
        OArray res2 = {res2_base, res1_m.length + y1.length, 1, res1_m.length + y1.length};
        CHECK(o_mul(res1_m, y1, res2));
        const uint32_t drop = (uint32_t)((L2 - m * 2) >> 5);
        OArray y2 = {y2_base, res2.length - drop, 1, res2.length - drop};
        CHECK(shift_right_large(res2, L2 - m * 2u, y2));
        if (negative) CHECK(o_add_small(y2, 1, y2));
        OArray y_new = make_subview(y, (uint32_t)((L - L2) >> 6), y.length - (uint32_t)((L - L2) >> 6));
        if (!negative) CHECK(o_add(y_new, y2, y_new));
        else CHECK(o_sub(y_new, y2, y_new));
        if (L2 != L) {
            OArray x_next = make_subview(x, (uint32_t)((L - L2) >> 6), x.length - (uint32_t)((L - L2) >> 6));
            res.data = res1_m.data - mw;
            res.length = res1_m.length + mw;
            res.batch_size = 1;
            res.stride = res.length;
            OArray y2x = {res2_base, y2.length + x_next.length, 1, y2.length + x_next.length};
            CHECK(o_mul(y2, x_next, y2x));
            if (!negative) CHECK(o_sub(res, y2x, res));
            else CHECK(o_sub(y2x, res, res));
            compact(&res);
        }
    }
    free(ws);
    *y_out = y;
    return 0;
}

/* ------------------------------------------------------------------ */
/* pi_fractional (pi_bs_gpu.cu:1524-1685)                               */
/* ------------------------------------------------------------------ */

int o_pi_fractional(int target_digits, OArray *ret_out, OPiInfo *info) {
    int N = 1;
    while (N * 17 * 14.3 < target_digits) N *= 2;
    const int prec_limbs = (int)(((long)(target_digits * 3.322) + 31) / 32 + 1);

    OArray Q_bs, R_bs;
    CHECK(o_binary_split((uint32_t)N, &Q_bs, &R_bs));
    const unsigned long long r_total_shift = 15ull * (unsigned long long)N * 17ull - 8ull;
    const int rs_bits = (int)(r_total_shift % 32);
    const int rs_limbs = (int)(r_total_shift / 32);
    OArray R_pre = {R_bs.data + rs_limbs, R_bs.length - (uint32_t)rs_limbs, 1, R_bs.length - (uint32_t)rs_limbs};
    OArray R_sh = new_array(1, R_pre.length + 1);
    CHECK(o_shift_bits(R_pre, R_sh, -rs_bits));
    free(R_bs.data);
    compact(&R_sh);

    long extra = (long)(Q_bs.length < R_sh.length ? Q_bs.length : R_sh.length) - prec_limbs;
    if (extra < 0) extra = 0;
    OArray Q_tr = {Q_bs.data + extra, Q_bs.length - (uint32_t)extra, 1, Q_bs.length - (uint32_t)extra};
    OArray R_tr = {R_sh.data + extra, R_sh.length - (uint32_t)extra, 1, R_sh.length - (uint32_t)extra};

    OArray P15, Q15;
    CHECK(o_sqrt_10005(target_digits, &P15, &Q15));
    OArray P_final = new_array(1, P15.length + Q_tr.length);
    OArray Q_final = new_array(1, Q15.length + R_tr.length);
    CHECK(o_mul(P15, Q_tr, P_final));
    CHECK(o_mul(R_tr, Q15, Q_final));
    compact(&P_final);
    compact(&Q_final);
    free(Q_bs.data); free(R_sh.data); free(P15.data); free(Q15.data);

    extra = (long)(P_final.length < Q_final.length ? P_final.length : Q_final.length) - prec_limbs;
    if (extra < 0) extra = 0;
    OArray P_tf = make_subview(P_final, (uint32_t)extra, P_final.length - (uint32_t)extra);
    OArray Q_tf = make_subview(Q_final, (uint32_t)extra, Q_final.length - (uint32_t)extra);
    compact(&Q_tf);
    uint32_t q_bits = 0;
    if (Q_tf.length > 0) {
        const uint32_t top = Q_tf.data[Q_tf.length - 1];
        q_bits = (Q_tf.length - 1) * 32u + (32u - (uint32_t)__builtin_clz(top));
    }
    const uint64_t inverse_bits = (uint64_t)q_bits * 2u - 1u;

    OArray inv;
    CHECK(o_recursive_inverse(Q_tf, inverse_bits, &inv));
    OArray prod = new_array(1, P_tf.length + inv.length);
    CHECK(o_mul(P_tf, inv, prod));
    compact(&prod);
    free(inv.data); free(P_final.data); free(Q_final.data);

    OArray ret = new_array(1, prod.length - (uint32_t)(inverse_bits / 32u) + (uint32_t)prec_limbs);
    const uint64_t final_shift = inverse_bits - (uint64_t)prec_limbs * 32u;
    CHECK(shift_right_large(prod, final_shift, ret));
    compact(&ret);
    free(prod.data);
    if (info) {
        info->target_digits = (uint32_t)target_digits;
        info->groups = (uint32_t)N;
        info->prec_limbs = (uint32_t)prec_limbs;
        info->inverse_bits = inverse_bits;
    }
    *ret_out = ret;
    return 0;
}

/* ------------------------------------------------------------------ */
/* sizing (pi_bs_gpu.cu:924-944; convert_decimal.cu:36-52)             */
/* ------------------------------------------------------------------ */

uint32_t o_round_up_decimal_digits(uint32_t requested, uint32_t *leaf_digits) {
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

uint32_t o_prec_limbs_for_digits(uint32_t digits) {
    const uint32_t prec_bits = (uint32_t)((double)digits * 3.322) + 31u;
    return prec_bits / 32u + 1u;
}

uint32_t o_bits_for_digits(uint32_t digits) {
    const double bits = ceil((double)digits * 3.32192809488736234787) + 34.0;
    return (uint32_t)bits;
}

/* ------------------------------------------------------------------ */
/* radix conversion (convert_decimal.cu:67-136, 198-474)                */
/* ------------------------------------------------------------------ */

/* convert_decimal.cu:67-96: dst row (2r + off) = src row r >> shift (window) */
static void extract_shifted_rows(OArray src, OArray dst, uint32_t rows, uint32_t shift_bits,
                                 uint32_t dst_row_offset) {
    const uint32_t ws = shift_bits >> 5, bs = shift_bits & 31u;
    for (uint32_t r = 0; r < rows; ++r) {
        const uint32_t *s = src.data + (size_t)r * src.stride;
        uint32_t *d = dst.data + (size_t)(r * 2u + dst_row_offset) * dst.stride;
        for (uint32_t j = 0; j < dst.length; ++j) {
            const uint32_t k = j + ws;
            uint64_t v = 0;
            if (k < src.length) {
                v = (uint64_t)s[k] >> bs;
                if (bs != 0 && k + 1 < src.length) v |= (uint64_t)s[k + 1] << (32u - bs);
            }
            d[j] = (uint32_t)v;
        }
    }
}

/* convert_decimal.cu:98-119 + 170-191 */
static void extract_low_bits_rows(OArray src, OArray dst, uint32_t rows, uint32_t valid_bits) {
    uint32_t mask = 0xffffffffu;
    if (valid_bits & 31u) mask = (1u << (valid_bits & 31u)) - 1u;
    for (uint32_t r = 0; r < rows; ++r) {
        const uint32_t *s = src.data + (size_t)r * src.stride;
        uint32_t *d = dst.data + (size_t)r * dst.stride;
        for (uint32_t j = 0; j < dst.length; ++j) {
            uint32_t v = s[j];
            if (mask != 0xffffffffu && j + 1 == dst.length) v &= mask;
            d[j] = v;
        }
    }
}

int o_convert_precompute_powers(uint32_t leaf_digits, uint32_t levels, OArray *powers) {
    if (levels == 0) return 0;
    mpz_t b;
    mpz_init(b);
    mpz_ui_pow_ui(b, 5u, leaf_digits);
    uint32_t limbs = (uint32_t)((mpz_sizeinbase(b, 2) + 31) / 32);
    if (limbs < 1) limbs = 1;
    powers[0] = new_array(1, limbs);
    size_t written = 0;
    uint32_t *tmp = xcalloc(limbs + 1);
    mpz_export(tmp, &written, -1, 4, 0, 0, b);
    memcpy(powers[0].data, tmp, (size_t)limbs * 4);
    free(tmp);
    mpz_clear(b);
    for (uint32_t l = 1; l < levels; ++l) {
        CHECK(mul_into(powers[l - 1], powers[l - 1], &powers[l], 1));
        compact(&powers[l]);
    }
    return 0;
}

static uint32_t levels_for_shape(uint32_t total, uint32_t leaf) {
    if (leaf == 0 || total == 0 || total % leaf) return UINT32_MAX;
    uint32_t ratio = total / leaf, lv = 0;
    while (ratio > 1 && (ratio & 1u) == 0) { ratio >>= 1; ++lv; }
    return ratio == 1 ? lv : UINT32_MAX;
}

int o_convert_equal_split(OArray input, uint32_t input_bits, uint32_t total_digits,
                          uint32_t leaf_digits, const OArray *powers, uint32_t levels_in,
                          char *output_digits) {
    if (input.batch_size != 1 || !input.data || !output_digits) return -1;
    const uint32_t levels = levels_for_shape(total_digits, leaf_digits);
    if (levels == UINT32_MAX || levels != levels_in) return -1;

    uint32_t *dig = xmalloc((levels + 1) * 4), *bits = xmalloc((levels + 1) * 4), *limbs = xmalloc((levels + 1) * 4);
    dig[0] = total_digits;
    bits[0] = input_bits;
    limbs[0] = input.length;
    for (uint32_t l = 1; l <= levels; ++l) {
        dig[l] = dig[l - 1] >> 1;
        bits[l] = o_bits_for_digits(dig[l]);
        limbs[l] = (bits[l] + 31) / 32;
    }
    OArray cur = input;
    int own = 0;
    if (levels == 0) {
        const uint32_t need = o_bits_for_digits(total_digits);
        if (input_bits < need) return -1;
        OArray t = new_array(1, (need + 31) / 32);
        extract_shifted_rows(input, t, 1, input_bits - need, 0);
        cur = t;
        own = 1;
    } else {
        for (uint32_t l = 0; l < levels; ++l) {
            const uint32_t child_digits = dig[l + 1], child_bits = bits[l + 1];
            if (cur.length < limbs[l]) return -1;
            if (bits[l] <= child_digits || bits[l] < child_bits) return -1;
            OArray next = new_array(cur.batch_size * 2u, limbs[l + 1]);
            extract_shifted_rows(cur, next, cur.batch_size, bits[l] - child_bits, 0);
            const uint32_t mant_bits = bits[l] - child_digits;
            OArray mant = new_array(cur.batch_size, (mant_bits + 31) / 32);
            extract_low_bits_rows(cur, mant, cur.batch_size, mant_bits);
            OArray pw = {powers[levels - 1 - l].data, powers[levels - 1 - l].length, cur.batch_size, 0};
            OArray prod = new_array(cur.batch_size, mant.length + pw.length);
            CHECK(o_mul(mant, pw, prod));
            extract_shifted_rows(prod, next, cur.batch_size, mant_bits - child_bits, 1);
            free(mant.data);
            free(prod.data);
            if (own) free(cur.data);
            cur = next;
            own = 1;
        }
    }
    OArray aligned = new_array(cur.batch_size, cur.length);
    const uint32_t leaf_bits = levels == 0 ? o_bits_for_digits(total_digits) : bits[levels];
    CHECK(o_shift_bits(cur, aligned, (int32_t)(cur.length * 32u - leaf_bits)));
    char *chunks = xmalloc((size_t)cur.batch_size * leaf_digits);
    o_decimal_naive(aligned.data, chunks, cur.batch_size, (int)aligned.length, aligned.stride,
                    (int)leaf_digits, leaf_digits);
    for (uint32_t c = 0; c < cur.batch_size; ++c)
        for (uint32_t d = 0; d < leaf_digits; ++d)
            output_digits[(size_t)c * leaf_digits + d] = chunks[(size_t)c * leaf_digits + (leaf_digits - 1 - d)];
    free(chunks);
    free(aligned.data);
    if (own) free(cur.data);
    free(dig); free(bits); free(limbs);
    return 0;
}

/* ------------------------------------------------------------------ */
/* whole program (pi_bs_gpu.cu:2073-2195) and hashes (:852-866)         */
/* ------------------------------------------------------------------ */

int o_pi_digits(uint32_t requested, int readme_prec, uint32_t **ret_limbs,
                uint32_t *ret_len, char **digits_out, uint32_t *convert_digits_out) {
    uint32_t leaf = 477;
    const uint32_t convert_digits = o_round_up_decimal_digits(requested, &leaf);
    const uint32_t prec_digits = readme_prec ? convert_digits : requested;
    const uint32_t prec_limbs = o_prec_limbs_for_digits(prec_digits);
    const uint32_t levels = convert_digits ? (uint32_t)__builtin_ctz(convert_digits / leaf) : 0u;

    OArray ret;
    OPiInfo info;
    CHECK(o_pi_fractional((int)prec_digits, &ret, &info));
    OArray *powers = xmalloc((levels + 1) * sizeof(OArray));
    CHECK(o_convert_precompute_powers(leaf, levels, powers));
    char *digits = xmalloc((size_t)convert_digits + 1);
    if (requested > 0) {
        OArray frac = make_subview(ret, 0, prec_limbs < ret.length ? prec_limbs : ret.length);
        CHECK(o_convert_equal_split(frac, prec_limbs * 32u, convert_digits, leaf, powers, levels, digits));
    }
    for (uint32_t l = 0; l < levels; ++l) free(powers[l].data);
    free(powers);
    *ret_limbs = ret.data;
    *ret_len = ret.length;
    *digits_out = digits;
    if (convert_digits_out) *convert_digits_out = convert_digits;
    return 0;
}

uint32_t o_mod_hash(const uint32_t *limbs, uint32_t length) {
    const uint64_t mod = 1000000007ull, base = (1ull << 32) % mod;
    uint64_t v = 0;
    for (int64_t i = (int64_t)length - 1; i >= 0; --i) v = (v * base + limbs[i]) % mod;
    return (uint32_t)v;
}

uint32_t o_decimal_hash(const char *digits, uint32_t length) {
    const uint64_t mod = 1000000007ull;
    uint64_t v = 3;
    for (uint32_t i = 0; i < length; ++i) v = (v * 10u + (uint32_t)(digits[i] - '0')) % mod;
    return (uint32_t)v;
}

void o_free(void *p) { free(p); }
