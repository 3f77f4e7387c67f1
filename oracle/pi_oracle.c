/*
 * pi_oracle: CPU runs of pi — TEST INFRASTRUCTURE and the benchmark's CPU
 * baseline.
 *
 *   pi_oracle <digits> [--readme-prec] [--full-print] [--threads T]
 *       the oracle's restatement of the reference pipeline (o_pi_digits);
 *       output mirrors the reference CLI's --benchmark lines
 *       (pi_bs_gpu.cu:2217-2245) so that GPU and CPU runs can be diffed.
 *       With T > 1 the rows of every batched product run on T threads (the
 *       binary-splitting tree's and the radix conversion's independent rows;
 *       the top levels hold one or two rows and stay serial).
 *   pi_oracle <digits> --classic [--threads T]
 *       a textbook Chudnovsky run on GMP, for the fastest-digits CPU figure:
 *       recursive binary splitting (P, Q, T with T = T1 Q2 + P1 T2) whose top
 *       log2(T) levels of subtrees run on separate threads, then
 *       floor(sqrt(10005 * 10^2d)), one division and mpz_get_str.  Prints the
 *       digits' hash and the first and last 20 digits.
 */
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "gmp_decl.h"
#include "oracle.h"

void __gmpz_mul(mpz_ptr, mpz_srcptr, mpz_srcptr);
void __gmpz_mul_ui(mpz_ptr, mpz_srcptr, unsigned long);
void __gmpz_add(mpz_ptr, mpz_srcptr, mpz_srcptr);
void __gmpz_neg(mpz_ptr, mpz_srcptr);
void __gmpz_sqrt(mpz_ptr, mpz_srcptr);
void __gmpz_tdiv_q(mpz_ptr, mpz_srcptr, mpz_srcptr);
char *__gmpz_get_str(char *, int, mpz_srcptr);

static double now_ms(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

/* ------------------------------------------------------------------ */
/* textbook Chudnovsky                                                 */
/* ------------------------------------------------------------------ */

typedef struct {
    unsigned long a, b;
    int depth;
    mpz_t P, Q, T;
} Split;

static void split(Split *s);

static void *split_thread(void *p) {
    split((Split *)p);
    return NULL;
}

/* terms [a, b): leaf p_k = (6k-5)(2k-1)(6k-1), q_k = k^3 C^3/24,
 * t_k = p_k (13591409 + 545140134 k) (-1)^k (p_0 = q_0 = 1) */
static void split(Split *s) {
    __gmpz_init(s->P);
    __gmpz_init(s->Q);
    __gmpz_init(s->T);
    if (s->b - s->a == 1) {
        const unsigned long k = s->a;
        if (k == 0) {
            __gmpz_set_ui(s->P, 1);
            __gmpz_set_ui(s->Q, 1);
        } else {
            __gmpz_set_ui(s->P, 6 * k - 5);
            __gmpz_mul_ui(s->P, s->P, 2 * k - 1);
            __gmpz_mul_ui(s->P, s->P, 6 * k - 1);
            __gmpz_set_ui(s->Q, k);
            __gmpz_mul_ui(s->Q, s->Q, k);
            __gmpz_mul_ui(s->Q, s->Q, k);
            __gmpz_mul_ui(s->Q, s->Q, 10939058860032000ul);
        }
        __gmpz_mul_ui(s->T, s->P, 13591409ul + 545140134ul * k);
        if (k & 1) __gmpz_neg(s->T, s->T);
        return;
    }
    const unsigned long m = (s->a + s->b) / 2;
    Split L = {s->a, m, s->depth - 1, {{0}}, {{0}}, {{0}}};
    Split R = {m, s->b, s->depth - 1, {{0}}, {{0}}, {{0}}};
    pthread_t th;
    const int threaded = s->depth > 0 && pthread_create(&th, NULL, split_thread, &L) == 0;
    if (!threaded) split(&L);
    split(&R);
    if (threaded) pthread_join(th, NULL);
    __gmpz_mul(s->P, L.P, R.P);
    __gmpz_mul(s->Q, L.Q, R.Q);
    __gmpz_mul(s->T, L.T, R.Q);
    __gmpz_mul(L.Q, L.P, R.T); /* reuse L.Q */
    __gmpz_add(s->T, s->T, L.Q);
    __gmpz_clear(L.P); __gmpz_clear(L.Q); __gmpz_clear(L.T);
    __gmpz_clear(R.P); __gmpz_clear(R.Q); __gmpz_clear(R.T);
}

static int run_classic(long d, int threads) {
    int depth = 0;
    while ((1 << depth) < threads) ++depth;
    const double t0 = now_ms();
    Split s = {0, (unsigned long)(d / 14.1816474627254776555 + 2), depth, {{0}}, {{0}}, {{0}}};
    split(&s);
    mpz_t c, pi;
    __gmpz_init(c);
    __gmpz_init(pi);
    __gmpz_ui_pow_ui(c, 10, 2 * (unsigned long)d); /* sqrt(10005) 10^d */
    __gmpz_mul_ui(c, c, 10005);
    __gmpz_sqrt(c, c);
    __gmpz_mul(pi, s.Q, c);
    __gmpz_mul_ui(pi, pi, 426880);
    __gmpz_tdiv_q(pi, pi, s.T);
    char *str = __gmpz_get_str(NULL, 10, pi);
    const double t1 = now_ms();
    const size_t n = strlen(str);
    printf("classic target_digits=%ld threads=%d elapsed_ms=%.3f\n", d, threads, t1 - t0);
    printf("DEC = %.20s...%.20s (%zu chars)\n", str, n > 20 ? str + n - 20 : str, n);
    if (n > 1) printf("DEC_hash = %u\n", o_decimal_hash(str + 1, (uint32_t)(n - 1)));
    free(str);
    __gmpz_clear(c); __gmpz_clear(pi);
    __gmpz_clear(s.P); __gmpz_clear(s.Q); __gmpz_clear(s.T);
    return 0;
}

int main(int argc, char **argv) {
    if (argc < 2) {
        fprintf(stderr, "usage: %s <digits> [--readme-prec] [--full-print] [--threads T] [--classic]\n", argv[0]);
        return 1;
    }
    const long d = strtol(argv[1], NULL, 10);
    if (d < 0) return 1;
    int readme = 0, full = 0, classic = 0, threads = 1;
    for (int i = 2; i < argc; ++i) {
        if (!strcmp(argv[i], "--readme-prec")) readme = 1;
        else if (!strcmp(argv[i], "--full-print")) full = 1;
        else if (!strcmp(argv[i], "--classic")) classic = 1;
        else if (!strcmp(argv[i], "--threads") && i + 1 < argc) threads = atoi(argv[++i]);
        else { fprintf(stderr, "unknown flag %s\n", argv[i]); return 1; }
    }
    if (threads < 1) threads = 1;
    if (classic) return run_classic(d, threads);
    o_set_threads(threads);
    uint32_t *ret = NULL, ret_len = 0, cdig = 0;
    char *digits = NULL;
    const double t0 = now_ms();
    if (o_pi_digits((uint32_t)d, readme, &ret, &ret_len, &digits, &cdig) != 0) {
        fprintf(stderr, "oracle failed\n");
        return 1;
    }
    const double t1 = now_ms();
    printf("target_digits=%ld convert_digits=%u RET_len=%u threads=%d elapsed_ms=%.3f\n", d, cdig, ret_len, threads,
           t1 - t0);
    printf("RET = ");
    if (ret_len == 0) printf("0");
    for (uint32_t i = 0; i < 2 && i < ret_len; ++i) printf(i ? "%08x" : "%x", ret[ret_len - 1 - i]);
    if (ret_len > 4) printf("...");
    for (uint32_t i = (ret_len < 2 ? ret_len : 2); i > 0; --i) {
        if (ret_len <= 4 && i - 1 >= ret_len - 2) continue;
        printf("%08x", ret[i - 1]);
    }
    printf(" hash = %u\n", o_mod_hash(ret, ret_len));
    const uint32_t pre = d < 63 ? (uint32_t)d : 63u, suf = d < 64 ? (uint32_t)d : 64u;
    printf("DEC = 3%.*s...%.*s\n", (int)pre, digits, (int)suf, digits + (d - suf));
    printf("DEC_hash = %u\n", o_decimal_hash(digits, (uint32_t)d));
    if (full) {
        printf("RET_FULL = ");
        for (int64_t i = (int64_t)ret_len - 1; i >= 0; --i)
            printf(i == (int64_t)ret_len - 1 ? "%x" : "%08x", ret[i]);
        printf("\nDEC_FULL = 3%.*s\n", (int)d, digits);
    }
    o_free(ret);
    o_free(digits);
    return 0;
}
