/*
 * pi_oracle: CPU run of the oracle's restated pipeline — TEST INFRASTRUCTURE.
 * Output mirrors the reference CLI's --benchmark lines (pi_bs_gpu.cu:2217-2245)
 * so that GPU and CPU runs can be diffed.
 *
 *   pi_oracle <digits> [--readme-prec] [--full-print]
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "oracle.h"

static double now_ms(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

int main(int argc, char **argv) {
    if (argc < 2) {
        fprintf(stderr, "usage: %s <digits> [--readme-prec] [--full-print]\n", argv[0]);
        return 1;
    }
    const long d = strtol(argv[1], NULL, 10);
    if (d < 0) return 1;
    int readme = 0, full = 0;
    for (int i = 2; i < argc; ++i) {
        if (!strcmp(argv[i], "--readme-prec")) readme = 1;
        else if (!strcmp(argv[i], "--full-print")) full = 1;
        else { fprintf(stderr, "unknown flag %s\n", argv[i]); return 1; }
    }
    uint32_t *ret = NULL, ret_len = 0, cdig = 0;
    char *digits = NULL;
    const double t0 = now_ms();
    if (o_pi_digits((uint32_t)d, readme, &ret, &ret_len, &digits, &cdig) != 0) {
        fprintf(stderr, "oracle failed\n");
        return 1;
    }
    const double t1 = now_ms();
    printf("target_digits=%ld convert_digits=%u RET_len=%u elapsed_ms=%.3f\n", d, cdig, ret_len, t1 - t0);
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
