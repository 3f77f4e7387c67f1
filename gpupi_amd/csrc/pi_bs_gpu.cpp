// pi_bs_gpu: command line front end of the library, same interface and output
// formats as the reference CLI (pi_bs_gpu.cu:2073-2431):
//
//   pi_bs_gpu <digits> [--benchmark|--profile-stages|--full-print] [--readme-prec]
//             [--output FILE] [--verify-bbp POS]...
//
// default: "3" + digits on stdout, the stats line on stderr.  Additions:
// --readme-prec sizes the precision from the rounded-up digit count, the
// convention under which the reference README's RET length and hash were
// published; --output writes "3." and the digits to FILE instead of stdout;
// --verify-bbp checks eight hexadecimal digits of RET from fractional
// position POS (1-based) against the Bailey-Borwein-Plouffe digit-extraction
// formula evaluated independently on the host (exit status 2 on a mismatch).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "batch_arith.h"
#include "gpupi_pi.h"

namespace {

constexpr uint64_t kValidationMod = 1000000007ull;

// pi_bs_gpu.cu:852-858: value mod 1e9+7 in base 2^32
uint32_t mod_hash(const uint32_t *limbs, uint32_t n) {
    const uint64_t base = (1ull << 32) % kValidationMod;
    uint64_t v = 0;
    for (int64_t i = (int64_t)n - 1; i >= 0; --i) v = (v * base + limbs[i]) % kValidationMod;
    return (uint32_t)v;
}

// pi_bs_gpu.cu:860-866
uint32_t decimal_hash(const char *d, uint32_t n) {
    uint64_t v = 3;
    for (uint32_t i = 0; i < n; ++i) v = (v * 10u + (uint32_t)(d[i] - '0')) % kValidationMod;
    return (uint32_t)v;
}

// pi_bs_gpu.cu:868-896
void print_hex_preview(const char *label, const std::vector<uint32_t> &v) {
    const uint32_t n = (uint32_t)v.size();
    const uint32_t h = mod_hash(v.data(), n);
    if (n == 0) {
        printf("%s = 0 hash = %u\n", label, h);
        return;
    }
    const uint32_t first = std::min(n, 2u), last = std::min(n, 2u);
    printf("%s = ", label);
    for (uint32_t i = 0; i < last; ++i) printf(i == 0 ? "%x" : "%08x", v[n - 1 - i]);
    if (n > first + last) printf("...");
    for (uint32_t i = first; i > 0; --i) {
        const uint32_t idx = i - 1;
        if (n <= first + last && idx >= n - last) continue;
        printf("%08x", v[idx]);
    }
    printf(" hash = %u\n", h);
}

void usage(const char *prog) {
    fprintf(stderr,
            "Usage: %s <target_digits> [--benchmark|--profile-stages|--full-print] [--readme-prec] "
            "[--output FILE] [--verify-bbp POS]...\n",
            prog);
}

// 16^e mod m for m < 2^32 (products fit in 64 bits).
uint64_t pow16_mod(uint64_t e, uint64_t m) {
    if (m == 1) return 0;
    uint64_t r = 1, b = 16 % m;
    while (e) {
        if (e & 1) r = r * b % m;
        b = b * b % m;
        e >>= 1;
    }
    return r;
}

// frac(sum_k 16^(n-k) / (8k + j)), k >= 0 (Bailey, Borwein and Plouffe,
// "On the rapid computation of various polylogarithmic constants", 1997).
long double bbp_series(uint64_t n, uint32_t j) {
    long double s = 0.0L;
    for (uint64_t k = 0; k <= n; ++k) {
        const uint64_t m = 8 * k + j;
        s += (long double)pow16_mod(n - k, m) / (long double)m;
        s -= floorl(s);
    }
    long double p = 1.0L / 16.0L;
    for (uint64_t k = n + 1; k <= n + 24; ++k, p /= 16.0L) s += p / (long double)(8 * k + j);
    return s - floorl(s);
}

// Eight hexadecimal digits of pi from fractional position n + 1:
// frac(16^n pi) = frac(4 S1 - 2 S4 - S5 - S6).
uint32_t bbp_hex8(uint64_t n) {
    const uint32_t js[4] = {1, 4, 5, 6};
    long double s[4];
    std::vector<std::thread> th;
    for (int i = 0; i < 4; ++i) th.emplace_back([&, i] { s[i] = bbp_series(n, js[i]); });
    for (auto &t : th) t.join();
    long double x = 4.0L * s[0] - 2.0L * s[1] - s[2] - s[3];
    x -= floorl(x);
    return (uint32_t)(x * 4294967296.0L);
}

// Eight hexadecimal digits of RET / 2^(32 (len - 1)) from fractional
// position p (1-based): RET's top limb is the integer part.
uint32_t ret_hex8(const std::vector<uint32_t> &ret, uint64_t p) {
    uint32_t v = 0;
    for (uint64_t q = p; q < p + 8; ++q) {
        const uint64_t limb = ret.size() - 2 - (q - 1) / 8;
        const uint32_t nib = (ret[limb] >> (28 - 4 * ((q - 1) % 8))) & 15u;
        v = (v << 4) | nib;
    }
    return v;
}

bool check(hipError_t e, const char *what) {
    if (e == hipSuccess) return true;
    fprintf(stderr, "%s failed: %s\n", what, hipGetErrorString(e));
    return false;
}

// pi_bs_gpu.cu:2258-2325: one line per stage, totals per kind, then the
// stats line with the accounted (sum of stage) time.
void print_stage_profile(BatchMPContext *ctx, const GpupiPiResult &res, int ws_mb) {
    const int n = gpupi_stage_profile_query(ctx, nullptr, 0);
    std::vector<GpupiStageEntry> e(n > 0 ? n : 0);
    if (n > 0) gpupi_stage_profile_query(ctx, e.data(), n);
    static const char *const kinds[] = {"leaf", "mul", "mul_small", "shift", "add", "sub", "compact",
                                        "exactdiv", "decimal_powers", "decimal_convert"};
    double total[10] = {}, ntt_ms = 0.0;
    for (int i = 0; i < n; ++i) {
        for (int k = 0; k < 10; ++k)
            if (!strcmp(e[i].kind, kinds[k])) total[k] += e[i].elapsed_ms;
        if (!strcmp(e[i].kind, "mul") && e[i].uses_ntt) ntt_ms += e[i].elapsed_ms;
        printf("stage[%03d] kind=%9s label=%18s digits=%8u batch=%8u lhs_len=%9u rhs_len=%9u kernel_ms=%6.3f "
               "uses_ntt=%d\n",
               i, e[i].kind, e[i].label, e[i].n1, e[i].batch_size, e[i].lhs_length, e[i].rhs_length,
               e[i].elapsed_ms, e[i].uses_ntt);
    }
    printf("leaf_total_ms=%.3f mul_total_ms=%.3f mul_small_total_ms=%.3f shift_total_ms=%.3f add_total_ms=%.3f "
           "sub_total_ms=%.3f compact_total_ms=%.3f exactdiv_total_ms=%.3f decimal_powers_total_ms=%.3f "
           "decimal_convert_total_ms=%.3f ntt_total_ms=%.3f\n",
           total[0], total[1], total[2], total[3], total[4], total[5], total[6], total[7], total[8], total[9],
           ntt_ms);
    double accounted = 0.0;
    for (double t : total) accounted += t;
    printf("target_digits=%u convert_digits=%u RET_len=%u elapsed_ms=%.3f accounted_ms=%.3f workspace_max=%dMB\n",
           res.requested_digits, res.convert_digits, res.ret_length, res.elapsed_ms, accounted, ws_mb);
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 2) {
        usage(argv[0]);
        return 1;
    }
    char *end = nullptr;
    const long parsed = strtol(argv[1], &end, 10);
    if (end == argv[1] || *end != '\0' || parsed < 0 || parsed > 2000000000l) {
        usage(argv[0]);
        return 1;
    }
    const uint32_t requested = (uint32_t)parsed;
    bool benchmark = false, profile = false, full = false;
    int convention = GPUPI_PREC_REQUESTED;
    int mode_flags = 0;
    const char *output = nullptr;
    std::vector<uint64_t> bbp_positions;
    for (int i = 2; i < argc; ++i) {
        if (!strcmp(argv[i], "--benchmark")) benchmark = true, ++mode_flags;
        else if (!strcmp(argv[i], "--profile-stages")) profile = true, ++mode_flags;
        else if (!strcmp(argv[i], "--full-print")) full = true, ++mode_flags;
        else if (!strcmp(argv[i], "--readme-prec")) convention = GPUPI_PREC_ROUNDED;
        else if (!strcmp(argv[i], "--output") && i + 1 < argc) output = argv[++i];
        else if (!strcmp(argv[i], "--verify-bbp") && i + 1 < argc) {
            char *e = nullptr;
            const long long pos = strtoll(argv[++i], &e, 10);
            // 16^e mod (8k + j) needs 8n + 6 < 2^32
            if (e == argv[i] || *e != '\0' || pos < 1 || pos > 500000000ll) {
                usage(argv[0]);
                return 1;
            }
            bbp_positions.push_back((uint64_t)pos);
        } else {
            usage(argv[0]);
            return 1;
        }
    }
    if (mode_flags > 1 || (output != nullptr && mode_flags > 0)) {
        usage(argv[0]);
        return 1;
    }
    BatchMPContext *ctx = batch_mp_init();
    if (ctx == nullptr) {
        fprintf(stderr, "batch_mp_init failed\n");
        return 1;
    }
    if (profile && !check(gpupi_stage_profile_enable(ctx, 1), "gpupi_stage_profile_enable")) {
        batch_mp_destroy(ctx);
        return 1;
    }
    GpupiPiResult res{};
    if (!check(gpupi_pi_run(ctx, requested, convention, &res), "gpupi_pi_run")) {
        batch_mp_destroy(ctx);
        return 1;
    }
    std::vector<uint32_t> ret(res.ret_length);
    std::vector<char> digits(requested);
    bool ok = check(hipMemcpy(ret.data(), res.ret, ret.size() * 4, hipMemcpyDeviceToHost), "hipMemcpy(RET)");
    if (ok && requested > 0)
        ok = check(hipMemcpy(digits.data(), res.digits, requested, hipMemcpyDeviceToHost), "hipMemcpy(digits)");
    if (!ok) {
        gpupi_pi_free(&res);
        batch_mp_destroy(ctx);
        return 1;
    }
    const int ws_mb = (int)(batch_mp_workspace_size(ctx) / (1024 * 1024));
    if (profile) {
        print_stage_profile(ctx, res, ws_mb);
        print_hex_preview("RET", ret);
        if (requested > 0) printf("DEC_preview_hash = %u\n", decimal_hash(digits.data(), std::min(requested, 4096u)));
    } else if (benchmark) {
        printf("target_digits=%u convert_digits=%u RET_len=%u elapsed_ms=%.3f workspace_max=%dMB\n", requested,
               res.convert_digits, res.ret_length, res.elapsed_ms, ws_mb);
        print_hex_preview("RET", ret);
        if (requested > 0) {
            const uint32_t pre = std::min(requested, 63u), suf = std::min(requested, 64u);
            printf("DEC = 3%.*s...%.*s\n", (int)pre, digits.data(), (int)suf, digits.data() + (requested - suf));
        } else {
            printf("DEC = 3\n");
        }
    } else {
        fprintf(stderr, "target_digits=%u convert_digits=%u RET_len=%u elapsed_ms=%.3f\n", requested,
                res.convert_digits, res.ret_length, res.elapsed_ms);
        if (full) {
            printf("RET = ");
            if (ret.empty()) printf("0");
            for (int64_t i = (int64_t)ret.size() - 1; i >= 0; --i)
                printf(i == (int64_t)ret.size() - 1 ? "%x" : "%08x", ret[i]);
            printf("\nDEC = 3");
            fwrite(digits.data(), 1, requested, stdout);
            printf("\n");
        } else if (output != nullptr) {
            FILE *f = fopen(output, "wb");
            const bool wrote = f != nullptr && fputs("3.", f) >= 0 &&
                               fwrite(digits.data(), 1, requested, f) == requested && fputc('\n', f) != EOF;
            if (f == nullptr || fclose(f) != 0 || !wrote) {
                fprintf(stderr, "could not write %s\n", output);
                ok = false;
            }
        } else {
            fputc('3', stdout);
            fwrite(digits.data(), 1, requested, stdout);
            fputc('\n', stdout);
        }
    }
    int status = ok ? 0 : 1;
    // the last two limbs of RET carry the truncation and inverse rounding
    const uint64_t exact_hex = ret.size() >= 4 ? 8ull * (ret.size() - 3) : 0;
    for (uint64_t p : bbp_positions) {
        if (p + 7 > exact_hex) {
            fprintf(stderr, "BBP position %llu is beyond RET's exact hexadecimal digits (%llu)\n",
                    (unsigned long long)p, (unsigned long long)exact_hex);
            status = status ? status : 1;
            continue;
        }
        const uint32_t want = bbp_hex8(p - 1), got = ret_hex8(ret, p);
        // long double carries ~64 bits; compare the leading 6 hex digits
        const bool match = (want >> 8) == (got >> 8);
        fprintf(stderr, "BBP hex digits at position %llu: formula %08x RET %08x %s\n", (unsigned long long)p, want,
                got, match ? "match" : "MISMATCH");
        if (!match) status = 2;
    }
    gpupi_pi_free(&res);
    batch_mp_destroy(ctx);
    return status;
}
