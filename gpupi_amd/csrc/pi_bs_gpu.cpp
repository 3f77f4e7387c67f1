// pi_bs_gpu: command line front end of the library, same interface and output
// formats as the reference CLI (pi_bs_gpu.cu:2073-2431):
//
//   pi_bs_gpu <digits> [--benchmark|--profile-stages|--full-print] [--readme-prec]
//
// default: "3" + digits on stdout, the stats line on stderr.  --readme-prec
// (an addition) sizes the precision from the rounded-up digit count, the
// convention under which the reference README's RET length and hash were
// published.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
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
    fprintf(stderr, "Usage: %s <target_digits> [--benchmark|--profile-stages|--full-print] [--readme-prec]\n", prog);
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
    for (int i = 2; i < argc; ++i) {
        if (!strcmp(argv[i], "--benchmark")) benchmark = true, ++mode_flags;
        else if (!strcmp(argv[i], "--profile-stages")) profile = true, ++mode_flags;
        else if (!strcmp(argv[i], "--full-print")) full = true, ++mode_flags;
        else if (!strcmp(argv[i], "--readme-prec")) convention = GPUPI_PREC_ROUNDED;
        else {
            usage(argv[0]);
            return 1;
        }
    }
    if (mode_flags > 1) {
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
        } else {
            fputc('3', stdout);
            fwrite(digits.data(), 1, requested, stdout);
            fputc('\n', stdout);
        }
    }
    gpupi_pi_free(&res);
    batch_mp_destroy(ctx);
    return 0;
}
