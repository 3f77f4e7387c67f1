// CPU check of the NTT's binary64 arithmetic (gpupi_amd/csrc/ntt_arith.hpp,
// the same header the kernels include): every butterfly kind, at every
// radix, for both primes, on inputs at and inside its bounds, must give
// exactly the modular DIF/DIT of its inputs (checked with 128-bit integer
// arithmetic) and outputs within the plan's bound.  Also checks the
// pointwise, column-twiddle and 1/K products at their bounds.
// Exit status 0 = all checks passed.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>

#include "ntt_arith.hpp"

using namespace gpupi::ntt;

static int fails = 0;
#define EXPECT(c, ...) do { if (!(c)) { if (fails++ < 20) { printf("FAIL: " __VA_ARGS__); printf("\n"); } } } while (0)

static const uint64_t kPrimes[2] = {562948879679489ull, 562941363486721ull};
static const uint64_t kGens[2] = {3, 13};

static uint64_t mulmod(uint64_t a, uint64_t b, uint64_t p) { return (uint64_t)((unsigned __int128)a * b % p); }
static uint64_t powmod(uint64_t b, uint64_t e, uint64_t p) {
    uint64_t r = 1;
    b %= p;
    while (e) {
        if (e & 1) r = mulmod(r, b, p);
        b = mulmod(b, b, p);
        e >>= 1;
    }
    return r;
}
static double centred(uint64_t v, uint64_t p) { return v > p / 2 ? -(double)(p - v) : (double)v; }
static uint64_t canon(double v, uint64_t p) {
    const int64_t x = (int64_t)v;
    const int64_t r = x % (int64_t)p;
    return (uint64_t)(r < 0 ? r + (int64_t)p : r);
}

struct Ctx {
    int P;
    uint64_t p;
    double w16[8], w16i[8];
    uint64_t u16[8], u16i[8];
};

static Ctx make_ctx(int P) {
    Ctx c{};
    c.P = P;
    c.p = kPrimes[P];
    const uint64_t w30 = powmod(kGens[P], (c.p - 1) >> 30, c.p);
    const uint64_t w16 = powmod(w30, 1ull << 26, c.p), w16i = powmod(w16, c.p - 2, c.p);
    for (int e = 0; e < 8; ++e) {
        c.u16[e] = powmod(w16, e, c.p);
        c.u16i[e] = powmod(w16i, e, c.p);
        c.w16[e] = centred(c.u16[e], c.p);
        c.w16i[e] = centred(c.u16i[e], c.p);
    }
    return c;
}

// exact radix-R DIF / DIT on canonical residues, same stage order as the kernels
static void dif_exact(uint64_t *v, int R, const Ctx &c) {
    for (int h = R / 2; h >= 1; h /= 2)
        for (int j = 0; j < R; ++j) {
            if (j & h) continue;
            const int e = (j & (h - 1)) * (8 / h);
            const uint64_t a = v[j], b = v[j + h];
            v[j] = (a + b) % c.p;
            v[j + h] = mulmod((a + c.p - b) % c.p, c.u16[e], c.p);
        }
}
static void dit_exact(uint64_t *v, int R, const Ctx &c) {
    for (int h = 1; h < R; h *= 2)
        for (int j = 0; j < R; ++j) {
            if (j & h) continue;
            const int e = (j & (h - 1)) * (8 / h);
            const uint64_t t = mulmod(v[j + h], c.u16i[e], c.p), a = v[j];
            v[j] = (a + t) % c.p;
            v[j + h] = (a + c.p - t) % c.p;
        }
}

static double draw(std::mt19937_64 &rng, double bound, uint64_t p, int mode) {
    const double lim = std::floor(bound * (double)p);
    switch (mode) {
        case 0: return lim;
        case 1: return -lim;
        case 2: return (rng() & 1) ? lim : -lim;
        default: {
            const double u = (double)(rng() >> 11) * (1.0 / 9007199254740992.0);  // [0, 1)
            return std::floor((2.0 * u - 1.0) * lim);
        }
    }
}

template <int KIND, int R>
static void check_kind(const Ctx &c, bool inverse, double b0, double bj) {
    constexpr RedPlan plan = PlanOf<KIND, R>::value;
    std::mt19937_64 rng(1000 + KIND * 16 + R + 7 * c.P);
    const Mod m = mod_of(c.P);
    // every sign pattern at the largest magnitudes, then random values
    const int patterns = 1 << R;
    for (int it = 0; it < patterns + 4000; ++it) {
        double v[16];
        uint64_t x[16];
        for (int j = 0; j < R; ++j) {
            const double b = j == 0 ? b0 : bj;
            v[j] = it < patterns ? draw(rng, b, c.p, (it >> j) & 1) : draw(rng, b, c.p, 3);
            x[j] = canon(v[j], c.p);
        }
        if (inverse) {
            dit<KIND, R>(v, c.w16i, m);
            dit_exact(x, R, c);
        } else {
            dif<KIND, R>(v, c.w16, m);
            dif_exact(x, R, c);
        }
        for (int j = 0; j < R; ++j) {
            EXPECT(v[j] == std::floor(v[j]), "non-integer kind=%d R=%d P=%d", KIND, R, c.P);
            EXPECT(std::fabs(v[j]) <= plan.out * (double)c.p, "bound kind=%d R=%d P=%d j=%d |v|/p=%.4f > %.4f",
                   KIND, R, c.P, j, std::fabs(v[j]) / (double)c.p, plan.out);
            EXPECT(canon(v[j], c.p) == x[j], "value kind=%d R=%d P=%d j=%d it=%d", KIND, R, c.P, j, it);
        }
    }
}

template <int R>
static void check_radix(const Ctx &c) {
    check_kind<kDifMid, R>(c, false, kBoundIn, kBoundIn);
    check_kind<kDifToTwiddle, R>(c, false, kBoundIn, kBoundIn);
    check_kind<kDifToPointA, R>(c, false, kBoundIn, kBoundIn);
    check_kind<kDifToPointB, R>(c, false, kBoundIn, kBoundIn);
    check_kind<kDitFirst, R>(c, true, kBoundIn, kBoundIn);
    check_kind<kDitMid, R>(c, true, kBoundInvMid, mul_bound(kBoundInvMid, kTwBound));
    check_kind<kDitLast, R>(c, true, kBoundInvMid, mul_bound(kBoundInvMid, kTwBound));
    check_kind<kDitOnly, R>(c, true, kBoundIn, kBoundIn);
    check_kind<kDitToGarner, R>(c, true, kBoundInvMid, mul_bound(kBoundInvMid, kTwBound));
}

// a * b at the stated bounds: exact mod p and within mul_bound
static void check_product(const Ctx &c, double ba, double bb, const char *what) {
    std::mt19937_64 rng(77 + c.P);
    const Mod m = mod_of(c.P);
    for (int it = 0; it < 20000; ++it) {
        const int mode = it < 9 ? it : 3;
        const double a = draw(rng, ba, c.p, it < 9 ? mode / 3 : 3);
        const double b = draw(rng, bb, c.p, it < 9 ? mode % 3 : 3);
        const double r = mulp(a, b, m);
        EXPECT(r == std::floor(r), "%s non-integer", what);
        EXPECT(std::fabs(r) <= mul_bound(ba, bb) * (double)c.p, "%s bound |r|/p=%.4f", what, std::fabs(r) / (double)c.p);
        EXPECT(canon(r, c.p) == mulmod(canon(a, c.p), canon(b, c.p), c.p), "%s value P=%d", what, c.P);
    }
}

int main() {
    for (int P = 0; P < 2; ++P) {
        const Ctx c = make_ctx(P);
        EXPECT(c.p == (uint64_t)(P ? kP1 : kP0), "prime constants");
        EXPECT(powmod(c.u16[1], 16, c.p) == 1 && powmod(c.u16[1], 8, c.p) != 1, "w16 order");
        check_radix<16>(c);
        check_radix<8>(c);
        check_radix<4>(c);
        check_radix<2>(c);
        check_product(c, kBoundPointA, kBoundPointB, "pointwise");
        check_product(c, kBoundToTwiddle, kColTwBound, "column twiddle");
        check_product(c, kBoundToTwiddle, kTwBound, "1/K");
        check_product(c, 7.9, kTwBound, "step twiddle");
        check_product(c, kColTwBound, kColTwBound, "twiddle recurrence");
        check_product(c, kTwBound, kTwBound, "table entry times 1/K");
        check_product(c, mul_bound(kTwBound, kTwBound), kTwBound, "column twiddle base with 1/K");
    }
    printf("plans (re-centrings per butterfly, R=16):");
    for (int k = 0; k < kBflyKinds; ++k) printf(" %d", bfly_plan(k, 16).count);
    printf("\n");
    printf(fails ? "bounds check: %d failures\n" : "bounds check: ok\n", fails);
    return fails ? 1 : 0;
}
