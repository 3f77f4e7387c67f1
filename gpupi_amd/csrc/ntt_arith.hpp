// NTT arithmetic in binary64: the modular product, re-centring, the
// radix-R butterflies and their re-centring plans.  Shared by the kernels
// and the CPU check of the value bounds (plain host/device code).
//
// Residues are integers held exactly in doubles, centred around 0.  A value
// is "bounded by b" when |v| <= b p.  The butterflies re-centre a sum only
// where its bound would otherwise break one of these limits (p < 2^49):
//   - exact sums:           bound < 16           (|v| < 2^53)
//   - modular product a*w:  b_a b_w <= 3.99      (|a w / p| < 2^51, so the
//                           rounded quotient is exact); the result is
//                           bounded by 0.5 + b_a b_w / 8 (quotient and
//                           low-part errors of at most b_a b_w / 16 each)
//   - re-centring:          bound 0.5 (+ 2^-10 slack)
// The plan for each butterfly kind is derived at compile time from these
// rules (greedy: re-centre any value above a threshold T, T searched for the
// fewest re-centrings) and checked on the CPU against exact arithmetic.
#pragma once

#include <stdint.h>

#include "ntt_plan.hpp"

namespace gpupi {
namespace ntt {

constexpr double kRoundMagic = 6755399441055744.0;  // 1.5 * 2^52
constexpr double kP0 = 562948879679489.0;          // 524287 * 2^30 + 1
constexpr double kP1 = 562941363486721.0;          // 524280 * 2^30 + 1

struct Mod {
    double p, pinv;
};
GPUPI_HD Mod mod_of(int P) { return P ? Mod{kP1, 1.0 / kP1} : Mod{kP0, 1.0 / kP0}; }

GPUPI_HD double qround(double x, Mod m) { return __builtin_fma(x, m.pinv, kRoundMagic) - kRoundMagic; }
GPUPI_HD double redp(double x, Mod m) { return __builtin_fma(-qround(x, m), m.p, x); }
// a * b mod p, centred: h + l = a b exactly (FMA two-product), q = round(h/p)
GPUPI_HD double mulp(double a, double b, Mod m) {
    const double h = a * b;
    const double l = __builtin_fma(a, b, -h);
    return __builtin_fma(-qround(h, m), m.p, h) + l;
}

constexpr int log2c(int x) { return x <= 1 ? 0 : 1 + log2c(x / 2); }

// ---------------------------------------------------------------- bounds
constexpr double kTwBound = 0.5;             // table twiddles, w16 constants
constexpr double kRedBound = 0.5 + 1.0 / 1024;
constexpr double kSumLimit = 15.9;
constexpr double kMulLimit = 3.99;            // b_a * b_w
constexpr double mul_bound(double ba, double bw) { return 0.5 + ba * bw / 8 + 1.0 / 1024; }

// Transform inputs (limbs, or outputs of a twiddle / pointwise product).
constexpr double kBoundIn = 1.0;
// Outputs of inverse steps that feed another inverse step.
constexpr double kBoundInvMid = 7.4;
// Column twiddles: table entries (0.5) -> base, step (<= 0.532) -> step^2
// and the chained products (<= 0.537).
constexpr double kColTwBound = 0.54;
// Outputs that go into a product with a column twiddle or 1/K.
constexpr double kBoundToTwiddle = 7.3;

// Re-centring plan of one radix-R butterfly: bit j of rs[st] / rd[st] asks to
// re-centre the sum / difference output of butterfly j at stage st.
struct RedPlan {
    uint16_t rs[4];
    uint16_t rd[4];
    int count;
    double out;  // bound of every output
    bool ok;
};

// DIF: stage st pairs (j, j + h), h = R >> (st + 1); v[j] = a + b,
// v[j + h] = (a - b) w16^e.  Inputs bounded by b_in; outputs slot 0 by out0
// and the others by outj.
constexpr RedPlan dif_plan_t(int R, double b_in, double out0, double outj, double T) {
    RedPlan r{};
    double v[16] = {};
    for (int j = 0; j < R; ++j) v[j] = b_in;
    const int S = log2c(R);
    for (int st = 0; st < S; ++st) {
        const int h = R >> (st + 1);
        const bool last = st == S - 1;
        for (int j = 0; j < R; ++j) {
            if (j & h) continue;
            const int e = (j & (h - 1)) * (8 / h);
            double s = v[j] + v[j + h], d = s;
            if (s >= kSumLimit) return r;
            const double ls = last ? (j == 0 ? out0 : outj) : T;
            const double ld = last ? outj : T;
            if (s > ls) {
                s = kRedBound;
                r.rs[st] |= (uint16_t)(1u << j);
                ++r.count;
            }
            if (e != 0) {
                if (d * kTwBound > kMulLimit) return r;
                d = mul_bound(d, kTwBound);
            }
            if (d > ld) {
                d = kRedBound;
                r.rd[st] |= (uint16_t)(1u << j);
                ++r.count;
            }
            v[j] = s;
            v[j + h] = d;
        }
    }
    for (int j = 0; j < R; ++j) r.out = v[j] > r.out ? v[j] : r.out;
    r.ok = true;
    return r;
}

// DIT: stage st pairs (j, j + h), h = 1 << st; t = v[j + h] w16^-e;
// v[j] = v[j] + t, v[j + h] = v[j] - t.  Input slot 0 bounded by b0, the
// others by bj; all outputs by out.
constexpr RedPlan dit_plan_t(int R, double b0, double bj, double out, double T) {
    RedPlan r{};
    double v[16] = {};
    v[0] = b0;
    for (int j = 1; j < R; ++j) v[j] = bj;
    const int S = log2c(R);
    for (int st = 0; st < S; ++st) {
        const int h = 1 << st;
        const bool last = st == S - 1;
        for (int j = 0; j < R; ++j) {
            if (j & h) continue;
            const int e = (j & (h - 1)) * (8 / h);
            double t = v[j + h];
            if (e != 0) {
                if (t * kTwBound > kMulLimit) return r;
                t = mul_bound(t, kTwBound);
            }
            double s = v[j] + t, d = s;
            if (s >= kSumLimit) return r;
            const double lim = last ? out : T;
            if (s > lim) {
                s = kRedBound;
                r.rs[st] |= (uint16_t)(1u << j);
                ++r.count;
            }
            if (d > lim) {
                d = kRedBound;
                r.rd[st] |= (uint16_t)(1u << j);
                ++r.count;
            }
            v[j] = s;
            v[j + h] = d;
        }
    }
    for (int j = 0; j < R; ++j) r.out = v[j] > r.out ? v[j] : r.out;
    r.ok = true;
    return r;
}

// Butterfly kinds (where its inputs come from and where its outputs go).
enum BflyKind : int {
    kDifMid = 0,     // forward step followed by another: slot 0 -> next step, others -> step twiddle
    kDifToTwiddle,   // last forward step of a column level: all -> column twiddle
    kDifToPointA,    // last forward step of A's transform: kept for the pointwise product
    kDifToPointB,    // last forward step of B's transform: multiplied by A's
    kDitFirst,       // first inverse step (inputs: pointwise / twiddle outputs) followed by another
    kDitMid,         // inverse step between two others
    kDitLast,        // last inverse step after another: all -> twiddle or 1/K
    kDitOnly,        // single inverse step (M <= 4)
    kDitToGarner,    // last inverse step after another, outputs re-centred: -> prime recombination
    kBflyKinds
};
constexpr double kBoundPointA = 0.51;   // A's transform
constexpr double kBoundPointB = 7.8;    // B's transform; 7.8 * 0.51 <= 3.99

constexpr RedPlan bfly_plan_t(int kind, int R, double T) {
    switch (kind) {
        case kDifMid: return dif_plan_t(R, kBoundIn, kBoundIn, 7.9, T);
        case kDifToTwiddle: return dif_plan_t(R, kBoundIn, kBoundToTwiddle, kBoundToTwiddle, T);
        case kDifToPointA: return dif_plan_t(R, kBoundIn, kBoundPointA, kBoundPointA, T);
        case kDifToPointB: return dif_plan_t(R, kBoundIn, kBoundPointB, kBoundPointB, T);
        case kDitFirst: return dit_plan_t(R, kBoundIn, kBoundIn, kBoundInvMid, T);
        case kDitMid:
            return dit_plan_t(R, kBoundInvMid, mul_bound(kBoundInvMid, kTwBound), kBoundInvMid, T);
        case kDitLast:
            return dit_plan_t(R, kBoundInvMid, mul_bound(kBoundInvMid, kTwBound), kBoundToTwiddle, T);
        case kDitToGarner:
            return dit_plan_t(R, kBoundInvMid, mul_bound(kBoundInvMid, kTwBound), kRedBound, T);
        default: return dit_plan_t(R, kBoundIn, kBoundIn, kBoundToTwiddle, T);
    }
}

// Fewest re-centrings over thresholds T = 0.5, 0.5625, ..., 7.9375.
constexpr RedPlan bfly_plan(int kind, int R) {
    RedPlan best{};
    best.count = 1 << 30;
    for (int i = 8; i < 128; ++i) {
        const RedPlan r = bfly_plan_t(kind, R, i / 16.0);
        if (r.ok && r.count < best.count) best = r;
    }
    return best;
}

// The step kinds' preconditions and postconditions chain:
static_assert(mul_bound(kTwBound, kTwBound) <= kColTwBound &&
                  mul_bound(mul_bound(kTwBound, kTwBound), mul_bound(kTwBound, kTwBound)) <= kColTwBound &&
                  mul_bound(kColTwBound, kColTwBound) <= kColTwBound,
              "column twiddle recurrence stays bounded");
static_assert(mul_bound(kBoundToTwiddle, kColTwBound) <= kBoundIn, "column twiddle output feeds a transform");
static_assert(kBoundToTwiddle * kColTwBound <= kMulLimit, "column twiddle product");
static_assert(kBoundToTwiddle * kTwBound <= kMulLimit && mul_bound(kBoundToTwiddle, kTwBound) < 1.0,
              "1/K product leaves |v| < p for the prime recombination");
static_assert(mul_bound(7.9, kTwBound) <= kBoundIn, "step twiddle output feeds the next step");
static_assert(kBoundPointA * kBoundPointB <= kMulLimit, "pointwise product");
static_assert(mul_bound(kBoundPointA, kBoundPointB) <= kBoundIn, "pointwise output feeds the inverse");
static_assert(kBoundInvMid * kTwBound <= kMulLimit, "inverse step twiddle");
static_assert(mul_bound(mul_bound(kTwBound, kTwBound), kTwBound) <= kColTwBound,
              "a column twiddle base carrying 1/K (a table entry times 1/K, times the other entry)");
static_assert(kRedBound < 1.0, "re-centred outputs leave |v| < p for the prime recombination");

template <int KIND, int R>
struct PlanOf {
    static constexpr RedPlan value = bfly_plan(KIND, R);
    static_assert(value.ok, "no valid re-centring plan");
};

// In-register radix-R DIF (natural in, bit-reversed out) following plan KIND.
template <int KIND, int R>
GPUPI_HD void dif(double *v, const double *w16, Mod m) {
    constexpr RedPlan kp = PlanOf<KIND, R>::value;
    constexpr int S = log2c(R);
#pragma unroll
    for (int st = 0; st < S; ++st) {
        const int h = R >> (st + 1);
#pragma unroll
        for (int j = 0; j < R; ++j) {
            if (j & h) continue;
            const int e = (j & (h - 1)) * (8 / h);
            const double a = v[j], b = v[j + h];
            const double s = a + b;
            double d = a - b;
            if (e != 0) d = mulp(d, w16[e], m);
            v[j] = ((kp.rs[st] >> j) & 1) ? redp(s, m) : s;
            v[j + h] = ((kp.rd[st] >> j) & 1) ? redp(d, m) : d;
        }
    }
}

// In-register radix-R DIT (bit-reversed in, natural out) following plan KIND.
template <int KIND, int R>
GPUPI_HD void dit(double *v, const double *w16i, Mod m) {
    constexpr RedPlan kp = PlanOf<KIND, R>::value;
    constexpr int S = log2c(R);
#pragma unroll
    for (int st = 0; st < S; ++st) {
        const int h = 1 << st;
#pragma unroll
        for (int j = 0; j < R; ++j) {
            if (j & h) continue;
            const int e = (j & (h - 1)) * (8 / h);
            const double t = (e == 0) ? v[j + h] : mulp(v[j + h], w16i[e], m);
            const double a = v[j];
            const double s = a + t, d = a - t;
            v[j] = ((kp.rs[st] >> j) & 1) ? redp(s, m) : s;
            v[j + h] = ((kp.rd[st] >> j) & 1) ? redp(d, m) : d;
        }
    }
}

}  // namespace ntt
}  // namespace gpupi
