"""NTT plan tuner (no counterpart in the reference, which hard-codes its
pass schedules for one GPU, batch_mul_ntt.cu:1810-1915, and its schoolbook
limit, batch_mul_naive.h:5).

For each transform size 2^k it times batched products (N * 2^k = 2^total
coefficients, operands of 2^(k-1) limbs) under every level split the
kernels cover (gpupi_ntt_set_split), checks that every split gives the same
product, and keeps the fastest; it also times the schoolbook and NTT paths
around the crossover.  The library's built-in plans are the result of this
tuner on MI355X.

    python -m gpupi_amd.tune --ks 13-28 [--shared] [--apply] [--json table.json]

Results are exact for every split; only speed changes."""
import argparse
import json
import sys
import time

import numpy as np

import gpupi_amd as g


def splits(k: int):
    """Every (m0, m1) the kernels cover for 2^k points (m1 = 0: two levels)."""
    out = []
    for m0 in range(5, 10):
        for m1 in ([0] if k <= 21 else range(5, 10)):
            rest = k - m0 - m1
            if 6 <= rest <= 12:
                out.append((m0, m1))
    return out


class Operands:
    """Random operands of 2^(total-1) limbs each and room for the products
    (a second product buffer with shared=True)."""

    def __init__(self, total_log2: int, seed: int = 1, shared: bool = False):
        rng = np.random.default_rng(seed)
        half = 1 << (total_log2 - 1)
        a = rng.integers(0, 1 << 32, size=half, dtype=np.uint64).astype(np.uint32)
        self.half = half
        self.ba, _ = g.upload_rows(a[None, :])
        self.bb, _ = g.upload_rows(a[None, ::-1].copy())
        self.bc = g.DeviceBuffer(4 * 2 * half)
        self.bc2 = g.DeviceBuffer(4 * 2 * half) if shared else None

    def arrays(self, L: int):
        N = self.half // L
        return (g.BatchMPArray(self.ba.ptr, L, N, L), g.BatchMPArray(self.bb.ptr, L, N, L),
                g.BatchMPArray(self.bc.ptr, 2 * L, N, 2 * L))

    def shared_arrays(self, L: int):
        """A1, A2, B, C1, C2 of the binary-splitting tree's pattern (C1 = A1 B,
        C2 = A2 B): two products of N rows sharing B."""
        A, B, C = self.arrays(L)
        return A, B, A, C, g.BatchMPArray(self.bc2.ptr, 2 * L, A.batch_size, 2 * L)


def time_mul(ctx, A, B, C, reps: int) -> float:
    g.check(ctx.mul(A, B, C), "mul")
    g.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        g.check(ctx.mul(A, B, C), "mul")
    g.synchronize()
    return (time.perf_counter() - t0) * 1e3 / reps


def time_mul_shared(ctx, arrays, reps: int) -> float:
    g.check(ctx.mul_shared(*arrays), "mul_shared")
    g.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        g.check(ctx.mul_shared(*arrays), "mul_shared")
    g.synchronize()
    return (time.perf_counter() - t0) * 1e3 / reps


def tune_splits(ctx, ops: Operands, k: int, reps: int = 3, log=None, shared: bool = False):
    """[(levels, ms)] for every split of 2^k, fastest first; raises if two
    splits disagree on the product.  shared: time pairs of products sharing
    an operand (the binary-splitting tree's launches) instead of single
    products."""
    A, B, C = ops.arrays(1 << (k - 1))
    sh = ops.shared_arrays(1 << (k - 1)) if shared else None
    ref = None
    res = []
    try:
        for m0, m1 in splits(k):
            g.ntt_set_split(k, m0, m1)
            ms = time_mul_shared(ctx, sh, reps) if shared else time_mul(ctx, A, B, C, reps)
            head = ops.bc.download(np.uint32, count=min(1 << 16, 2 * A.length * A.batch_size))
            if ref is None:
                ref = head
            elif not np.array_equal(head, ref):
                raise RuntimeError(f"split {g.ntt_get_split(k)} changed the product")
            res.append((g.ntt_get_split(k), ms))
            if log:
                log(f"2^{k} levels {g.ntt_get_split(k)}: {ms:.3f} ms")
    finally:
        g.ntt_set_split(k, 0)
    return sorted(res, key=lambda r: r[1])


def tune_crossover(ctx, ops: Operands, totals=(128, 192, 256, 320, 384, 512, 768, 1024), reps: int = 5):
    """Smallest L_a + L_b (of the candidates) from which the NTT is faster
    than the schoolbook kernels on full batches."""
    best = None
    for t in totals:
        A, B, C = ops.arrays(t // 2)
        g.check(ctx.set_ntt_threshold(1 << 30))
        school = time_mul(ctx, A, B, C, reps)
        g.check(ctx.set_ntt_threshold(2))
        ntt = time_mul(ctx, A, B, C, reps)
        if ntt < school and best is None:
            best = t
        elif ntt >= school:
            best = None
    return best


def apply(table: dict) -> None:
    """table: {k: [m0, m1, ...levels]} as written by main(); sets the splits."""
    for k, levels in table.items():
        k = int(k)
        if len(levels) >= 2:
            g.ntt_set_split(k, levels[0], levels[1] if len(levels) == 3 else 0)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--ks", default="13-28", help="transform sizes, e.g. 13-28 or 20,24")
    ap.add_argument("--total-log2", type=int, default=28, help="coefficients per timed call (log2)")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--shared", action="store_true",
                    help="time pairs of products sharing an operand (the tree's launches)")
    ap.add_argument("--crossover", action="store_true", help="also time the schoolbook/NTT crossover")
    ap.add_argument("--apply", action="store_true", help="keep the fastest splits for this process")
    ap.add_argument("--json", help="write the table here")
    args = ap.parse_args(argv)
    if "-" in args.ks:
        lo, hi = (int(x) for x in args.ks.split("-"))
        ks = list(range(lo, hi + 1))
    else:
        ks = [int(x) for x in args.ks.split(",")]
    ks = [k for k in ks if 13 <= k <= min(30, args.total_log2)]
    ctx = g.Context()
    ops = Operands(args.total_log2, shared=args.shared)
    table, report = {}, {}
    for k in ks:
        res = tune_splits(ctx, ops, k, args.reps, log=lambda s: print(s, file=sys.stderr, flush=True),
                          shared=args.shared)
        table[k] = list(res[0][0])
        report[k] = {"best": list(res[0][0]), "best_ms": round(res[0][1], 3),
                     "builtin": list(g.ntt_get_split(k)),
                     "builtin_ms": round(next(ms for lv, ms in res if lv == g.ntt_get_split(k)), 3)}
    out = {"splits": report}
    if args.crossover:
        out["crossover_total_limbs"] = tune_crossover(ctx, ops)
    ctx.close()
    if args.apply:
        apply(table)
    if args.json:
        with open(args.json, "w") as f:
            json.dump({"table": table, **out}, f, indent=1)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
