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

With --launch it tunes the launch shapes instead (tile pairing per column
level kind, the XCD chunk of the block order, the fused group stage's groups
per block): each candidate times batched products at three transform sizes
(2^22, 2^25, 2^28 points, 2^28 coefficients each) or the group stage of a
10^7-digit run (65536 groups), and must leave the same products and rows
as the built-in shape.

    python -m gpupi_amd.tune --launch [--apply] [--json launch.json]

Results are exact for every split and shape; only speed changes."""
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


def _products_digest(ops: Operands, A, B, C) -> np.ndarray:
    return ops.bc.download(np.uint32, count=min(1 << 18, C.length * C.batch_size))


def tune_launch(ctx, ops: Operands, ks=(22, 25, 28), reps: int = 3, log=None) -> dict:
    """Times every tile-pairing mask, XCD chunk and group-stage shape; each
    must give the built-in shape's products / rows.  Restores the built-in
    shapes and returns {"pair_mask": [...], "xcd_chunk_log2": [...],
    "group_blocks": [...]} of (value, ms) fastest first, plus the built-ins."""
    arrays = [ops.arrays(1 << (k - 1)) for k in ks]
    builtin_launch, builtin_gpb = g.ntt_get_launch(), g.group_stage_get_shape()

    def products_ms():
        ms, digests = 0.0, []
        for A, B, C in arrays:
            ms += time_mul(ctx, A, B, C, reps)
            digests.append(_products_digest(ops, A, B, C))
        return ms, digests

    def check(ref, got, what):
        if any(not np.array_equal(a, b) for a, b in zip(ref, got)):
            raise RuntimeError(f"{what} changed the products")

    out = {"builtin": {"pair_mask": builtin_launch[0], "xcd_chunk_log2": builtin_launch[1],
                       "group_blocks": builtin_gpb}}
    try:
        _, ref = products_ms()
        res = []
        for mask in range(16):
            g.ntt_set_launch(mask, builtin_launch[1])
            ms, got = products_ms()
            check(ref, got, f"pair mask {mask}")
            res.append((mask, ms))
            if log:
                log(f"pair mask {mask:2d}: {ms:.3f} ms")
        out["pair_mask"] = sorted(res, key=lambda r: r[1])
        g.ntt_set_launch(builtin_launch[0], builtin_launch[1])
        res = []
        for c in range(0, 9):
            g.ntt_set_launch(builtin_launch[0], c)
            ms, got = products_ms()
            check(ref, got, f"XCD chunk 2^{c}")
            res.append((c, ms))
            if log:
                log(f"XCD chunk 2^{c}: {ms:.3f} ms")
        out["xcd_chunk_log2"] = sorted(res, key=lambda r: r[1])
        g.ntt_set_launch(builtin_launch[0], builtin_launch[1])
        res, ref_rows = [], None
        stage = g.LeafStage(ctx, 65536, g.LEAF_FUSED)
        try:
            for gpb in (1, 2, 4, 8, 16):
                g.group_stage_set_shape(gpb)
                stage.run()
                g.synchronize()
                t0 = time.perf_counter()
                for _ in range(reps * 4):
                    P, Q, R = stage.run()
                g.synchronize()
                ms = (time.perf_counter() - t0) * 1e3 / (reps * 4)
                rows = [g.rows_of(x) for x in (P, Q, R)]
                if ref_rows is None:
                    ref_rows = rows
                elif any(not np.array_equal(a, b) for a, b in zip(ref_rows, rows)):
                    raise RuntimeError(f"group shape {gpb} changed the rows")
                res.append((gpb, ms))
                if log:
                    log(f"group stage {gpb} groups per block: {ms:.3f} ms")
        finally:
            stage.close()
        out["group_blocks"] = sorted(res, key=lambda r: r[1])
    finally:
        g.ntt_set_launch(*builtin_launch)
        g.group_stage_set_shape(builtin_gpb)
    return out


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
    ap.add_argument("--launch", action="store_true",
                    help="tune the launch shapes (tile pairing, XCD chunk, group-stage blocks) instead of splits")
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
    if args.launch:
        res = tune_launch(ctx, ops, reps=args.reps, log=lambda s: print(s, file=sys.stderr, flush=True))
        ctx.close()
        best = {"pair_mask": res["pair_mask"][0][0], "xcd_chunk_log2": res["xcd_chunk_log2"][0][0],
                "group_blocks": res["group_blocks"][0][0]}
        if args.apply:
            g.ntt_set_launch(best["pair_mask"], best["xcd_chunk_log2"])
            g.group_stage_set_shape(best["group_blocks"])
        out = {"best": best, **{k: [[v, round(ms, 3)] for v, ms in res[k]]
                                for k in ("pair_mask", "xcd_chunk_log2", "group_blocks")},
               "builtin": res["builtin"]}
        if args.json:
            with open(args.json, "w") as f:
                json.dump(out, f, indent=1)
        print(json.dumps(out))
        return 0
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
