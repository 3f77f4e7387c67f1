"""Benchmark: milliseconds to 10^9 decimal digits of pi on one MI355X, with
the other measured workloads of this path reported in the same JSON line.

Headline step = the reference's timed region (pi_bs_gpu.cu:2137-2193): the
binary-splitting series, sqrt(10005), the inverse and final division, the
5^k power table and the radix conversion to decimal, with the result left in
HBM.  The default workload is 1,000,000,000 digits in the precision
convention of the reference README (RET of 103,847,955 limbs); the last
step's RET hash and decimal head/tail are checked against the published ones.
The steps are timed with kernel statistics off; a few more steps with HIP
events around every NTT launch give the roofline of the dominant kernel; one
step of the plain per-op form (one launch per op, a compaction after each
product, no shared transforms, no block cache) gives batched/plain.

"configs" (unless --headline-only) adds, each with its own roofline and CPU
baseline:
  pi_1e6          pi to 10^6 digits, batched and plain forms, vs the oracle
  ragged_256      256 products of log-uniform lengths in [2^10, 2^24] limbs
                  in one ragged call vs one call per pair
  group_stage_1e7 the 17-term group stage of a 10^7-digit run (65,536
                  groups) in its plain, batched, graph and fused forms
  serving         the serving loop on a Poisson stream of pi and multiply jobs

Printed: one JSON line.  Progress goes to stderr.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import re
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
README_1E9 = {
    "ret_len": 103847955,
    "ret_hash": 354748112,
    "ret_preview": "3243f6a88...d6e80bbbf2a7f738",
    "dec_head": "3141592653589793238462643383279502884197169399375105820974944592",
    "dec_tail": "1990548327874671398682093196353628204612755715171395115275045519",
}


def log(msg: str) -> None:
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def hex_preview(limbs) -> str:
    n = len(limbs)
    top = format(int(limbs[-1]), "x") + format(int(limbs[-2]), "08x")
    low = format(int(limbs[1]), "08x") + format(int(limbs[0]), "08x")
    return f"{top}...{low}" if n > 4 else top


def measured_traffic(profile: str, kernel: str):
    """Calibrated HBM bytes per launch of `kernel`, from a committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary under profiles/."""
    path = os.path.join(ROOT, "profiles", profile)
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) == 4 and parts[0] == "traffic" and parts[1] == kernel:
            return float(parts[3])
    return None


def hbm_roofline(stats, profile: str | None = None) -> dict | None:
    """Roofline of the NTT kernel with the largest summed time: algorithmic
    bytes per launch / average launch time (HIP events on the launch stream)."""
    if not stats:
        return None
    name, launches, total_ms, total_bytes = max(stats, key=lambda s: s[2])
    achieved = total_bytes / (total_ms / 1e3) / 1e9
    r = {"bound": "hbm", "kernel": name, "launches": int(launches), "avg_launch_ms": round(total_ms / launches, 4),
         "algorithmic_bytes_per_launch": round(total_bytes / launches), "achieved": round(achieved, 1),
         "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": None}
    if profile is not None:
        t = measured_traffic(profile, name)
        if t is not None:
            r["traffic"] = t
            r["traffic_source"] = (f"profiles/{profile}: rocprofv3 FETCH_SIZE x2.000 + WRITE_SIZE x1.000 bytes "
                                   "per launch (copy-kernel calibration)")
    return r


def ntt_table(stats, calls: int) -> dict:
    return {s[0]: {"launches": int(s[1]), "ms_per_call": round(s[2] / calls, 3),
                   "GBps": round(s[3] / (s[2] / 1e3) / 1e9, 1),
                   "frac_of_hbm_peak": round(s[3] / (s[2] / 1e3) / 1e9 / HBM_PEAK_GBPS, 3)} for s in stats}


# ------------------------------------------------------------ pi steps

class PiRunner:
    """pi to `digits` digits through gpupi_pi_run on one context."""

    def __init__(self, g, digits: int, conv: int, form: int):
        self.g, self.digits, self.conv = g, digits, conv
        self.lib = g.lib()
        self.ctx = g.Context()
        self.ctx.set_form(form)
        self.last = None

    def step(self):
        g = self.g
        if self.last is not None:
            self.lib.gpupi_pi_free(ctypes.byref(self.last))
        self.last = g.GpupiPiResult()
        g.check(self.lib.gpupi_pi_run(self.ctx.ptr, self.digits, self.conv, ctypes.byref(self.last)), "gpupi_pi_run")

    def timed(self, steps: int) -> float:
        g = self.g
        g.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            self.step()
        g.synchronize()
        return (time.perf_counter() - t0) * 1000.0 / max(steps, 1)

    def ret(self):
        return self.g.download_words(self.last.ret, self.last.ret_length)

    def result(self):
        g = self.g
        ret = self.ret()
        dec = g.download_bytes(self.last.digits, self.digits).decode() if self.digits else ""
        return ret, dec

    def close(self):
        if self.last is not None:
            self.lib.gpupi_pi_free(ctypes.byref(self.last))
            self.last = None
        self.ctx.close()


def host_threads() -> int:
    """Host threads for the CPU baselines: the GPU's share of the host (16
    cores per GPU) or fewer if the machine has fewer."""
    return max(1, min(16, os.cpu_count() or 1))


def run_oracle_cli(*argv) -> str:
    cli = os.path.join(ROOT, "oracle", "pi_oracle")
    if not os.path.exists(cli):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    return subprocess.run([cli, *map(str, argv)], capture_output=True, text=True, check=True).stdout


def cpu_baseline_pi(target: int) -> dict:
    """CPU pi on this host's cores (GMP is single-threaded; the parallelism is
    over independent subtrees / rows).  value: a textbook GMP Chudnovsky
    (binary splitting, sqrt, one division, mpz_get_str) measured at 10^8
    digits on all host threads and scaled to `target` with the growth factor
    measured between 10^7 and 10^8.  Also reported: the oracle's restatement
    of the reference pipeline at 10^7 digits on one thread and on all of
    them."""
    T = host_threads()
    ms = lambda out: float(re.search(r"elapsed_ms=([0-9.]+)", out).group(1))
    classic = {}
    for d in (10_000_000, 100_000_000):
        log(f"CPU baseline: classic GMP Chudnovsky, {d:.0e} digits, {T} threads")
        classic[d] = ms(run_oracle_cli(d, "--classic", "--threads", T))
    growth = classic[100_000_000] / classic[10_000_000]
    value = classic[100_000_000] * growth ** math.log10(target / 100_000_000)
    log("CPU baseline: restated pipeline, 1e7 digits")
    restated_1 = ms(run_oracle_cli(10_000_000))
    restated_t = ms(run_oracle_cli(10_000_000, "--threads", T))
    return {"value": round(value, 1), "unit": f"ms to {target:.0e} digits (scaled from 1e8)", "cores": T,
            "kind": "port",
            "sample": (f"textbook GMP Chudnovsky (oracle/pi_oracle --classic: binary splitting with subtrees on "
                       f"{T} threads, mpz_sqrt, mpz_tdiv_q, mpz_get_str) measured {classic[10_000_000]:.0f} ms at "
                       f"1e7 and {classic[100_000_000]:.0f} ms at 1e8 digits; scaled to {target:.0e} by the measured "
                       f"growth {growth:.2f}x per decade"),
            "measured_ms": {"classic_1e7": round(classic[10_000_000], 1), "classic_1e8": round(classic[100_000_000], 1),
                            "restated_1e7_1_thread": round(restated_1, 1),
                            f"restated_1e7_{T}_threads": round(restated_t, 1)}}


def bench_headline(args, g) -> dict:
    conv = g.PREC_ROUNDED if args.prec == "rounded" else g.PREC_REQUESTED
    readme = args.digits == 1_000_000_000 and conv == g.PREC_ROUNDED
    run = PiRunner(g, args.digits, conv, g.FORM_BATCHED)
    log(f"pi {args.digits} digits: {args.warmup} warmup + {args.steps} timed steps")
    for _ in range(args.warmup):
        run.step()
    ms_per_step = run.timed(args.steps)
    ret = run.ret()
    head = "3" + g.download_bytes(run.last.digits, min(63, args.digits)).decode()
    tail_n = min(64, args.digits)
    tail = g.download_bytes(run.last.digits + args.digits - tail_n, tail_n).decode()
    event_ms = run.last.elapsed_ms
    check = {"ret_len": int(len(ret)), "ret_hash": g.ret_mod_hash(ret), "ret_preview": hex_preview(ret),
             "dec_head": head, "dec_tail": tail}
    bit_exact = all(check[k] == README_1E9[k] for k in README_1E9) if readme else None

    # roofline steps: HIP events around every NTT launch
    rsteps = max(1, min(3, args.steps))
    log(f"roofline: {rsteps} steps with kernel events")
    g.kernel_stats_reset(True)
    ms_events = run.timed(rsteps)
    stats = g.kernel_stats()
    g.kernel_stats_reset(False)
    roofline = hbm_roofline(stats, "pi_1e9_ntt_traffic.md" if readme else None)
    if roofline is not None:
        roofline["steps"] = rsteps
        roofline["ms_per_step_with_kernel_events"] = round(ms_events, 3)
    ntt = {s[0]: {"launches": int(s[1]), "ms_per_step": round(s[2] / rsteps, 2),
                  "GBps": round(s[3] / (s[2] / 1e3) / 1e9, 1)} for s in stats}
    run.ctx.cache_trim()
    run.close()

    plain = None
    if not args.no_plain:
        log("plain form: 1 step")
        prun = PiRunner(g, args.digits, conv, g.FORM_PLAIN)
        plain_ms = prun.timed(1)
        pret = prun.ret()
        prun.close()
        plain = {"ms_per_step": round(plain_ms, 3), "steps": 1,
                 "speedup_batched_vs_plain": round(plain_ms / ms_per_step, 3),
                 "identical_ret": bool(len(pret) == len(ret) and g.ret_mod_hash(pret) == check["ret_hash"]
                                       and (pret == ret).all())}
    del ret

    cpu = None
    if not args.no_cpu_baseline:
        log("CPU baseline (oracle)")
        cpu = cpu_baseline_pi(args.digits)
    return {
        "metric": "ms to 1e9 pi digits (bit-exact)" if args.digits == 1_000_000_000 else f"ms to {args.digits} pi digits",
        "value": round(ms_per_step, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": False,
        "scaling": "strong",
        # the reference publishes 4623 ms at 10^9 digits (RTX 5090, README.md:8); < 1 is faster
        "vs_baseline": round(ms_per_step / 4623.0, 4) if args.digits == 1_000_000_000 else None,
        "dtype": "u32 limbs, f64 NTT residues",
        "data": "none (pi has no input; the computation is deterministic)",
        "config": {"workload": f"pi to {args.digits} decimal digits, single GPU, "
                               f"{'README' if conv == g.PREC_ROUNDED else 'requested-digit'} precision convention",
                   "digits": args.digits, "prec_convention": args.prec, "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "bit_exact_vs_reference_readme": bit_exact,
        "check": check,
        "event_ms_last_step": round(event_ms, 3),
        "plain_form": plain,
        "ntt_kernels": ntt,
    }


# -------------------------------------------------------- pi at 10^6

def bench_pi_small(args, g, digits: int = 1_000_000) -> dict:
    """pi to `digits` digits (requested-digit convention, as the reference's
    pi_bs_gpu runs it) in the batched and the plain form; RET limbs and
    digits checked against the oracle, which is also the CPU baseline."""
    steps = max(args.steps, 5)
    out, results = {}, {}
    for name, form in (("batched", g.FORM_BATCHED), ("plain", g.FORM_PLAIN)):
        run = PiRunner(g, digits, g.PREC_REQUESTED, form)
        for _ in range(max(args.warmup, 2)):
            run.step()
        out[name] = run.timed(steps)
        results[name] = run.result()
        if name == "batched":
            g.kernel_stats_reset(True)
            run.timed(3)
            stats = g.kernel_stats()
            g.kernel_stats_reset(False)
        run.close()
    same = (results["batched"][1] == results["plain"][1] and
            (results["batched"][0] == results["plain"][0]).all())
    r = {"metric": f"ms to {digits} pi digits", "value": round(out["batched"], 3), "unit": "ms",
         "higher_is_better": False, "steps": steps,
         "plain_ms": round(out["plain"], 3), "speedup_batched_vs_plain": round(out["plain"] / out["batched"], 3),
         "plain_identical": bool(same), "roofline": hbm_roofline(stats), "cpu_baseline": None,
         "bit_exact_vs_oracle": None}
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_ffi as of
        t0 = time.perf_counter()
        oret, odec, _ = of.pi_digits(digits)
        t_cpu = (time.perf_counter() - t0) * 1e3
        r["cpu_baseline"] = {"value": round(t_cpu, 1), "unit": f"ms to {digits} digits", "cores": 1, "kind": "port",
                             "sample": "the whole workload: oracle/liboracle (GMP restatement of the pipeline), "
                                       "one thread"}
        ret, dec = results["batched"]
        r["bit_exact_vs_oracle"] = bool(dec == odec and len(ret) == len(oret) and (ret == oret).all())
    return r


# ------------------------------------------------- 17-term group stage

def bench_leaf(args, g, digits: int = 10_000_000) -> dict:
    """The 17-term group stage of a `digits` run (N groups: the reference's
    N*17*14.3 >= digits rule, pi_bs_gpu.cu:1539) in its four forms: the
    reference's per-op form (a synchronizing compaction after every product),
    the batched form (lengths fixed in advance), the batched form replayed
    from a hipGraph, and the fused one-kernel form (a wave per group in LDS).
    Checks that all leave the same rows; reports ms per run, schoolbook
    multiply-adds per second, and the fused kernel against the integer-VALU
    multiply-add peak."""
    import numpy as np
    N = 1
    while N * 17 * 14.3 < digits:
        N *= 2
    steps = max(args.steps, 5)
    ctx = g.Context()
    forms = {"plain": g.LEAF_PLAIN, "batched": g.LEAF_BATCHED, "graph": g.LEAF_GRAPH, "fused": g.LEAF_FUSED}
    out, rows = {}, {}
    fused_kernel_ms = None
    for name, mode in forms.items():
        st = g.LeafStage(ctx, N, mode)
        for _ in range(max(args.warmup, 2)):
            st.run()
        g.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            res = st.run()
        g.synchronize()
        ms = (time.perf_counter() - t0) * 1000.0 / steps
        rows[name] = [g.rows_of(a) for a in res]
        # work: sum over the stage's products of batch * L_a * L_b (32x32->64
        # multiply-adds), from one profiled run of this form
        ctx.stage_profile(True)
        st.run()
        entries = ctx.stage_entries()
        ctx.stage_profile(False)
        macs = sum(e["batch"] * e["lhs"] * e["rhs"] for e in entries if e["kind"] == "mul")
        if name in ("graph", "fused"):  # one graph launch / one kernel: the batched form's products
            macs = out["batched"]["macs"]
        if name == "fused":
            fused_kernel_ms = sum(e["ms"] for e in entries if e["label"] == "fused_group_stage")
        out[name] = {"ms": round(ms, 4), "macs": int(macs), "gmac_per_s": round(macs / (ms / 1e3) / 1e9, 1),
                     "stages": len(entries)}
        st.close()
    same = all(all(np.array_equal(a, b) for a, b in zip(rows["plain"], rows[k])) for k in ("batched", "graph", "fused"))
    ctx.close()
    dev = g.device_info()
    peak = g.mac_peak() / 1e12
    achieved = out["fused"]["macs"] / (fused_kernel_ms / 1e3) / 1e12
    roofline = {"bound": "valu", "kernel": "group_stage_kernel", "launches": 1,
                "avg_launch_ms": round(fused_kernel_ms, 4), "achieved": round(achieved, 3), "peak": round(peak, 2),
                "unit": "T 32x32->64 MAC/s", "frac": round(achieved / peak, 4), "traffic": None,
                "peak_basis": (f"measured: back-to-back v_mad_u64_u32 on {dev['compute_units']} CUs at "
                               f"{dev['clock_khz'] / 1e6:.2f} GHz (gpupi_calib_mac_peak); one MAC of the group "
                               "stage = one v_mad_u64_u32 plus a 64-bit add and two LDS reads")}
    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_ffi as of
        op, oq, orr = of.OArray(), of.OArray(), of.OArray()
        t0 = time.perf_counter()
        assert of.lib().o_leaf_stage(N, ctypes.byref(op), ctypes.byref(oq), ctypes.byref(orr)) == 0
        t_cpu = (time.perf_counter() - t0) * 1e3
        want = (of.take_rows(op), of.take_rows(oq), of.take_rows(orr))
        exact = all(a.shape == b.shape and np.array_equal(a, b) for a, b in zip(rows["fused"], want))
        cpu = {"value": round(t_cpu, 2), "unit": "ms per group stage", "cores": 1, "kind": "port",
               "sample": f"the whole stage: oracle o_leaf_stage ({N} groups, GMP), one thread",
               "gpu_rows_checked_bit_exact": bool(exact)}
    return {"metric": f"ms per 17-term group stage, {N} groups ({digits} digits)",
            "value": out["fused"]["ms"], "unit": "ms", "higher_is_better": False, "steps": steps,
            "groups": N, "forms": out,
            "speedup_fused_vs_plain": round(out["plain"]["ms"] / out["fused"]["ms"], 2),
            "speedup_graph_vs_plain": round(out["plain"]["ms"] / out["graph"]["ms"], 2),
            "speedup_batched_vs_plain": round(out["plain"]["ms"] / out["batched"]["ms"], 2),
            "identical_rows": same, "lengths": {k: int(v.shape[1]) for k, v in zip("PQR", rows["plain"])},
            "roofline": roofline, "cpu_baseline": cpu}


# ------------------------------------------------- ragged multiply

def ragged_pairs(seed: int, count: int, lo: float, hi: float):
    """count pairs with lengths floor(2^U), U uniform in [lo, hi) (drawn L_a
    then L_b per pair), from numpy's default generator seeded with `seed`."""
    import numpy as np
    rng = np.random.default_rng(seed)
    u = rng.uniform(lo, hi, size=(count, 2))
    return rng, [(int(2 ** a), int(2 ** b)) for a, b in u]


def bench_ragged(args, g) -> dict:
    """256 pairs, lengths log-uniform in [2^10, 2^24] limbs, uniform random
    limbs (seed 1234): one ragged call (gpupi_mul_ragged) against one
    batch_mp_mul call per pair in the reference's plain form (each pair one
    whole product; GPUPI_FORM_PLAIN) and in the batched form (unbalanced
    pairs in pieces), with inputs resident in HBM.  Reports ms for each,
    per-pass NTT GB/s of the ragged form, and the reference's compulsory
    figure 4*sum(L_a+L_b+L_c)/t (batch_mul_ntt_test.cu:491-492)."""
    import numpy as np
    rng, lengths = ragged_pairs(1234, args.pairs, 10.0, 24.0)
    ctx = g.Context()
    ctx_plain = g.Context()
    ctx_plain.set_form(g.FORM_PLAIN)
    bufs, items, uni = [], [], []
    for la, lb in lengths:
        A = rng.integers(0, 1 << 32, size=la, dtype=np.uint64).astype(np.uint32)
        B = rng.integers(0, 1 << 32, size=lb, dtype=np.uint64).astype(np.uint32)
        ba, bb = g.DeviceBuffer.from_array(A), g.DeviceBuffer.from_array(B)
        bc, bu = g.DeviceBuffer(4 * (la + lb)), g.DeviceBuffer(4 * (la + lb))
        bufs.append((ba, bb, bc, bu))
        items.append((ba.ptr, bb.ptr, bc.ptr, la, lb))
        uni.append((g.BatchMPArray(ba.ptr, la, 1, la), g.BatchMPArray(bb.ptr, lb, 1, lb),
                    g.BatchMPArray(bu.ptr, la + lb, 1, la + lb)))
    del A, B
    steps = max(3, min(args.steps, 5))

    def plain():
        for a_, b_, c_ in uni:
            g.check(ctx_plain.mul(a_, b_, c_), "mul")

    def single():
        for a_, b_, c_ in uni:
            g.check(ctx.mul(a_, b_, c_), "mul")

    def ragged():
        g.check(ctx.mul_ragged(items), "gpupi_mul_ragged")

    res = {}
    same_single = True
    for name, fn in (("plain", plain), ("single", single), ("ragged", ragged)):
        for _ in range(max(args.warmup, 1)):
            fn()
        g.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        g.synchronize()
        res[name] = (time.perf_counter() - t0) * 1000.0 / steps
        if name == "plain":
            plain_out = [bu.download(np.uint32) for _, _, _, bu in bufs]
        elif name == "single":
            same_single = all(np.array_equal(bu.download(np.uint32), p_) for (_, _, _, bu), p_ in zip(bufs, plain_out))
            del plain_out
    g.kernel_stats_reset(True)
    ragged()
    g.synchronize()
    stats = g.kernel_stats()
    g.kernel_stats_reset(False)
    same = all(np.array_equal(bc.download(np.uint32), bu.download(np.uint32)) for _, _, bc, bu in bufs)
    limbs = sum(2 * (la + lb) for la, lb in lengths)

    cpu = None
    if not args.no_cpu_baseline:
        # GMP mpn_mul on a bounded sample of the pairs (the smallest ones up
        # to ~10 s of CPU work), each checked against the GPU product
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_ffi as of
        order = sorted(range(len(lengths)), key=lambda i: lengths[i][0] + lengths[i][1])
        done, t_cpu, limbs_cpu, exact = 0, 0.0, 0, True
        for i in order:
            la, lb = lengths[i]
            ba, bb, bc, _ = bufs[i]
            A = ba.download(np.uint32).reshape(1, la)
            B = bb.download(np.uint32).reshape(1, lb)
            hc = of.HostArray.empty(1, la + lb)
            t0 = time.perf_counter()
            assert of.lib().o_mul(of.HostArray(A).oarray(), of.HostArray(B).oarray(), hc.oarray()) == 0
            t_cpu += time.perf_counter() - t0
            exact = exact and np.array_equal(hc.rows()[0], bc.download(np.uint32))
            done += 1
            limbs_cpu += 2 * (la + lb)
            if t_cpu > 10.0:
                break
        cpu = {"value": round(4.0 * limbs_cpu / t_cpu / 1e9, 3), "unit": "GB/s (A+B+ret)", "cores": 1,
               "kind": "port", "sample": f"GMP mpn_mul on the {done} smallest of {len(lengths)} pairs "
                                         f"({limbs_cpu} limbs in+out) in {t_cpu:.1f} s, one thread",
               "gpu_products_checked_bit_exact": exact}
    ctx.close()
    ctx_plain.close()
    same = same and same_single
    return {"metric": f"ms per batched product of {len(lengths)} ragged pairs (one ragged call)",
            "value": round(res["ragged"], 3), "unit": "ms", "higher_is_better": False, "steps": steps,
            "plain_ms": round(res["plain"], 3), "speedup_ragged_vs_plain": round(res["plain"] / res["ragged"], 3),
            "single_calls_batched_form_ms": round(res["single"], 3),
            "speedup_ragged_vs_single_calls": round(res["single"] / res["ragged"], 3),
            "a_b_ret_GBps": round(4.0 * limbs / (res["ragged"] / 1e3) / 1e9, 1),
            "total_limbs_in_out": limbs, "identical_to_plain": same, "ntt_passes": ntt_table(stats, 1),
            "roofline": hbm_roofline(stats), "cpu_baseline": cpu}


# ------------------------------------------------------------ serving

def bench_serve(args, g) -> dict:
    """The serving loop on a Poisson request stream (seed 7): 40% pi jobs of
    floor(10^U(5,8)) digits, 60% multiply jobs of floor(2^U(12,26)) limbs per
    operand with uniform random limbs.  Every job first runs alone (service
    time, reference result), then the stream arrives at 70% of the measured
    sequential capacity; reports p50/p99 latency and throughput, and checks
    that every result in the mix equals its result alone."""
    import hashlib
    import threading
    import numpy as np
    rng = np.random.default_rng(7)
    n = args.jobs
    jobs = []
    for _ in range(n):
        if rng.random() < 0.4:
            jobs.append(("pi", int(10 ** rng.uniform(5, 8))))
        else:
            la, lb = int(2 ** rng.uniform(12, 26)), int(2 ** rng.uniform(12, 26))
            A = rng.integers(0, 1 << 32, size=la, dtype=np.uint64).astype(np.uint32)
            B = rng.integers(0, 1 << 32, size=lb, dtype=np.uint64).astype(np.uint32)
            jobs.append(("mul", (A, B)))
    gaps = rng.exponential(1.0, size=n)  # scaled by the rate below

    def digest(kind, r, data):
        if kind == "pi":
            return (r.ret_length, r.ret_hash, r.dec_hash)
        return hashlib.sha256(data.tobytes()).hexdigest()

    srv = g.Server(budget_bytes=200 * 10 ** 9)
    alone, service = {}, {}
    log(f"serve: {n} jobs generated; running each alone")
    for i, (kind, x) in enumerate(jobs):
        jid = srv.submit_pi(x) if kind == "pi" else srv.submit_mul(*x)
        assert jid >= 0
        (r, data), = srv.results(wait=True)
        assert r.status == 0, (kind, r.status)
        alone[i] = digest(kind, r, data)
        service[i] = r.finish_ms - r.start_ms
    mean_service = sum(service.values()) / n
    rate = 0.7 / (mean_service / 1e3)  # jobs per second
    log(f"serve: alone done (mean service {mean_service:.1f} ms); mix at {rate:.1f} jobs/s")
    arrivals = np.cumsum(gaps / rate)

    results, id2job = {}, {}

    def collect():
        while len(results) < n:
            got = srv.results(wait=True)
            for r, data in got:
                results[r.id] = (r, data)
            if not got:  # nothing in flight yet: the next arrival is pending
                time.sleep(0.0005)

    th = threading.Thread(target=collect)
    th.start()
    t0 = time.perf_counter()
    for i, (kind, x) in enumerate(jobs):
        delay = arrivals[i] - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        jid = srv.submit_pi(x) if kind == "pi" else srv.submit_mul(*x)
        assert jid >= 0
        id2job[jid] = i
    th.join()
    srv.close()
    log("serve: mix done")
    lat = {"pi": [], "mul": []}
    same = True
    first_submit = min(r.submit_ms for r, _ in results.values())
    last_finish = max(r.finish_ms for r, _ in results.values())
    batch_max = pi_batch_max = 0
    for jid, (r, data) in results.items():
        i = id2job[jid]
        kind = jobs[i][0]
        lat[kind].append(r.finish_ms - r.submit_ms)
        same = same and digest(kind, r, data) == alone[i]
        if kind == "mul":
            batch_max = max(batch_max, r.batch_size)
        else:
            pi_batch_max = max(pi_batch_max, r.batch_size)
    allv = np.array(lat["pi"] + lat["mul"])
    pct = lambda v, q: round(float(np.percentile(np.array(v), q)), 2) if v else None

    cpu = None
    if not args.no_cpu_baseline:
        # GMP mpn_mul on the smallest multiply jobs (~10 s of CPU work), each
        # checked against the GPU result of the mix
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_ffi as of
        muls = sorted((i for i, j in enumerate(jobs) if j[0] == "mul"), key=lambda i: jobs[i][1][0].size + jobs[i][1][1].size)
        by_job = {id2job[jid]: data for jid, (r, data) in results.items()}
        t_cpu, done, exact, limbs = 0.0, 0, True, 0
        for i in muls:
            A, B = jobs[i][1]
            hc = of.HostArray.empty(1, A.size + B.size)
            t1 = time.perf_counter()
            assert of.lib().o_mul(of.HostArray(A.reshape(1, -1)).oarray(), of.HostArray(B.reshape(1, -1)).oarray(),
                                  hc.oarray()) == 0
            t_cpu += time.perf_counter() - t1
            exact = exact and np.array_equal(hc.rows()[0], by_job[i])
            done += 1
            limbs += 2 * (A.size + B.size)
            if t_cpu > 10.0:
                break
        cpu = {"value": round(done / t_cpu, 2), "unit": "multiply jobs/s", "cores": 1, "kind": "port",
               "sample": f"GMP mpn_mul on the {done} smallest multiply jobs ({limbs} limbs in+out), one thread",
               "gpu_results_checked_bit_exact": exact}
    return {"metric": f"serving p99 latency, {n} jobs, Poisson arrivals at 70% of sequential capacity",
            "value": pct(allv.tolist(), 99), "unit": "ms", "higher_is_better": False,
            "p50_ms": pct(allv.tolist(), 50), "p99_ms": pct(allv.tolist(), 99),
            "pi": {"jobs": len(lat["pi"]), "p50_ms": pct(lat["pi"], 50), "p99_ms": pct(lat["pi"], 99),
                   "max_jobs_per_shared_tree": pi_batch_max},
            "mul": {"jobs": len(lat["mul"]), "p50_ms": pct(lat["mul"], 50), "p99_ms": pct(lat["mul"], 99),
                    "max_jobs_per_shared_launch": batch_max},
            "throughput_jobs_per_s": round(n / ((last_finish - first_submit) / 1e3), 2),
            "arrival_rate_jobs_per_s": round(rate, 2), "mean_service_alone_ms": round(mean_service, 2),
            "identical_alone_and_in_mix": same, "hbm_budget_gb": 200,
            "roofline": None, "cpu_baseline": cpu}


def bench_serve_burst(args, g) -> dict:
    """A burst of 16 pi jobs of one tree size (floor(10^U(6, 6.27)) digits,
    seed 5: 8192 groups each) submitted at once to a server with 4 pi
    workers, with shared trees (pi_batch_max 8) and without (1).  Each server
    first serves the same burst once (contexts, memory pool), then three
    timed bursts; three servers per mode, alternating: medians over the nine
    timed bursts of each mode.  Every result must be identical in both
    modes, and equal the oracle."""
    import numpy as np
    rng = np.random.default_rng(5)
    digits = [int(10 ** rng.uniform(6.0, 6.27)) for _ in range(16)]
    assert len({g.lib().gpupi_pi_groups_for_digits(d) for d in digits}) == 1

    def burst(batch_max):
        srv = g.Server(pi_workers=4, pi_batch_max=batch_max)
        samples, out = [], None
        try:
            for rep in range(4):
                t0 = time.perf_counter()
                ids = {srv.submit_pi(d, keep_digits=True): d for d in digits}
                res = {}
                while len(res) < len(ids):
                    for r, data in srv.results(wait=True):
                        assert r.status == 0
                        res[r.id] = (r, bytes(data))
                wall = (time.perf_counter() - t0) * 1e3
                got = {ids[i]: (r.ret_length, r.ret_hash, data) for i, (r, data) in res.items()}
                assert out is None or got == out
                out = got
                if rep > 0:  # the first burst warms the contexts and the pool
                    lat = [r.finish_ms - r.submit_ms for r, _ in res.values()]
                    samples.append((wall, float(np.mean(lat)), max(r.batch_size for r, _ in res.values())))
        finally:
            srv.close()
        return samples, out

    runs = {1: [], 8: []}
    outs = {}
    for _ in range(3):
        for bm in (1, 8):
            samples, out = burst(bm)
            runs[bm].extend(samples)
            outs.setdefault(bm, out)
            assert out == outs[bm]
    identical = outs[1] == outs[8]

    def median(v, i):
        return float(np.median([x[i] for x in v]))
    med = {bm: (median(v, 0), median(v, 1)) for bm, v in runs.items()}
    cpu, exact = None, None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_ffi as of
        t_cpu, exact = 0.0, True
        for d in digits:
            t1 = time.perf_counter()
            _, odec, _ = of.pi_digits(d)
            t_cpu += time.perf_counter() - t1
            exact = exact and odec.encode() == outs[8][d][2]
        cpu = {"value": round(len(digits) / t_cpu, 3), "unit": "pi jobs/s", "cores": 1, "kind": "port",
               "sample": f"the whole burst: oracle/liboracle (GMP restatement of the pipeline) on the {len(digits)} "
                         "jobs one after the other, one thread", "gpu_digits_checked_bit_exact": exact}
    return {"metric": "pi jobs/s, burst of 16 jobs of one tree size (1.0-1.9 M digits), 4 pi workers",
            "value": round(len(digits) / (med[8][0] / 1e3), 1), "unit": "jobs/s", "higher_is_better": True,
            "shared_trees": {"makespan_ms": round(med[8][0], 2), "mean_latency_ms": round(med[8][1], 2),
                             "max_jobs_per_tree": max(r[2] for r in runs[8])},
            "one_tree_per_job": {"makespan_ms": round(med[1][0], 2), "mean_latency_ms": round(med[1][1], 2)},
            "speedup_shared_vs_single": round(med[1][0] / med[8][0], 3),
            "runs_ms": {str(bm): [round(r[0], 2) for r in v] for bm, v in runs.items()},
            "digits": [min(digits), max(digits)], "identical_results": identical,
            "roofline": None, "cpu_baseline": cpu}


# ------------------------------------------------------ kernel study

def bench_mul(args, g) -> dict:
    """Times batch_mp_mul on N random pairs of L_a x L_b limbs (inputs
    resident in HBM) and reports per-kernel NTT timings."""
    import numpy as np
    rng = np.random.default_rng(1234)
    ctx = g.Context()
    A = rng.integers(0, 1 << 32, size=(args.n, args.la), dtype=np.uint64).astype(np.uint32)
    B = rng.integers(0, 1 << 32, size=(args.n, args.lb), dtype=np.uint64).astype(np.uint32)
    ba, Aa = g.upload_rows(A)
    bb, Ba = g.upload_rows(B)
    del A, B
    bc = g.DeviceBuffer(4 * args.n * (args.la + args.lb))
    Ca = g.BatchMPArray(bc.ptr, args.la + args.lb, args.n, args.la + args.lb)
    for _ in range(args.warmup):
        g.check(ctx.mul(Aa, Ba, Ca), "mul")
    g.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        g.check(ctx.mul(Aa, Ba, Ca), "mul")
    g.synchronize()
    ms = (time.perf_counter() - t0) * 1000.0 / args.steps
    g.kernel_stats_reset(True)
    g.check(ctx.mul(Aa, Ba, Ca), "mul")
    g.synchronize()
    stats = g.kernel_stats()
    g.kernel_stats_reset(False)
    # the reference's compulsory-traffic figure (batch_mul_ntt_test.cu:491-492)
    abr = 4.0 * args.n * (2 * (args.la + args.lb)) / (ms / 1e3) / 1e9
    ctx.close()
    return {"metric": f"ms per batched product N={args.n} {args.la}x{args.lb} limbs",
            "value": round(ms, 4), "unit": "ms", "steps": args.steps, "warmup": args.warmup,
            "a_b_ret_GBps": round(abr, 1), "ntt_kernels": ntt_table(stats, 1), "roofline": hbm_roofline(stats)}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--digits", type=int, default=1_000_000_000)
    ap.add_argument("--prec", choices=["rounded", "requested"], default="rounded")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="headline: skip the plain-form step")
    ap.add_argument("--headline-only", action="store_true", help="skip the other workloads")
    ap.add_argument("--pairs", type=int, default=256)
    ap.add_argument("--jobs", type=int, default=100)
    ap.add_argument("--workload", choices=["all", "pi", "pi1e6", "mul", "leaf", "ragged", "serve", "burst"],
                    default="all",
                    help="all: the headline run plus every other workload in one line; pi: the headline run; "
                         "pi1e6: pi to 10^6 digits, batched vs plain; mul: one batched product shape "
                         "(kernel study); leaf: the 17-term group stage of a 10^7-digit run in its four forms; "
                         "ragged: --pairs log-uniform pairs in one ragged call vs one call per pair; "
                         "serve: the serving loop on a Poisson stream of --jobs pi and multiply jobs; "
                         "burst: 16 pi jobs of one tree size at once, shared trees vs one tree per job")
    ap.add_argument("--la", type=int, default=1 << 27)
    ap.add_argument("--lb", type=int, default=1 << 27)
    ap.add_argument("--n", type=int, default=1)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: the gpupi_amd pipeline runs on one GPU; --gpus {args.gpus} is not supported "
              "(every stage of it is single-device).", file=sys.stderr)
        return 2

    import gpupi_amd as g

    single = {"pi1e6": bench_pi_small, "mul": bench_mul, "leaf": bench_leaf, "ragged": bench_ragged,
              "serve": bench_serve, "burst": bench_serve_burst}
    if args.workload in single:
        r = single[args.workload](args, g)
        print(json.dumps(r))
        return 0
    line = bench_headline(args, g)
    ok = line["bit_exact_vs_reference_readme"] in (None, True)
    if args.workload == "all" and not args.headline_only:
        configs = {}
        for key, fn in (("pi_1e6", bench_pi_small), ("ragged_256", bench_ragged),
                        ("group_stage_1e7", bench_leaf), ("serving", bench_serve),
                        ("serving_burst", bench_serve_burst)):
            log(f"config {key}")
            t0 = time.perf_counter()
            configs[key] = fn(args, g)
            configs[key]["wall_s"] = round(time.perf_counter() - t0, 1)
        line["configs"] = configs
        line["batched_vs_plain"] = {
            "pi_1e9": line["plain_form"]["speedup_batched_vs_plain"] if line["plain_form"] else None,
            "pi_1e6": configs["pi_1e6"]["speedup_batched_vs_plain"],
            "ragged_256": configs["ragged_256"]["speedup_ragged_vs_plain"],
            "group_stage_1e7_fused": configs["group_stage_1e7"]["speedup_fused_vs_plain"],
            "group_stage_1e7_graph": configs["group_stage_1e7"]["speedup_graph_vs_plain"],
            "serving_burst_shared_trees": configs["serving_burst"]["speedup_shared_vs_single"],
        }
        ok = ok and configs["pi_1e6"]["plain_identical"] and configs["pi_1e6"]["bit_exact_vs_oracle"] in (None, True)
        ok = ok and configs["ragged_256"]["identical_to_plain"] and configs["group_stage_1e7"]["identical_rows"]
        ok = ok and configs["serving"]["identical_alone_and_in_mix"]
        ok = ok and configs["serving_burst"]["identical_results"]
        burst_cpu = configs["serving_burst"]["cpu_baseline"]
        ok = ok and (burst_cpu is None or burst_cpu["gpu_digits_checked_bit_exact"])
    print(json.dumps(line))
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
