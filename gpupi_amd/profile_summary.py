"""Summaries of rocprofv3 output for the profiles kept with this library.

    python -m gpupi_amd.profile_summary counters OUT.md DIR [DIR ...]
        DIR: rocprofv3 --pmc ... --output-format csv -d DIR -o run runs
        (one per counter pass, same program).  Writes per-kernel sums of
        every counter and derived figures: share of wave time waiting,
        VALU issue rate per CU cycle, FP64 share of VALU instructions, LDS
        bank-conflict share, HBM bytes per launch from FETCH_SIZE (x2, the
        copy-kernel calibration) and WRITE_SIZE (x1).
    python -m gpupi_amd.profile_summary stats OUT.md DIR
        DIR: a rocprofv3 --kernel-trace --stats --output-format csv run;
        writes calls, total and average time per kernel.

Kernel names are shortened to the kernel template's name; the first
forward NTT level, which reads limbs (ntt_col_fwd_limbs_kernel), keeps the
label of the live kernel statistics, 'ntt_col_fwd_kernel<limbs>'."""
import collections
import csv
import glob
import os
import re
import sys

# FETCH_SIZE reads half the bytes a copy kernel of known size reads, and
# WRITE_SIZE the bytes it writes (1 GiB copy: 524,300 and 1,048,576 KiB).
FETCH_SCALE = 2.0
WRITE_SCALE = 1.0


def short(name: str) -> str:
    m = re.search(r"(\w+_kernel)\b", name)
    base = m.group(1) if m else name.split("(")[0][-48:]
    if base == "ntt_col_fwd_limbs_kernel" or (base == "ntt_col_fwd_kernel" and "<true" in name):
        base = "ntt_col_fwd_kernel<limbs>"
    return base


def load_counters(dirs):
    sums = collections.defaultdict(lambda: collections.defaultdict(float))
    launches = collections.defaultdict(set)
    dur = collections.defaultdict(float)
    for d in dirs:
        for path in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            seen = set()
            for r in csv.DictReader(open(path)):
                k = short(r["Kernel_Name"])
                sums[k][r["Counter_Name"]] += float(r["Counter_Value"])
                key = (path, r["Dispatch_Id"])
                if key not in seen:
                    seen.add(key)
                    launches[k].add(key)
                    dur[(k, path)] += (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-6
    return sums, launches, dur


def counters(out, dirs):
    sums, launches, dur = load_counters(dirs)
    runs = collections.defaultdict(list)
    for (k, path), ms in dur.items():
        runs[k].append(ms)
    rows = []
    for k, c in sums.items():
        n_pass = len(runs[k])
        ms = sum(runs[k]) / max(n_pass, 1)  # kernel time per counter pass
        n = len(launches[k]) / max(n_pass, 1)
        rows.append((ms, k, n, c))
    rows.sort(reverse=True)
    lines = ["| kernel | launches | ms | wait % of wave cycles | VALU instr / CU cycle | FP64 % of VALU | int64 % of VALU | "
             "LDS conflict % | HBM read GB/launch | HBM write GB/launch |", "|" + "---|" * 10]
    for ms, k, n, c in rows:
        def pct(a, b):
            return f"{100.0 * c[a] / c[b]:.0f}" if c.get(a) is not None and c.get(b) else ""
        fp64 = sum(c.get(x, 0.0) for x in ("SQ_INSTS_VALU_ADD_F64", "SQ_INSTS_VALU_MUL_F64", "SQ_INSTS_VALU_FMA_F64"))
        valu = c.get("SQ_INSTS_VALU")
        issue = f"{valu / c['SQ_BUSY_CU_CYCLES']:.2f}" if valu and c.get("SQ_BUSY_CU_CYCLES") else ""
        fp = f"{100.0 * fp64 / valu:.0f}" if valu and "SQ_INSTS_VALU_FMA_F64" in c else ""
        i64 = f"{100.0 * c['SQ_INSTS_VALU_INT64'] / valu:.0f}" if valu and "SQ_INSTS_VALU_INT64" in c else ""
        rd = f"{c['FETCH_SIZE'] * 1024 * FETCH_SCALE / n / 1e9:.3f}" if "FETCH_SIZE" in c and n else ""
        wr = f"{c['WRITE_SIZE'] * 1024 * WRITE_SCALE / n / 1e9:.3f}" if "WRITE_SIZE" in c and n else ""
        lines.append(f"| {k} | {n:.0f} | {ms:.2f} | {pct('SQ_WAIT_ANY', 'SQ_WAVE_CYCLES')} | {issue} | {fp} | {i64} | "
                     f"{pct('SQ_LDS_BANK_CONFLICT', 'SQ_LDS_IDX_ACTIVE')} | {rd} | {wr} |")
    lines += ["", "Raw sums per kernel (all launches of one pass):", ""]
    for ms, k, n, c in rows:
        lines.append(f"- {k}: " + ", ".join(f"{name}={v:.4g}" for name, v in sorted(c.items())))
    with open(out, "w") as f:
        f.write("\n".join(lines) + "\n")


def stats(out, d):
    path = glob.glob(os.path.join(d, "**", "*kernel_stats.csv"), recursive=True)[0]
    agg = collections.defaultdict(lambda: [0, 0.0])
    for r in csv.DictReader(open(path)):
        a = agg[short(r["Name"])]
        a[0] += int(r["Calls"])
        a[1] += float(r["TotalDurationNs"]) * 1e-6
    total = sum(v[1] for v in agg.values())
    lines = ["| kernel | calls | total ms | average ms | % |", "|---|---|---|---|---|"]
    for k, (n, ms) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        lines.append(f"| {k} | {n} | {ms:.2f} | {ms / n:.4f} | {100 * ms / total:.1f} |")
    lines.append(f"| all kernels | {sum(v[0] for v in agg.values())} | {total:.2f} | | 100 |")
    with open(out, "w") as f:
        f.write("\n".join(lines) + "\n")


def main() -> int:
    if len(sys.argv) < 4 or sys.argv[1] not in ("counters", "stats"):
        print(__doc__)
        return 2
    if sys.argv[1] == "counters":
        counters(sys.argv[2], sys.argv[3:])
    else:
        stats(sys.argv[2], sys.argv[3])
    return 0


if __name__ == "__main__":
    sys.exit(main())
