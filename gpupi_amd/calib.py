"""Calibration runs for the rooflines (no counterpart in the reference).

    python -m gpupi_amd.calib mac            measured v_mad_u64_u32 peak (MAC/s)
    python -m gpupi_amd.calib copy BYTES     one plain device copy of BYTES

The copy is the reference kernel for the HBM traffic counters: run it under
`rocprofv3 --pmc FETCH_SIZE` (then WRITE_SIZE) and divide BYTES by the
counter to get the factor that turns an NTT kernel's counter into bytes."""
import argparse
import json
import sys

import numpy as np

import gpupi_amd as g


def main() -> int:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    sub.add_parser("mac")
    cp = sub.add_parser("copy")
    cp.add_argument("bytes", type=int)
    args = ap.parse_args()
    if args.cmd == "mac":
        dev = g.device_info()
        peak = g.mac_peak()
        per_clock = peak / dev["compute_units"] / (dev["clock_khz"] * 1e3)
        print(json.dumps({"macs_per_second": peak, "per_cu_per_clock": round(per_clock, 2), **dev}))
        return 0
    n = args.bytes
    src, dst = g.DeviceBuffer(n), g.DeviceBuffer(n)
    src.fill(0x5A)
    g.check(g.lib().gpupi_calib_copy(dst.ptr, src.ptr, n, None), "gpupi_calib_copy")
    g.synchronize()
    ok = bool((dst.download(np.uint8) == 0x5A).all())
    print(json.dumps({"copied_bytes": n, "ok": ok}))
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
