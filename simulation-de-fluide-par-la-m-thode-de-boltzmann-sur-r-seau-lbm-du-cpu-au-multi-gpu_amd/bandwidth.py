#!/usr/bin/env python3
"""The reference's copy-ceiling tool and ceiling report on the MI355X (bench/gpu_bandwidth.cu,
bench/ceiling.py).

    bandwidth.py [output.csv] [--bytes N] [--reps R] [--peak-gbs G]

Measures the HBM copy ceiling with the library's copy kernel and hipMemcpy device-to-device
(best of R runs over two arrays of N bytes, gpu_bandwidth.cu:17-22,66-81), writes one CSV row
with gpu_bandwidth's columns (gpu_bandwidth.cu:92-96) to output.csv (appended to an existing
file without a second header) or to stdout, prints gpu_bandwidth's summary line, then
ceiling.py's MLUPS ceilings for float and double (2 * 9 * sizeof(real) bytes per update;
with --peak-gbs also the hardware ceiling, ceiling.py:42-59).
"""
from __future__ import annotations

import argparse
import csv
import io
import pathlib
import sys

HERE = pathlib.Path(__file__).resolve().parent
if __package__:
    from .driver import package
else:
    sys.path.insert(0, str(HERE))
    from driver import package  # noqa: E402


def gpu_name(info: str) -> str:
    """The device name of lbm_device_info's text, or "AMD <arch>" where the runtime reports
    none (as bench_info does)."""
    name, _, rest = info.partition(" (")
    arch = rest.split(")")[0].split(":")[0]
    return name or f"AMD {arch}"


def measure(nbytes: int, reps: int) -> dict:
    lbm = package()
    from lbm_amd import formats

    c = lbm.copy_ceiling(nbytes, reps)
    return formats.bandwidth_row(gpu_name(lbm.device_info()), c["copy_kernel"] * 1e3,
                                 c["memcpy"] * 1e3, nbytes, reps)


def report(rows, peak_gbs=None, out=None) -> None:
    """ceiling.py's printout (ceiling.py:51-59)."""
    from lbm_amd import formats

    out = out or sys.stdout
    best = max(rows, key=lambda r: max(float(r["copy_MBps"]), float(r["memcpy_MBps"] or 0)))
    rate = max(float(best["copy_MBps"]), float(best["memcpy_MBps"] or 0)) * 1e6
    print(f"Measured best copy: {rate / 1e9:.2f} GB/s ({best['gpu']})", file=out)
    for name, size in (("float", 4), ("double", 8)):
        measured, hw = formats.ceiling_mlups(rows, size, peak_gbs, gpu=True)
        line = f"  {name:6s}: measured ceiling = {measured:7.1f} MLUPS"
        if hw is not None:
            line += f"   hardware ceiling = {hw:7.1f} MLUPS"
        print(line, file=out)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("output", nargs="?", default=None)
    ap.add_argument("--bytes", type=int, default=1 << 30,
                    help="bytes per array (default 1 GiB: well above the 256 MB Infinity Cache; "
                         "the reference's T4 runs used 256 MiB)")
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--peak-gbs", type=float, default=None,
                    help="published HBM bandwidth in GB/s (MI355X: 8000)")
    args = ap.parse_args(argv)
    package()
    from lbm_amd import formats

    row = measure(args.bytes, args.reps)
    if args.output:
        formats.write_csv([row], args.output, formats.BANDWIDTH_COLUMNS, append=True)
    else:
        buf = io.StringIO()
        w = csv.DictWriter(buf, fieldnames=formats.BANDWIDTH_COLUMNS)
        w.writeheader()
        w.writerow(row)
        sys.stdout.write(buf.getvalue())
    print(f"{row['gpu']}: copy kernel {float(row['copy_MBps']) / 1e3:.1f} GB/s, hipMemcpy D2D "
          f"{float(row['memcpy_MBps']) / 1e3:.1f} GB/s")
    report([row], args.peak_gbs)
    return 0


if __name__ == "__main__":
    sys.exit(main())
