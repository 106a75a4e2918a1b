"""HBM traffic per launch from rocprofv3 counters: the MI355X counterpart of the reference's
bench/ncu_profile.py (which reads dram__bytes_read / dram__bytes_write of one launch with Nsight
Compute and reports DRAM bytes per cell update against the 72 B ideal, ncu_profile.py:27-86).

Two rocprofv3 passes (one per counter: FETCH_SIZE needs 3 of the 4 TCC counters of a pass,
WRITE_SIZE 2) over the same command, optionally a third for SQ_INSTS_VALU; each value is KiB
(or instructions) per dispatch. The raw counters are scaled by what the same counters read for
a copy kernel of known size run in the same process (the copy-ceiling kernel copy_kernel<4,
true> over two arrays of 2,415,919,104 bytes: the command must run the copy_ceiling form), so
the result is bytes that really crossed the HBM interface. The
summary goes to a Markdown file with a JSON block that bench.py reads:

    {"kernel": ..., "traffic_bytes_per_launch": ..., "read_bytes_per_launch": ...,
     "write_bytes_per_launch": ..., "bytes_per_cell_update": ..., "algorithmic_bytes": ...,
     "calibration": {"fetch": f, "write": w}, "valu_instructions_per_launch": ...,
     "source_tag": ...}

`source_tag` hashes the device sources the library is built from; bench.py flags the figures as
stale when the current sources hash differently.

Run on a machine with the GPU, from the repository root (the program itself right after `--`):

    python <package>/counters.py run --kernel step2_kernel<float, PlanesOne --cells 67108864 \\
        --updates-per-launch 2 --bytes-per-cell 72 --out profiles/step2_fp32_8192x8192.traffic.md \\
        -- python3 bench.py --steps 6 --warmup 2 --no-cpu --forms copy_ceiling

or, on CSV files that rocprofv3 wrote elsewhere, `summarize --fetch F.csv --write W.csv ...`.

`waves` reads the wave-state counters instead (two passes of eight SQ counters, WAVE_PASSES) and
reports per kernel the engine clock under the load (SQ_BUSY_CYCLES over the shader engines and
the dispatch time), the fraction of the SIMD time the vector ALUs were active
(SQ_ACTIVE_INST_VALU quad-cycles over SIMDs x time x clock / 4) and the resident waves per SIMD
(SQ_WAVE_CYCLES over the same quad-cycles):

    python <package>/counters.py waves --kernel step2_fp16_kernel --kernel 'step2_kernel<double' \
        --last 90 --out profiles/step2_8192x8192.waves.md \
        -- python3 bench.py --steps 20 --warmup 4 --settle 200 --no-cpu --forms precisions
"""
from __future__ import annotations

import argparse
import csv
import hashlib
import json
import os
import pathlib
import re
import shlex
import subprocess
import sys
import time

HERE = pathlib.Path(__file__).resolve().parent
CALIBRATION_KERNEL = "copy_kernel<4, true>"  # 16-byte non-temporal grid-stride copy
CALIBRATION_BYTES = 2_415_919_104            # bytes read (and written) per copy dispatch
# Wave-state passes (at most 8 SQ counters per rocprofv3 pass) and the MI355X shape they are
# normalised by: 256 CUs of 4 SIMDs, 32 shader engines (8 XCDs of 4), 2.4 GHz peak engine clock.
WAVE_PASSES = ("SQ_WAVES SQ_WAVE_CYCLES SQ_BUSY_CYCLES SQ_WAIT_ANY SQ_WAIT_INST_ANY "
               "SQ_ACTIVE_INST_VALU SQ_ACTIVE_INST_LDS SQ_INSTS_SALU",
               "SQ_INSTS_VALU SQ_INSTS_LDS SQ_INSTS_VMEM_RD SQ_INSTS_VMEM_WR SQ_ACTIVE_INST_ANY "
               "SQ_INST_CYCLES_VMEM_RD SQ_WAIT_INST_LDS SQ_ACTIVE_INST_MISC")
SIMDS = 1024
SHADER_ENGINES = 32
PEAK_CLOCK_HZ = 2.4e9


def source_tag() -> str:
    """Hash of the device sources (csrc/*.hip, *.hpp) the library is built from."""
    h = hashlib.sha256()
    for p in sorted((HERE / "csrc").glob("*")):
        if p.suffix in (".hip", ".hpp"):
            h.update(p.name.encode())
            h.update(p.read_bytes())
    return h.hexdigest()[:16]


def _col(header: list[str], *names: str) -> int:
    low = [h.strip().lower() for h in header]
    for n in names:
        if n.lower() in low:
            return low.index(n.lower())
    raise ValueError(f"none of the columns {names} in {header}")


def read_counters(path: str | os.PathLike) -> list[dict]:
    """Rows of a rocprofv3 counter-collection CSV as dicts: dispatch, kernel, counter, value
    (one row per dispatch and counter; values of one counter in one dispatch summed)."""
    with open(path, newline="") as f:
        rows = list(csv.reader(f))
    if not rows:
        return []
    head = rows[0]
    ki = _col(head, "Kernel_Name", "KernelName", "Kernel-Name")
    ci = _col(head, "Counter_Name", "CounterName")
    vi = _col(head, "Counter_Value", "CounterValue", "Value")
    di = _col(head, "Dispatch_Id", "DispatchId", "Correlation_Id", "Index")
    try:  # dispatch start and end (ns), where the file has them
        ti = (_col(head, "Start_Timestamp"), _col(head, "End_Timestamp"))
    except ValueError:
        ti = None
    acc: dict[tuple, dict] = {}
    for r in rows[1:]:
        if len(r) <= max(ki, ci, vi, di, *(ti or ())):
            continue
        key = (r[di], r[ci])
        e = acc.setdefault(key, {"dispatch": int(float(r[di])), "kernel": r[ki], "counter": r[ci],
                                 "value": 0.0})
        e["value"] += float(r[vi])
        if ti:
            e["ns"] = float(r[ti[1]]) - float(r[ti[0]])
    return sorted(acc.values(), key=lambda e: e["dispatch"])


def per_dispatch(rows: list[dict], kernel: str, counter: str, skip: int = 0,
                 last: int = 0) -> list[float]:
    """Values of `counter` for the dispatches whose kernel name contains `kernel`, after
    leaving out the first `skip` of them; with last > 0 only the last `last` of those."""
    vals = [e["value"] for e in rows if kernel in e["kernel"] and e["counter"] == counter][skip:]
    return vals[-last:] if last > 0 else vals


def _mean(v: list[float]) -> float:
    if not v:
        raise ValueError("no dispatch of the kernel in the counter file")
    return sum(v) / len(v)


def summarize(fetch_rows, write_rows, kernel: str, cells: int, updates: int,
              bytes_per_cell: float, valu_rows=None, skip: int = 0, last: int = 0,
              dispatches_per_launch: int = 1,
              calibration_kernel: str = CALIBRATION_KERNEL,
              calibration_bytes: int = CALIBRATION_BYTES) -> dict:
    """Calibrated HBM bytes per launch of `kernel`: the mean over its selected dispatches times
    dispatches_per_launch (a launch of the batched form is one dispatch per tile-width class;
    `last` then selects whole launches)."""
    if last and last % dispatches_per_launch:
        raise ValueError("last must be a whole number of launches")
    f_cal = calibration_bytes / 1024 / _mean(per_dispatch(fetch_rows, calibration_kernel, "FETCH_SIZE"))
    w_cal = calibration_bytes / 1024 / _mean(per_dispatch(write_rows, calibration_kernel, "WRITE_SIZE"))
    f = per_dispatch(fetch_rows, kernel, "FETCH_SIZE", skip, last)
    w = per_dispatch(write_rows, kernel, "WRITE_SIZE", skip, last)
    k = dispatches_per_launch
    read = _mean(f) * k * f_cal * 1024
    write = _mean(w) * k * w_cal * 1024
    algorithmic = bytes_per_cell * cells
    out = {"kernel": kernel, "dispatches": {"fetch": len(f), "write": len(w)},
           "fetch_kib_per_dispatch": round(_mean(f), 1), "write_kib_per_dispatch": round(_mean(w), 1),
           "calibration": {"fetch": round(f_cal, 4), "write": round(w_cal, 4),
                           "kernel": calibration_kernel, "bytes": calibration_bytes},
           "read_bytes_per_launch": round(read), "write_bytes_per_launch": round(write),
           "traffic_bytes_per_launch": round(read + write),
           "algorithmic_bytes": round(algorithmic), "cells": cells,
           "updates_per_launch": updates, "dispatches_per_launch": k,
           "traffic_vs_algorithmic": round((read + write) / algorithmic, 4),
           "read_vs_algorithmic": round(read / (algorithmic / 2), 4),
           "bytes_per_cell_update": round((read + write) / (cells * updates), 2),
           "source_tag": source_tag()}
    if valu_rows is not None:
        v = per_dispatch(valu_rows, kernel, "SQ_INSTS_VALU", skip, last)
        if v:
            out["valu_instructions_per_launch"] = round(_mean(v) * k)
            out["valu_instructions_per_update"] = round(_mean(v) * k * 64 / (cells * updates), 1)
    return out


def wave_states(rows: list[dict], kernel: str, skip: int = 0, last: int = 0,
                simds: int = SIMDS, shader_engines: int = SHADER_ENGINES) -> dict:
    """Wave-state figures of `kernel` from the rows of the SQ passes (the counters of both
    passes read into one list): per dispatch means of each counter over the selected dispatches,
    the mean dispatch time of each pass, the engine clock SQ_BUSY_CYCLES / shader engines / time,
    and, at that clock, the SIMD time in quad-cycles (simds x time x clock / 4) over which
    SQ_ACTIVE_INST_VALU gives the vector ALUs' busy fraction and SQ_WAVE_CYCLES the resident waves
    per SIMD."""
    names = sorted({e["counter"] for e in rows if kernel in e["kernel"]})
    if not names:
        raise ValueError("no dispatch of the kernel in the counter file")
    mean = {c: _mean(per_dispatch(rows, kernel, c, skip, last)) for c in names}

    def seconds(counter):
        sel = [e for e in rows if kernel in e["kernel"] and e["counter"] == counter][skip:]
        sel = sel[-last:] if last > 0 else sel
        ns = [e["ns"] for e in sel if "ns" in e]
        return _mean(ns) * 1e-9 if ns else None

    out = {"kernel": kernel, "counters": {c: mean[c] for c in names},
           "dispatches": len(per_dispatch(rows, kernel, names[0], skip, last))}
    t = seconds("SQ_BUSY_CYCLES")
    if t and "SQ_BUSY_CYCLES" in mean:
        clock = mean["SQ_BUSY_CYCLES"] / shader_engines / t
        quads = simds * t * clock / 4
        out.update(dispatch_seconds=t, clock_hz=clock, simd_quad_cycles=quads)
        if "SQ_ACTIVE_INST_VALU" in mean:
            out["valu_active_fraction"] = mean["SQ_ACTIVE_INST_VALU"] / quads
            out["valu_active_fraction_at_peak_clock"] = (
                mean["SQ_ACTIVE_INST_VALU"] / (simds * t * PEAK_CLOCK_HZ / 4))
        if "SQ_WAVE_CYCLES" in mean:
            out["waves_per_simd"] = mean["SQ_WAVE_CYCLES"] / quads
        if "SQ_INSTS_VALU" in mean:
            out["valu_instruction_fraction"] = mean["SQ_INSTS_VALU"] / quads
    return out


def write_waves(path: str | os.PathLike, results: list[dict], commands: list[str]) -> None:
    md = ["# Wave states (rocprofv3 SQ counters)", "",
          f"Written by `counters.py` on {time.strftime('%Y-%m-%d %H:%M')} UTC; device sources "
          f"tag `{source_tag()}`. Per dispatch, means over {results[0]['dispatches']} dispatches "
          f"of each kernel; {SIMDS} SIMDs, {SHADER_ENGINES} shader engines. Commands (repository "
          "root):", ""]
    md += [f"    {c}" for c in commands]
    md += ["", "| quantity | " + " | ".join(f"`{r['kernel']}`" for r in results) + " |",
           "|---|" + "---|" * len(results)]

    def row(label, f):
        md.append(f"| {label} | " + " | ".join(f(r) for r in results) + " |")

    def opt(key, fmt):
        return lambda r: fmt.format(r[key]) if key in r else "-"

    row("dispatch time (ms)", lambda r: f"{r['dispatch_seconds'] * 1e3:.4f}" if "dispatch_seconds" in r else "-")
    row("engine clock (GHz, SQ_BUSY_CYCLES)", lambda r: f"{r['clock_hz'] / 1e9:.3f}" if "clock_hz" in r else "-")
    row("vector ALUs active (at that clock; at 2.4 GHz)",
        lambda r: (f"{r['valu_active_fraction']:.3f}; {r['valu_active_fraction_at_peak_clock']:.3f}"
                   if "valu_active_fraction" in r else "-"))
    row("VALU instructions / SIMD quad-cycles", opt("valu_instruction_fraction", "{:.3f}"))
    row("resident waves per SIMD", opt("waves_per_simd", "{:.2f}"))
    for c in sorted({c for r in results for c in r["counters"]}):
        row(c, lambda r, c=c: f"{r['counters'][c]:.4e}" if c in r["counters"] else "-")
    md += ["", "```json", json.dumps({"source_tag": source_tag(), "kernels": results}, indent=1),
           "```", ""]
    pathlib.Path(path).write_text("\n".join(md))


def write_summary(path: str | os.PathLike, s: dict, commands: list[str]) -> None:
    md = [f"# HBM traffic of `{s['kernel']}` (rocprofv3 FETCH_SIZE / WRITE_SIZE, calibrated)", "",
          f"Written by `counters.py` on {time.strftime('%Y-%m-%d %H:%M')} UTC; device sources "
          f"tag `{s['source_tag']}`. Commands (repository root):", ""]
    md += [f"    {c}" for c in commands]
    md += ["", "| quantity | value |", "|---|---|",
           f"| dispatches of the kernel (fetch / write pass) | {s['dispatches']['fetch']} / "
           f"{s['dispatches']['write']} |",
           f"| dispatches per launch | {s.get('dispatches_per_launch', 1)} |",
           f"| FETCH_SIZE KiB per dispatch | {s['fetch_kib_per_dispatch']:,} |",
           f"| WRITE_SIZE KiB per dispatch | {s['write_kib_per_dispatch']:,} |",
           f"| calibration ({s['calibration']['kernel']}, {s['calibration']['bytes']:,} B each "
           f"way) | FETCH x {s['calibration']['fetch']}, WRITE x {s['calibration']['write']} |",
           f"| read / written bytes per launch | {s['read_bytes_per_launch']:,} / "
           f"{s['write_bytes_per_launch']:,} |",
           f"| algorithmic bytes per launch | {s['algorithmic_bytes']:,} |",
           f"| traffic / algorithmic (reads alone) | {s['traffic_vs_algorithmic']} "
           f"({s['read_vs_algorithmic']}) |",
           f"| HBM bytes per cell update | {s['bytes_per_cell_update']} |"]
    if "valu_instructions_per_launch" in s:
        md.append(f"| VALU wave-instructions per launch (per update) | "
                  f"{s['valu_instructions_per_launch']:,} ({s['valu_instructions_per_update']}) |")
    md += ["", "```json", json.dumps(s, indent=1), "```", ""]
    pathlib.Path(path).write_text("\n".join(md))


def read_summary(path: str | os.PathLike) -> dict | None:
    """The JSON block of a summary written by write_summary (None if absent)."""
    p = pathlib.Path(path)
    if not p.exists():
        return None
    m = re.search(r"```json\n(.*?)\n```", p.read_text(), re.S)
    return json.loads(m.group(1)) if m else None


def find_csv(directory: str | os.PathLike) -> pathlib.Path:
    hits = sorted(pathlib.Path(directory).rglob("*counter_collection*.csv"))
    if not hits:
        raise FileNotFoundError(f"no counter_collection CSV under {directory}")
    return hits[-1]


def run_pass(counters: str, out_dir: str, cmd: list[str], timeout: int) -> tuple[pathlib.Path, str]:
    """One rocprofv3 pass; returns the counter CSV and the command line it ran."""
    full = ["rocprofv3", "--pmc", *counters.split(), "--output-format", "csv", "-d", out_dir,
            "-o", "pmc", "--", *cmd]
    line = " ".join(shlex.quote(c) for c in full)
    print(f"[counters] {line}", flush=True)
    subprocess.run(["timeout", "-s", "KILL", str(timeout), *full], check=True)
    return find_csv(out_dir), line


def main(argv=None) -> None:
    ap = argparse.ArgumentParser(description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="what", required=True)
    for name in ("run", "summarize"):
        sp = sub.add_parser(name)
        sp.add_argument("--kernel", required=True, help="substring of the kernel name")
        sp.add_argument("--cells", type=int, required=True, help="cells per launch")
        sp.add_argument("--updates-per-launch", type=int, default=1)
        sp.add_argument("--bytes-per-cell", type=float, required=True,
                        help="algorithmic bytes per cell per launch")
        sp.add_argument("--skip", type=int, default=0, help="dispatches left out at the start")
        sp.add_argument("--last", type=int, default=0,
                        help="only the last N dispatches (after --skip)")
        sp.add_argument("--dispatches-per-launch", type=int, default=1,
                        help="dispatches of the kernel that make one launch of the path")
        sp.add_argument("--out", required=True, help="summary Markdown file")
    r = sub.choices["run"]
    r.add_argument("--work", default="rocprof_counters", help="rocprofv3 output directory")
    r.add_argument("--valu", action="store_true", help="third pass with SQ_INSTS_VALU")
    r.add_argument("--timeout", type=int, default=300)
    r.add_argument("cmd", nargs=argparse.REMAINDER, help="-- program and arguments")
    s = sub.choices["summarize"]
    s.add_argument("--fetch", required=True)
    s.add_argument("--write", required=True)
    s.add_argument("--valu-csv", default=None)
    wv = sub.add_parser("waves", help="wave-state counters (two SQ passes) per kernel")
    wv.add_argument("--kernel", action="append", required=True,
                    help="substring of a kernel name (repeat for several kernels)")
    wv.add_argument("--skip", type=int, default=0, help="dispatches left out at the start")
    wv.add_argument("--last", type=int, default=0, help="only the last N dispatches (after --skip)")
    wv.add_argument("--out", required=True, help="summary Markdown file")
    wv.add_argument("--work", default="rocprof_counters", help="rocprofv3 output directory")
    wv.add_argument("--timeout", type=int, default=300)
    wv.add_argument("--csv", action="append", default=None,
                    help="summarise these counter CSV files instead of running the passes")
    wv.add_argument("cmd", nargs=argparse.REMAINDER, help="-- program and arguments")
    a = ap.parse_args(argv)
    rel = pathlib.Path(__file__).resolve().relative_to(pathlib.Path.cwd()) \
        if pathlib.Path.cwd() in pathlib.Path(__file__).resolve().parents else pathlib.Path(__file__)
    made_by = " ".join(shlex.quote(c) for c in ["python3", str(rel), *(argv or sys.argv[1:])])
    if a.what == "waves":
        rows, commands = [], [made_by]
        if a.csv:
            for c in a.csv:
                rows += read_counters(c)
        else:
            cmd = a.cmd[1:] if a.cmd and a.cmd[0] == "--" else a.cmd
            if not cmd:
                ap.error("give the program after --, or --csv files")
            commands.append("  which runs:")
            for k, counters in enumerate(WAVE_PASSES):
                csv_path, line = run_pass(counters, f"{a.work}/sq{k}", cmd, a.timeout)
                commands.append(line)
                rows += read_counters(csv_path)
        results = [wave_states(rows, k, a.skip, a.last) for k in a.kernel]
        write_waves(a.out, results, commands)
        print(json.dumps(results), flush=True)
        return
    if a.what == "run":
        cmd = a.cmd[1:] if a.cmd and a.cmd[0] == "--" else a.cmd
        if not cmd:
            ap.error("give the program after --")
        fcsv, fl = run_pass("FETCH_SIZE", f"{a.work}/fetch", cmd, a.timeout)
        wcsv, wl = run_pass("WRITE_SIZE", f"{a.work}/write", cmd, a.timeout)
        commands = [made_by, "  which runs:", fl, wl]
        vrows = None
        if a.valu:
            vcsv, vl = run_pass("SQ_WAVES SQ_INSTS_VALU", f"{a.work}/valu", cmd, a.timeout)
            commands.append(vl)
            vrows = read_counters(vcsv)
        summary = summarize(read_counters(fcsv), read_counters(wcsv), a.kernel, a.cells,
                            a.updates_per_launch, a.bytes_per_cell, vrows, a.skip, a.last,
                            a.dispatches_per_launch)
    else:
        commands = [made_by]
        summary = summarize(read_counters(a.fetch), read_counters(a.write), a.kernel, a.cells,
                            a.updates_per_launch, a.bytes_per_cell,
                            read_counters(a.valu_csv) if a.valu_csv else None, a.skip, a.last,
                            a.dispatches_per_launch)
    write_summary(a.out, summary, commands)
    print(json.dumps(summary), flush=True)


if __name__ == "__main__":
    main()
