"""Write tests/golden/bench_protocol.json: the reference's recorded benchmark data that pins the
benchmark-protocol and ceiling code (data only, read from the reference checkout's results/
and README; no reference code is run).

  * run_bench CSV header and rows (nx, ny, steps, warmup, reps, time_mean_s, mlups_mean,
    mlups_max) of results/v1_i5-10310U.csv and results/v3_blocks_Tesla-T4.csv;
  * the copy-ceiling CSV header and rows of results/gpu_bandwidth_Tesla-T4.csv and the STREAM
    rows of results/bandwidth_i5-10310U.csv;
  * the ceilings README.md:41-50 quotes for them (3,221 MLUPS GPU, 199 MLUPS CPU, float);
  * the results/strouhal.csv header and rows (the reference's cylinder runs: 1000 x 320,
    D 40, Re 100, u_in 0.05, 160,000 steps, St from the probe over the last 45 %).

Usage (development container, reference checkout at /root/reference):
    python tests/golden/make_protocol_fixture.py
"""
from __future__ import annotations

import csv
import json
import pathlib

REF = pathlib.Path("/root/reference")
HERE = pathlib.Path(__file__).resolve().parent


def rows(name, keep=None, limit=None):
    with open(REF / "results" / name, newline="") as f:
        r = csv.DictReader(f)
        out = [{k: v for k, v in row.items() if keep is None or k in keep} for row in r]
        return r.fieldnames, out[:limit] if limit else out


def main():
    keep = ["version", "nx", "ny", "steps", "warmup", "reps", "threads", "time_mean_s",
            "mlups_mean", "mlups_max", "block"]
    h_cpu, r_cpu = rows("v1_i5-10310U.csv", keep)
    h_gpu, r_gpu = rows("v3_blocks_Tesla-T4.csv", keep)
    h_bw, r_bw = rows("gpu_bandwidth_Tesla-T4.csv")
    _, r_stream = rows("bandwidth_i5-10310U.csv", ["threads", "copy_MBps"])
    h_st, r_st = rows("strouhal.csv")
    fixture = {
        "source": "reference results/*.csv and README.md:41-50",
        "run_bench_header_cpu": h_cpu, "run_bench_rows_cpu": r_cpu,
        "run_bench_header_gpu": h_gpu, "run_bench_rows_gpu": r_gpu,
        "gpu_bandwidth_header": h_bw, "gpu_bandwidth_rows": r_bw,
        "stream_rows": r_stream,
        "readme_ceiling_mlups": {"gpu_float": 3221, "cpu_float": 199},
        "strouhal_header": h_st, "strouhal_rows": r_st,
    }
    (HERE / "bench_protocol.json").write_text(json.dumps(fixture, indent=1) + "\n")
    print(f"{len(r_cpu)} + {len(r_gpu)} run_bench rows, {len(r_bw)} bandwidth rows")


if __name__ == "__main__":
    main()
