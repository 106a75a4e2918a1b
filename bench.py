"""Benchmark of the fused D2Q9 step on one MI355X, printed as one JSON line.

Workload (the instance the headline metric is quoted on): plain fp32 step on an 8192 x 8192
force-driven channel with walls at y = 0 and y = ny - 1, periodic in x, tau = 0.6, gx = 1e-6,
rest initial state; `--steps` timed steps (default 200) after `--warmup` untimed ones.
MLUPS = nx * ny * steps / t / 1e6, counting every cell (reference src/common/driver.hpp:120).
Timing: HIP events recorded on the solver's stream around the timed steps (lbm_timed_run).

Besides the headline value the line carries:
  roofline      72 algorithmic bytes per cell update (9 reads + 9 writes of fp32) x cells per
                launch / average launch time, against the 8 TB/s HBM peak; `traffic` is the
                rocprofv3-measured HBM bytes per launch (FETCH_SIZE + WRITE_SIZE, scaled by the
                copy-kernel calibration), read from the committed profile summary;
  cpu_baseline  the reference's own v1 OpenMP solver compiled from its sources (oracle/_ref),
                or the oracle's OpenMP port when that is absent, timed on this host on a
                bounded sample of the same channel;
  forms         the other measured forms (copy ceiling, fp64, FP16 storage, batched domains,
                adjoint cost) when --forms is given.
"""
from __future__ import annotations

import argparse
import json
import os
import pathlib
import re
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

import __graft_entry__  # noqa: E402

HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL = {"float": 72, "double": 144, "fp16": 36}
PROFILE_SUMMARY = ROOT / "profiles" / "step_fp32_8192x8192.md"


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def traffic_from_profile():
    """HBM bytes per launch of the plain fp32 step, from the committed rocprofv3 summary."""
    if not PROFILE_SUMMARY.exists():
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", PROFILE_SUMMARY.read_text())
    return float(m.group(1)) if m else None


def run_plain(lbm, nx, ny, steps, warmup, precision="float", block=None):
    s = lbm.Solver(nx, ny, 0.6, 1e-6, 0.0, lbm.make_channel(nx, ny), precision)
    if block:
        s.set_block(*block)
    s.init_equilibrium(1.0, 0.0, 0.0)
    s.run(warmup)
    t = s.timed_run(steps)
    info = s.bench_info()
    mass = s.total_mass()
    s.close()
    return t, info, mass


def cpu_baseline(nx, ny, steps=10, warmup=2):
    """Reference v1 (oracle/_ref) or the oracle's OpenMP port, on this host's cores."""
    import oracle

    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    sample = f"{nx}x{ny} walled channel, tau 0.6, gx 1e-6, {warmup} untimed + {steps} timed steps"
    if oracle.Reference.available("float"):
        ref = oracle.Reference("float")
        sec, used = ref.v1_channel_timed(nx, ny, 0.6, 1e-6, warmup, steps)
        kind = "reference"
    else:
        o = oracle.Oracle(nx, ny, 0.6, 1e-6, 0.0, oracle.make_channel(nx, ny), "float", "pull")
        o.init_equilibrium(1.0, 0.0, 0.0)
        o.run(warmup)
        sec = o.timed_run(steps)
        used = o.threads()
        kind = "port"
    return {"value": nx * ny * steps / sec / 1e6, "unit": "MLUPS", "cores": used or threads,
            "kind": kind, "sample": sample, "seconds": sec}


def forms(lbm, args):
    out = {}
    t0 = time.time()
    copy_gbs, memcpy_gbs = lbm.copy_ceiling(1 << 31, 20)
    ceiling = max(copy_gbs, memcpy_gbs)
    out["copy_ceiling"] = {"copy_kernel_gbs": copy_gbs, "memcpy_dtod_gbs": memcpy_gbs,
                           "bytes_per_array": 1 << 31, "mlups_ceiling_fp32": ceiling * 1e3 / 72}
    log(f"copy ceiling {copy_gbs:.0f} / {memcpy_gbs:.0f} GB/s ({time.time() - t0:.1f} s)")
    for prec, steps in (("double", 100), ("fp16", 200)):
        t, info, _ = run_plain(lbm, args.nx, args.ny, steps, 20, prec)
        mlups = args.nx * args.ny * steps / t / 1e6
        out[f"plain_{prec}"] = {"mlups": mlups, "steps": steps,
                                "hbm_gbs": mlups * 1e6 * BYTES_PER_CELL[prec] / 1e9}
        log(f"{prec}: {mlups:.0f} MLUPS")
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=40)
    ap.add_argument("--nx", type=int, default=8192)
    ap.add_argument("--ny", type=int, default=8192)
    ap.add_argument("--block", type=str, default=None, help="BXxROWS launch shape")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--forms", action="store_true", help="also time the other forms")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: every box here has a single MI355X and nothing in "
              f"this build shards the domain across GPUs (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)

    lbm = __graft_entry__.load_package()
    log(f"device: {lbm.device_info()}")
    block = tuple(int(v) for v in args.block.split("x")) if args.block else None
    t, info, mass = run_plain(lbm, args.nx, args.ny, args.steps, args.warmup, "float", block)
    cells = args.nx * args.ny
    mlups = cells * args.steps / t / 1e6
    launch_s = t / args.steps
    achieved = BYTES_PER_CELL["float"] * cells / launch_s / 1e9
    log(f"plain fp32 {args.nx}x{args.ny}: {mlups:.1f} MLUPS, {launch_s * 1e3:.4f} ms/step, "
        f"{achieved:.0f} GB/s algorithmic; {info}")

    line = {
        "metric": "MLUPS & % HBM roofline, 8192² fp32 plain vs fp16-store/batched; adjoint cost",
        "value": round(mlups, 2),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(launch_s * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"plain fp32 D2Q9 step (SoA pull + BGK + Guo + halfway "
                               f"bounce-back, f-w storage), {args.nx}x{args.ny} walled channel, "
                               f"tau 0.6, gx 1e-6, rest init",
                   "nx": args.nx, "ny": args.ny, "parallelism": "single GPU",
                   "solver": json.loads("{" + info + "}")},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic_from_profile()},
    }
    if args.forms:
        line["forms"] = forms(lbm, args)
        ceil = line["forms"]["copy_ceiling"]["mlups_ceiling_fp32"]
        line["forms"]["plain_float_frac_of_copy_ceiling"] = round(mlups / ceil, 4)
    if not args.no_cpu:
        log("cpu baseline ...")
        line["cpu_baseline"] = cpu_baseline(args.nx, args.ny)
        log(f"cpu baseline: {line['cpu_baseline']['value']:.1f} MLUPS on "
            f"{line['cpu_baseline']['cores']} threads ({line['cpu_baseline']['kind']})")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
