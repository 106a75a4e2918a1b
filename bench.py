"""Benchmark of the fused D2Q9 step on one MI355X, printed as one JSON line.

Workload (the instance the headline metric is quoted on): plain fp32 step on an 8192 x 8192
force-driven channel with walls at y = 0 and y = ny - 1, periodic in x, tau = 0.6, gx = 1e-6,
rest initial state; `--steps` timed steps (default 200) after `--warmup` untimed ones, which
follow `--settle` untimed steps (default 2000, about 0.9 s: the engine clock ramps for this
kernel after the idle state upload; reported as `settle_steps`).
MLUPS = nx * ny * steps / t / 1e6, counting every cell (reference src/common/driver.hpp:120).
Timing: HIP events recorded on the solver's stream around the timed steps (lbm_timed_run).
The solver's run() advances pairs of steps in one launch of the two-step kernel (bit-identical
to two single steps); an odd --steps ends with one single-step launch.

Besides the headline value the line carries:
  roofline      of the two-step kernel: 72 algorithmic bytes per cell per launch (each of the 9
                fp32 populations read once and written once for two updates) x cells / average
                launch time, against the 8 TB/s HBM peak; `traffic` is the rocprofv3-measured
                HBM bytes per launch (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel
                calibration), read from the committed profile summary; `valu_issue_frac` the
                profiled VALU instructions of a launch (SQ_INSTS_VALU, same summary) against
                the vector-ALU issue ceiling (256 CUs x 4 SIMDs x 2.4 GHz / 4 cycles per wave64
                instruction) over the measured launch time: the kernel's other bound;
  cpu_baseline  the reference's own v1 OpenMP solver compiled from its sources (oracle/_ref),
                or the oracle's OpenMP port when that is absent, timed on this host on the
                reference's protocol (warm-up steps / 5, 5 runs, mean / std / max) at one
                thread and at all physical cores of the process's CPU share, with the CPU
                model from lscpu; `value` is the all-core mean on the 8192^2 channel;
  forms         the other measured forms (one step per launch, the reference's 5-repetition
                protocol on the 8192^2 and 2048^2 instances, copy ceiling, fp64, FP16 storage,
                batched domains, strip share, adjoint cost, cylinder wake).
"""
from __future__ import annotations

import argparse
import json
import os
import pathlib
import subprocess
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

import __graft_entry__  # noqa: E402

HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL = {"float": 72, "double": 144, "fp16": 36}
TRAFFIC_SUMMARY = ROOT / "profiles" / "step2_fp32_8192x8192.traffic.md"
BATCH_TRAFFIC_SUMMARY = ROOT / "profiles" / "batch_step2_256_domains.traffic.md"


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def traffic_summary(lbm, path=None):
    """A kernel's calibrated HBM traffic and VALU count per launch (default: the two-step fp32
    kernel), from the summary the package's counter tool (counters.py: rocprofv3 FETCH_SIZE /
    WRITE_SIZE / SQ_INSTS_VALU passes) wrote; `stale` when the device sources changed since."""
    import importlib

    path = TRAFFIC_SUMMARY if path is None else path
    counters = importlib.import_module(lbm.__name__ + ".counters")
    s = counters.read_summary(path)
    if not s:
        return None
    s["stale"] = s.get("source_tag") != counters.source_tag()
    s["file"] = str(path.relative_to(ROOT))
    return s


# Vector-ALU issue ceiling of MI355X: 256 CUs x 4 SIMDs, one wave64 instruction per SIMD every
# 4 cycles at the 2.4 GHz peak engine clock.
VALU_WAVE_INSTR_PER_S = 256 * 4 * 2.4e9 / 4


def run_plain(lbm, nx, ny, steps, warmup, precision="float", block=None, steps_per_launch=2,
              settle=0):
    """`settle` untimed steps, then `warmup` untimed steps, then `steps` timed ones. The settle
    steps exist because the engine clock ramps for this kernel: after the state upload (seconds
    of host work, the GPU idle) the first 20 timed steps ran 127-138k MLUPS against 145-148k
    settled, also when a second of HBM-bound copies ran just before; after 2000 steps of the
    kernel itself the first 20 ran at the settled rate."""
    s = lbm.Solver(nx, ny, 0.6, 1e-6, 0.0, lbm.make_channel(nx, ny), precision)
    if block:
        s.set_block(*block)
    s.set_steps_per_launch(steps_per_launch)
    s.init_equilibrium(1.0, 0.0, 0.0)
    if settle:
        s.run(settle)
    s.run(warmup)
    t = s.timed_run(steps)
    info = s.bench_info()
    mass = s.total_mass()
    s.close()
    return t, info, mass


def host_cpu():
    """The host's CPU as run_bench.py records it (bench/run_bench.py:135-160): model, sockets,
    cores and threads per core from lscpu, the CPUs this process may run on, the frequency
    governor, and the physical cores of that share."""
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
    except (OSError, subprocess.SubprocessError):
        out = ""
    keys = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
            "Thread(s) per core": "threads_per_core", "NUMA node(s)": "numa_nodes", "CPU(s)": "cpus"}
    for line in out.splitlines():
        k, _, v = line.partition(":")
        if k.strip() in keys:
            v = v.strip()
            info[keys[k.strip()]] = int(v) if v.isdigit() else v
    gov = pathlib.Path("/sys/devices/system/cpu/cpu0/cpufreq/scaling_governor")
    info["governor"] = gov.read_text().strip() if gov.exists() else "unknown"
    info["process_cpus"] = len(os.sched_getaffinity(0))
    info["process_physical_cores"] = max(1, info["process_cpus"] // int(info.get("threads_per_core") or 1))
    return info


# CPU baseline instances: the 2048^2 throughput instance of the
# plain method (walls + seeded 8x8 block, tau 0.8, F = (1e-5, 2e-6)) and the 8192^2 channel
# (tau 0.6, gx 1e-6), rest init, float. Timed steps per repetition: the protocol's own for all
# cores; bounded samples for one thread (the full protocol there would take minutes).
CPU_INSTANCES = (
    ("block_2048", 2048, ["--mask", "block", "--tau", "0.8", "--fx", "1e-5", "--fy", "2e-6"],
     {"all": 1000, "one": 100}),
    ("channel_8192", 8192, ["--mask", "channel", "--tau", "0.6", "--fx", "1e-6"],
     {"all": 200, "one": 10}),
)


def cpu_baseline(full: bool = False):
    """The reference's v1 (oracle/_ref) or the oracle's OpenMP port on this host, on the
    reference's protocol: warm-up of steps / 5, then 5 timed runs in one process, mean / std /
    max (src/common/driver.hpp:115-132, bench/run_bench.py:124-160), with OMP_PROC_BIND=close
    and OMP_PLACES=cores, at 1 thread and at all physical cores of this process's CPU share
    (the OMP_NUM_THREADS the host sets, when it sets one). Each run is a child process
    (oracle/cpu_protocol.py) so the placement is in force when OpenMP starts. `value` is the
    all-core rate on the headline's 8192^2 channel."""
    cpu = host_cpu()
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    cores = min(cpu["process_physical_cores"], share) if share else cpu["process_physical_cores"]
    runs = {}
    for name, n, extra, steps in CPU_INSTANCES:
        for label, threads in (("one_thread", 1), ("all_cores", cores)):
            k = steps["all"] if (label == "all_cores" or full) else steps["one"]
            env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close",
                       OMP_PLACES="cores")
            cmd = [sys.executable, "-m", "oracle.cpu_protocol", "--nx", str(n), "--ny", str(n),
                   "--steps", str(k), "--reps", "5", *extra]
            t0 = time.time()
            r = subprocess.run(cmd, cwd=str(ROOT), env=env, capture_output=True, text=True,
                               timeout=1800)
            if r.returncode != 0:
                raise RuntimeError(f"cpu protocol {name} {label}: {r.stderr[-500:]}")
            res = json.loads(r.stdout.strip().splitlines()[-1])
            runs[f"{name}_{label}"] = res
            log(f"cpu {name} {label} ({res['threads']} threads, {k} steps x 5): "
                f"{res['mean']:.1f} +- {res['std']:.1f} MLUPS ({time.time() - t0:.1f} s)")
    head = runs["channel_8192_all_cores"]
    return {"value": head["mean"], "unit": "MLUPS", "cores": head["threads"], "kind": head["kind"],
            "sample": f"8192x8192 walled channel, tau 0.6, gx 1e-6, float: warm-up "
                      f"{head['warmup']} steps, then 5 runs of {head['steps']} steps (mean; "
                      f"reference protocol), {head['threads']} OpenMP threads, "
                      f"OMP_PROC_BIND=close, OMP_PLACES=cores",
            "std": head["std"], "max": head["max"], "cpu": cpu, "runs": runs,
            "note": "one-thread runs time 100 (2048^2) and 10 (8192^2) steps per repetition "
                    "(bounded sample) unless --cpu-full"}


def form_single_step(lbm, args):
    """The same workload with one step per launch (step_kernel): 72 B (fp32) or 144 B (fp64)
    per cell update; the fp32 figures at the top level, fp64 and FP16 storage under "double" and
    "fp16"."""
    out = {}
    for prec in ("float", "double", "fp16"):
        t, info, _ = run_plain(lbm, args.nx, args.ny, args.steps, args.warmup, prec, None, 1,
                               settle=args.settle // 2)
        mlups = args.nx * args.ny * args.steps / t / 1e6
        gbs = mlups * 1e6 * BYTES_PER_CELL[prec] / 1e9
        log(f"{prec} single step per launch: {mlups:.0f} MLUPS, {gbs:.0f} GB/s")
        r = {"mlups": round(mlups, 1), "timed_steps": args.steps,
             "hbm_gbs_algorithmic": round(gbs, 1), "frac_of_hbm_peak": round(gbs / HBM_PEAK_GBS, 4),
             "solver": json.loads("{" + info + "}")}
        if prec == "float":
            out.update(r)
        else:
            out[prec] = r
    return out


def reference_protocol(s, cells, steps, reps=5):
    """The reference's timing protocol (src/common/driver.hpp:115-132, bench/run_bench.py:138-139):
    untimed warm-up of steps / 5, then `reps` timed runs of `steps` steps in one process."""
    import numpy as np

    s.run(max(1, steps // 5))
    m = np.array([cells * steps / s.timed_run(steps) / 1e6 for _ in range(reps)])
    return {"mlups_mean": round(float(m.mean()), 1), "mlups_std": round(float(m.std()), 1),
            "mlups_max": round(float(m.max()), 1), "reps": reps, "timed_steps": steps,
            "warmup_steps": max(1, steps // 5)}


def form_protocol(lbm, args):
    """The headline instance and the config-0 throughput instance on the reference's protocol:
    8192^2 channel (tau 0.6, gx 1e-6, 200 steps) and 2048^2 walls + seeded 8x8 block (tau 0.8,
    F = (1e-5, 2e-6), rest init, 1000 steps), fp32, default path."""
    out = {}
    s = lbm.Solver(args.nx, args.ny, 0.6, 1e-6, 0.0, lbm.make_channel(args.nx, args.ny), "float")
    s.init_equilibrium(1.0, 0.0, 0.0)
    out["channel_8192"] = reference_protocol(s, args.nx * args.ny, 200)
    s.close()
    n = 2048
    s = lbm.Solver(n, n, 0.8, 1e-5, 2e-6, lbm.seeded_block_mask(0, n, n), "float")
    s.init_equilibrium(1.0, 0.0, 0.0)
    out["block_2048"] = reference_protocol(s, n * n, 1000)
    out["block_2048"]["block_seed"] = 0
    out["block_2048"]["block_xy"] = list(lbm.seeded_block_position(0, n, n))
    out["block_2048"]["note"] = "fp32 pair 302 MB: partly held by the 256 MB Infinity Cache"
    s.close()
    log(f"protocol: 8192^2 {out['channel_8192']['mlups_mean']:.0f} MLUPS, 2048^2 block "
        f"{out['block_2048']['mlups_mean']:.0f} MLUPS (mean of 5)")
    return out


def form_copy_ceiling(lbm, args):
    nbytes = 2_415_919_104  # one fp32 population buffer of the 8192^2 grid (>> 256 MB MALL)
    c = lbm.copy_ceiling(nbytes, 20)
    ceiling = max(c["copy_kernel"], c["memcpy"])
    return {"copy_kernel_gbs": round(c["copy_kernel"], 1), "memcpy_dtod_gbs": round(c["memcpy"], 1),
            "reference_method_copy_gbs": round(c["reference_kernel"], 1),
            "nine_array_copy_gbs": round(c["nine_array_copy"], 1),
            "bytes_per_array": nbytes, "reps": 20,
            "mlups_ceiling_fp32": round(ceiling * 1e3 / 72, 1),
            "mlups_ceiling_fp16": round(ceiling * 1e3 / 36, 1),
            "note": "ceiling = max(best copy kernel, memcpy); the reference's own copy "
                    "configuration and a nine-array copy (the planes' stream structure) are "
                    "reported beside it"}


def form_precisions(lbm, args):
    """fp64 and FP16-storage steps on the same instance; FP16 velocity error vs FP64."""
    import numpy as np

    nx, ny = args.nx, args.ny
    solid = lbm.make_channel(nx, ny)
    fluid = ~solid.astype(bool)
    out, ux = {}, {}
    for prec, steps in (("double", 200), ("fp16", 200)):
        s = lbm.Solver(nx, ny, 0.6, 1e-6, 0.0, solid, prec)
        s.init_equilibrium(1.0, 0.0, 0.0)
        m0 = s.total_mass()
        s.run(20)
        t = s.timed_run(steps - 20)
        mass_drift = (s.total_mass() - m0) / m0  # over the 200 steps the accuracy figures cover
        ux[prec] = s.macroscopic()[1][fluid].astype(np.float64)
        # the rate: the same solver after `settle` more steps (the engine clock ramps, run_plain)
        s.run(args.settle)
        t = s.timed_run(steps - 20)
        mlups = nx * ny * (steps - 20) / t / 1e6
        per_launch = 2 if s.steps_per_launch()["in_use"] else 1
        out[f"plain_{prec}"] = {"mlups": round(mlups, 1), "timed_steps": steps - 20,
                                "settle_steps": args.settle,
                                "steps_per_launch": per_launch,
                                "hbm_gbs_algorithmic": round(
                                    mlups * 1e6 * BYTES_PER_CELL[prec] / per_launch / 1e9, 1),
                                "mass_drift_fp64_reduction": mass_drift}
        s.close()
        log(f"{prec}: {mlups:.0f} MLUPS")
    d = ux["fp16"] - ux["double"]
    out["plain_fp16"]["ux_rel_l2_vs_fp64_after_200_steps"] = float(
        np.sqrt(np.sum(d * d) / np.sum(ux["double"] ** 2)))
    return out


def form_batched(lbm, args):
    doms = lbm.random_domains(seed=0, n=256)
    b = lbm.Batch(doms, "float")
    if args.batch_tile:  # a fixed two-step shape instead of the autotuner (profiling runs)
        btx, bty = (int(v) for v in args.batch_tile.split("x"))
        b.set_steps_per_launch(2, btx, bty, 0)
        tx = tr = wg = None  # the one-step shape is not used
    else:
        tx, tr, wg, tuned = b.autotune(steps=3)
    order, nt = b.order()
    two = b.steps_per_launch()
    layout = b.two_step_layout()
    b.set_graph_steps(100)
    b.run(100)
    t = b.timed_run(1000)
    cells = b.total_cells()
    mlups = cells * 1000 / t / 1e6
    b.close()
    per_launch = two["steps_per_launch"]
    log(f"batched: {mlups:.0f} MLUPS over {cells} cells ({per_launch} steps per launch, two-step "
        f"sub-tiles {two['tile_x']}x{two['tile_rows']} in {layout['items']} items, fill "
        f"{layout['fill']:.3f}; "
        f"one-step order {order}, tile {tx}x{tr}, {wg} WGs)")
    return {"domains": 256, "seed": 0, "cells": cells, "mlups": round(mlups, 1),
            "timed_steps": 1000, "graph_steps": 100, "steps_per_launch": per_launch,
            "two_step_subtile": f"{two['tile_x']}x{two['tile_rows']}",
            "two_step_band_rows_modelled": layout["modelled"],
            "two_step_items": layout["items"], "two_step_threads": layout["threads"],
            "two_step_fill": round(layout["fill"], 4),
            "two_step_workgroups": two["workgroups"],
            "two_step_grid": "one workgroup per work item" if two["workgroups"] == 0
            else "persistent",
            "one_step": {"order": order, "streaming_stores": nt, "tile_x": tx, "tile_rows": tr,
                         "workgroups": wg},
            "hbm_gbs_algorithmic": round(mlups * 1e6 * 72 / per_launch / 1e9, 1),
            "traffic": batch_traffic(lbm)}


def batch_traffic(lbm):
    """Calibrated traffic of the batched two-step kernel on these 256 domains (counter tool)."""
    s = traffic_summary(lbm, BATCH_TRAFFIC_SUMMARY)
    if not s:
        return None
    return {k: s.get(k) for k in ("file", "stale", "source_tag", "traffic_bytes_per_launch",
                                  "traffic_vs_algorithmic", "read_vs_algorithmic",
                                  "valu_instructions_per_update")}


def form_strip(lbm, args):
    """One GPU's share of 8 strips of the headline grid, 8192 x 1024 owned rows, with the ring
    closed on itself (ghost rows refreshed on the device after every launch): one ghost row
    (the reference's strip, one step per exchange) and two ghost rows (two fused steps per
    launch and per exchange)."""
    import numpy as np

    nx, nyl = args.nx, args.ny // 8
    out = {"shape": f"{nx}x({nyl}+2*ghost)", "timed_steps": 200, "runs": 3,
           "note": "one GPU's share of 8 strips, ring closed on itself: one ghost row, ghost "
                   "rows refreshed by a kernel after every step; two, each two-step launch "
                   "stores its halo into its own next ghost rows; mlups = best of 3 runs"}
    for ghost in (1, 2):
        st = lbm.Strip(nx, nyl, "float", ghost)
        st.set_params(0.6, 1e-6, 0.0)
        st.upload_solid(np.zeros((nyl + 2 * ghost, nx), np.uint8))
        st.timed_steps(1000)  # about 0.1 s of launches: the clocks settle after the upload
        ts = [st.timed_steps(200) for _ in range(3)]
        mlups = nx * nyl * 200 / min(ts) / 1e6
        st.close()
        per = 2 if ghost == 2 else 1
        out[f"ghost_{ghost}"] = {"mlups": round(mlups, 1), "steps_per_launch": per,
                                 "mlups_runs": [round(nx * nyl * 200 / s / 1e6, 1) for s in ts],
                                 "halo_values_per_message": (3 if ghost == 1 else 12) * nx}
        log(f"strip {nx}x({nyl}+2x{ghost}): {mlups:.0f} MLUPS, {per} steps per launch")
    out["mlups"] = out["ghost_2"]["mlups"]
    return out


def form_adjoint(lbm, args):
    import numpy as np

    nx, ny, T, avg = 512, 128, 2000, 200
    solid = lbm.add_disc(lbm.make_channel(nx, ny), nx, ny, 128.0, 64.0, 12.0)
    rng = np.random.default_rng(2024)
    F = rng.normal(0.0, 1e-7, (2, ny, nx))
    F[0] += 1e-5
    out = {}
    for snaps in (45, T + 1):
        a = lbm.Adjoint(nx, ny, 0.7, solid, T, (384, 48, 16, 32), avg, snaps, "float")
        a.set_force_field(F)
        a.forward()
        J, tf = a.forward()
        a.gradient()
        J2, dF, dtau, tg, fsteps = a.gradient()
        sched = a.schedule_info()
        out[f"snapshots_{snaps}"] = {"forward_ms": round(tf * 1e3, 3), "gradient_ms": round(tg * 1e3, 3),
                                     "cost_vs_forward": round(tg / tf, 3),
                                     "forward_steps": fsteps, "schedule": sched["kind"],
                                     "stored_states": sched["slots"], "levels": sched["levels"],
                                     "launches": sched["launches"], "J": J, "dJ_dtau": dtau}
        a.close()
        log(f"adjoint snapshots={snaps}: forward {tf * 1e3:.2f} ms, gradient {tg * 1e3:.2f} ms "
            f"(x{tg / tf:.2f}), {sched}")
    # the binomial schedule at the sqrt(T) budget, for comparison
    a = lbm.Adjoint(nx, ny, 0.7, solid, T, (384, 48, 16, 32), avg, 45, "float", "binomial")
    a.set_force_field(F)
    a.forward()
    J, tf = a.forward()
    a.gradient()
    _, _, _, tg, fsteps = a.gradient()
    sched = a.schedule_info()
    out["snapshots_45_binomial"] = {"forward_ms": round(tf * 1e3, 3), "gradient_ms": round(tg * 1e3, 3),
                                    "cost_vs_forward": round(tg / tf, 3), "forward_steps": fsteps,
                                    "schedule": sched["kind"], "stored_states": sched["slots"],
                                    "launches": sched["launches"]}
    a.close()
    log(f"adjoint binomial snapshots=45: gradient {tg * 1e3:.2f} ms (x{tg / tf:.2f})")
    return out


def form_cylinder(lbm, args):
    """The 1024 x 256 cylinder instance at Re 100: the body force holding the window-mean fluid
    velocity at U = 0.05 is found once by FP64 runs (cylinder_force), then float, double and
    FP16 storage run 40,000 steps; St from the lift (and the wake probe as a cross-check)."""
    t0 = time.time()
    force, uw, cal = lbm.cylinder_force()
    out = {"force": force, "calibration_runs": [[f, u] for f, u in cal],
           "calibration_s": round(time.time() - t0, 2)}
    for prec in ("float", "double", "fp16"):
        r = lbm.cylinder_instance(prec, force=force)
        out[prec] = {"st_lift_fft": round(r["st_lift_fft"], 5),
                     "st_lift_zero_crossing": round(r["st_lift_zero"], 5),
                     "lift_periods": r["lift_periods"],
                     "st_probe_fft": round(r["st_fft"], 5),
                     "st_probe_zero_crossing": round(r["st_zero"], 5),
                     "probe_periods": r["periods"],
                     "U_window": round(r["U_window"], 6), "Re_window": round(r["Re_window"], 2)}
        log(f"cylinder {prec}: Re {r['Re_window']:.2f}, lift St {r['st_lift_fft']:.4f} (FFT) / "
            f"{r['st_lift_zero']:.4f} (zero crossings, {r['lift_periods']} periods), probe St "
            f"{r['st_fft']:.4f} / {r['st_zero']:.4f}")
    out["strouhal"] = out["double"]["st_lift_fft"]
    out["signal"] = ("lift on the cylinder (momentum exchange, FP64) every 20 steps, second half "
                     "of the run; St = f D / U_window, U_window = window mean of the fluid "
                     "velocity (FP64 momentum / fluid cells)")
    out["fp16_st_rel_error_vs_fp64"] = abs(out["fp16"]["st_lift_fft"] - out["double"]["st_lift_fft"]) / out["double"]["st_lift_fft"]
    out["tau"] = r["tau"]
    return out


def form_edges(lbm, args):
    """Open x (inlet column, outlet column) and a per-cell force field on the two-step walk
    against one step per launch: a 4096 x 1024 walled channel with a disc of diameter 64,
    float and double, 200 timed steps after 20; the field is the adjoint form's (N(0, 1e-7)
    per cell from seed 0, plus gx = 1e-5)."""
    import numpy as np

    nx, ny = 4096, 1024
    solid = lbm.make_channel(nx, ny)
    lbm.add_disc(solid, nx, ny, 512.0, 512.0, 32.0)
    rng = np.random.default_rng(0)
    field = rng.normal(0.0, 1e-7, 2 * nx * ny)
    field[: nx * ny] += 1e-5
    out = {"grid": f"{nx}x{ny}", "steps": 200}
    for case in ("open_x", "force_field"):
        out[case] = {}
        for prec in ("float", "double"):
            rate = {}
            for spl in (2, 1):
                s = lbm.Solver(nx, ny, 0.6, 0.0 if case == "force_field" else 1e-6, 0.0, solid, prec)
                if case == "open_x":
                    s.set_inflow_outflow(0.05)
                else:
                    s.set_force_field(field.astype(np.float32 if prec == "float" else np.float64))
                s.set_steps_per_launch(spl)
                assert s.steps_per_launch()["in_use"] == (spl == 2)
                s.init_equilibrium(1.0, 0.05 if case == "open_x" else 0.0, 0.0)
                s.run(20)
                rate[spl] = nx * ny * 200 / s.timed_run(200) / 1e6
                s.close()
            out[case][prec] = {"two_step_mlups": round(rate[2], 1),
                               "single_step_mlups": round(rate[1], 1),
                               "speedup": round(rate[2] / rate[1], 3)}
            log(f"{case} {prec}: two steps per launch {rate[2]:.0f} MLUPS, one step "
                f"{rate[1]:.0f} MLUPS (x{rate[2] / rate[1]:.2f})")
    return out


FORMS = {"single_step": form_single_step, "protocol": form_protocol,
         "copy_ceiling": form_copy_ceiling,
         "precisions": form_precisions,
         "batched": form_batched, "strip": form_strip, "edges": form_edges,
         "adjoint": form_adjoint,
         "cylinder": form_cylinder}


def wanted_forms(args):
    return list(FORMS) if args.forms in ("all", None) else args.forms.split(",")


def forms(lbm, args, out=None):
    out = {} if out is None else out
    for name in wanted_forms(args):
        if name in out:
            continue
        fn = FORMS[name]
        t0 = time.time()
        try:
            out[name] = fn(lbm, args)
        except Exception as e:  # one form failing must not lose the headline line
            out[name] = {"error": f"{type(e).__name__}: {e}"}
            log(f"form {name} failed: {e}")
        log(f"form {name}: {time.time() - t0:.1f} s")
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=40)
    ap.add_argument("--nx", type=int, default=8192)
    ap.add_argument("--ny", type=int, default=8192)
    ap.add_argument("--settle", type=int, default=2000,
                    help="untimed steps before the warm-up (engine clock ramp, see run_plain)")
    ap.add_argument("--block", type=str, default=None, help="BXxROWS launch shape")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-forms", action="store_true", help="skip the other forms")
    ap.add_argument("--forms", default="all",
                    help="comma-separated subset of: " + ",".join(FORMS) + " (default all)")
    ap.add_argument("--batch-tile", default=None,
                    help="batched form: the two-step maximum tile WxH instead of autotuning")
    ap.add_argument("--cpu-full", action="store_true",
                    help="CPU baseline: the full protocol steps at one thread too (slow)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: the library computes one GPU's share and does "
              f"not shard the domain across GPUs (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)

    lbm = __graft_entry__.load_package()
    log(f"device: {lbm.device_info()}")
    done = {}
    if not args.no_forms and "copy_ceiling" in wanted_forms(args):
        done["copy_ceiling"] = form_copy_ceiling(lbm, args)
        log(f"copy ceiling: {done['copy_ceiling']['copy_kernel_gbs']:.0f} GB/s")
    block = tuple(int(v) for v in args.block.split("x")) if args.block else None
    t, info, mass = run_plain(lbm, args.nx, args.ny, args.steps, args.warmup, "float", block,
                              settle=args.settle)
    solver = json.loads("{" + info + "}")
    cells = args.nx * args.ny
    mlups = cells * args.steps / t / 1e6
    per_launch = solver.get("steps_per_launch", 1)  # 2: the two-step kernel
    launches = args.steps / per_launch
    launch_s = t / launches
    # each launch reads and writes every population once, whether it advances one or two steps
    achieved = BYTES_PER_CELL["float"] * cells / launch_s / 1e9
    log(f"plain fp32 {args.nx}x{args.ny}: {mlups:.1f} MLUPS, {t / args.steps * 1e3:.4f} ms/step, "
        f"{per_launch} steps per launch, {achieved:.0f} GB/s algorithmic per launch; {info}")

    line = {
        "metric": "MLUPS & % HBM roofline, 8192² fp32 plain vs fp16-store/batched; adjoint cost",
        "value": round(mlups, 2),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "settle_steps": args.settle,
        "ms_per_step": round(t / args.steps * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"plain fp32 D2Q9 step (SoA pull + BGK + Guo + halfway "
                               f"bounce-back, f-w storage), {args.nx}x{args.ny} walled channel, "
                               f"tau 0.6, gx 1e-6, rest init",
                   "nx": args.nx, "ny": args.ny, "parallelism": "single GPU",
                   "solver": solver},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None,
                     "kernel": "step2_kernel<float>" if per_launch == 2 else "step_kernel<float>",
                     "updates_per_launch": per_launch,
                     "bytes_per_launch": BYTES_PER_CELL["float"] * cells,
                     "launch_ms": round(launch_s * 1e3, 5)},
    }
    ts = traffic_summary(lbm) if per_launch == 2 else None
    valu = None
    if ts:
        line["roofline"]["traffic"] = ts["traffic_bytes_per_launch"]
        line["roofline"]["traffic_source"] = {
            "file": ts["file"], "stale": ts["stale"], "source_tag": ts["source_tag"],
            "traffic_vs_algorithmic": ts["traffic_vs_algorithmic"],
            "calibration": ts["calibration"]}
        valu = ts.get("valu_instructions_per_launch")
    if valu and per_launch == 2:
        # the two-step kernel is bound by vector-ALU issue as much as by HBM: the profiled VALU
        # instruction count of a launch against the issue ceiling, over the launch time measured here
        line["roofline"]["valu_issue_frac"] = round(valu / VALU_WAVE_INSTR_PER_S / launch_s, 4)
        line["roofline"]["valu_instructions_per_launch"] = valu
    if not args.no_forms:
        line["forms"] = forms(lbm, args, done)
        f = line["forms"]
        cc = f.get("copy_ceiling", {})
        if cc.get("mlups_ceiling_fp32"):
            # MLUPS against the copy ceiling / 72 B per update; >1 once a launch advances two steps
            f["plain_float_mlups_vs_copy_ceiling"] = round(mlups / cc["mlups_ceiling_fp32"], 4)
            f["plain_float_launch_bytes_frac_of_copy_ceiling"] = round(
                achieved / max(cc["copy_kernel_gbs"], cc["memcpy_dtod_gbs"]), 4)
            ss = f.get("single_step", {})
            if ss.get("mlups"):
                f["single_step_frac_of_copy_ceiling"] = round(ss["mlups"] / cc["mlups_ceiling_fp32"], 4)
                f["single_step_frac_of_nine_array_copy"] = round(
                    ss["hbm_gbs_algorithmic"] / cc["nine_array_copy_gbs"], 4)
        pr, ss = f.get("precisions", {}), f.get("single_step", {})
        # like for like: both one step per launch (the FP16 form's >= 1.7 x target), both two
        if ss.get("mlups") and ss.get("fp16", {}).get("mlups"):
            f["fp16_vs_fp32_one_step_per_launch"] = round(ss["fp16"]["mlups"] / ss["mlups"], 3)
        if pr.get("plain_fp16", {}).get("mlups"):
            f["fp16_vs_fp32_two_steps_per_launch"] = round(pr["plain_fp16"]["mlups"] / mlups, 3)
            if ss.get("mlups"):
                f["fp16_default_vs_fp32_one_step_per_launch"] = round(
                    pr["plain_fp16"]["mlups"] / ss["mlups"], 3)
    if not args.no_cpu:
        log("cpu baseline ...")
        try:
            line["cpu_baseline"] = cpu_baseline(args.cpu_full)
            log(f"cpu baseline: {line['cpu_baseline']['value']:.1f} MLUPS on "
                f"{line['cpu_baseline']['cores']} threads ({line['cpu_baseline']['kind']})")
        except Exception as e:  # the CPU leg failing must not lose the headline line
            line["cpu_baseline"] = {"error": f"{type(e).__name__}: {e}"}
            log(f"cpu baseline failed: {e}")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
