"""The reference's data formats on either side of the step, written from this library's solvers
so that the reference's analysis and plotting scripts read MI355X runs unchanged.

* Field files (src/common/io.cpp:11-52, driver.hpp:133-163,236-249): rho.npy, ux.npy, uy.npy,
  vorticity.npy (the solver real, shape (ny, nx)), solid.npy (uint8) and meta.json; the
  cylinder case adds probe.csv and frames/vort_<step>.npy (driver.hpp:186-232). The .npy files
  have the reference's bytes: format 1.0, the header dict padded with spaces so that the data
  starts at the first 64-byte boundary (numpy's own writer pads further). meta.json and
  probe.csv print numbers as C++ streams do by default (%g, 6 significant digits).
* Benchmark protocol (driver.hpp:115-132, bench/run_bench.py:24-31,127-160): one BENCH record
  per timed repetition after an untimed warm-up, the CSV row run_bench.py aggregates them into
  (same columns: the CUDA-specific ones stay empty, the HIP versions go to gpu_driver), the copy
  ceiling CSV of bench/gpu_bandwidth.cu:92-96, and bench/ceiling.py's MLUPS ceiling.
* Strouhal report (scripts/strouhal.py:57-100): the analysis of a cylinder run's probe.csv and
  the strouhal.csv row.
"""
from __future__ import annotations

import csv
import datetime
import json
import os
import pathlib
import platform
import statistics
import struct

import numpy as np

# bench/run_bench.py:24-33
BENCH_COLUMNS = [
    "version", "nx", "ny", "steps", "warmup", "reps", "threads",
    "time_mean_s", "time_std_s", "mlups_mean", "mlups_max", "mlups_std",
    "real", "cpu", "governor", "compiler", "flags", "build_type",
    "omp_proc_bind", "omp_places", "date", "commit",
    "gpu", "gpu_driver", "gpu_driver_cuda", "cuda_runtime", "cuda_compiled", "ecc", "block",
    "variant",
]
GPU_FIELDS = ["gpu", "gpu_driver_cuda", "cuda_runtime", "cuda_compiled", "ecc", "block",
              "variant"]
# bench/gpu_bandwidth.cu:92-94
BANDWIDTH_COLUMNS = ["gpu", "copy_MBps", "memcpy_MBps", "bytes_per_array", "reps", "ecc",
                     "gpu_driver_cuda", "cuda_runtime", "cuda_compiled", "date"]
# scripts/strouhal.py:84-100 (results/strouhal.csv)
STROUHAL_COLUMNS = ["run", "version", "real", "nx", "ny", "diameter", "u_in", "re", "tau",
                    "steps", "window_start", "st_fft", "st_zero_crossing", "periods",
                    "uy_amplitude", "commit", "date"]

VERSION = "v5_hip"
Q = 9
_DESCR = {np.dtype(np.float32): "<f4", np.dtype(np.float64): "<f8", np.dtype(np.uint8): "|u1"}


# ---- fields -----------------------------------------------------------------------------------

def npy_bytes(a, ny: int, nx: int) -> bytes:
    """The bytes of the reference's write_npy (io.cpp:11-52) for an (ny, nx) float32, float64 or
    uint8 field: magic, version 1.0, little-endian header length, the header dict padded with
    spaces and a newline to the next 64-byte boundary, then the data in C order."""
    a = np.ascontiguousarray(a)
    descr = _DESCR.get(a.dtype)
    if descr is None:
        raise ValueError("write_npy takes float32, float64 or uint8 data")
    if a.size != nx * ny:
        raise ValueError("write_npy: size does not match shape")
    head = f"{{'descr': '{descr}', 'fortran_order': False, 'shape': ({ny}, {nx}), }}"
    total = (10 + len(head) + 1 + 63) // 64 * 64
    head = head + " " * (total - 10 - len(head) - 1) + "\n"
    return (b"\x93NUMPY\x01\x00" + struct.pack("<H", len(head)) + head.encode("latin1")
            + a.astype(a.dtype.newbyteorder("<"), copy=False).tobytes())


def write_npy(path, a, ny: int, nx: int) -> None:
    pathlib.Path(path).write_bytes(npy_bytes(a, ny, nx))


def cformat(v) -> str:
    """A value as `std::ostream << v` prints it with default flags: integers in full, floating
    values as %g (6 significant digits), strings quoted (the reference writes them quoted)."""
    if isinstance(v, str):
        return f'"{v}"'
    if isinstance(v, (bool, np.bool_)):
        raise TypeError("booleans have no meta.json form")
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    return "%g" % float(v)


def meta_text(items) -> str:
    """A meta.json line from (key, value) pairs in the reference's key order."""
    return "{" + ", ".join(f'"{k}": {cformat(v)}' for k, v in items) + "}\n"


def channel_meta(nx, ny, tau, umax, force_x, steps, real, case="poiseuille", version=VERSION):
    """driver.hpp:157-162 (tau, umax, force_x: the solver's real values)."""
    return [("version", version), ("case", case), ("nx", int(nx)), ("ny", int(ny)),
            ("tau", float(tau)), ("umax", float(umax)), ("force_x", float(force_x)),
            ("steps", int(steps)), ("real", real)]


def cylinder_meta(nx, ny, tau, u_in, re, diameter, xc, yc, probe_x, probe_y, probe_every,
                  frame_every, steps, real, version=VERSION):
    """driver.hpp:242-249."""
    return [("version", version), ("case", "cylinder"), ("nx", int(nx)), ("ny", int(ny)),
            ("tau", float(tau)), ("u_in", float(u_in)), ("re", float(re)),
            ("diameter", int(diameter)), ("xc", float(xc)), ("yc", float(yc)),
            ("probe_x", int(probe_x)), ("probe_y", int(probe_y)),
            ("probe_every", int(probe_every)), ("frame_every", int(frame_every)),
            ("steps", int(steps)), ("real", real)]


def real_name(solver) -> str:
    """The reference's real_name of the solver's arithmetic: "double", or "float" (FP16
    storage computes in float)."""
    return "double" if np.dtype(solver.dtype) == np.float64 else "float"


def write_fields(solver, out_dir, meta, periodic_x: int = 1) -> None:
    """rho, ux, uy, vorticity, solid of the solver's current state and meta.json (driver.hpp:
    145-162, 236-249). periodic_x: 1 the channel's periodic x (io.hpp:20-21 default), 0 the
    cylinder's one-sided open ends. The vorticity is computed on the device."""
    out = pathlib.Path(out_dir)
    out.mkdir(parents=True, exist_ok=True)
    ny, nx = solver.ny(), solver.nx()
    rho, ux, uy = solver.macroscopic()
    write_npy(out / "rho.npy", rho, ny, nx)
    write_npy(out / "ux.npy", ux, ny, nx)
    write_npy(out / "uy.npy", uy, ny, nx)
    write_npy(out / "vorticity.npy", solver.vorticity(periodic_x), ny, nx)
    write_npy(out / "solid.npy", np.asarray(solver.solid(), np.uint8), ny, nx)
    (out / "meta.json").write_text(meta_text(meta))


def probe_moments(f, dtype):
    """(ux, uy, rho) at a probe from its nine populations f - w, in the solver real with the
    reference's sequential sums (driver.hpp:205-214)."""
    t = np.dtype(dtype).type
    m0 = mx = my = t(0)
    cx = (0, 1, 0, -1, 0, 1, -1, -1, 1)
    cy = (0, 0, 1, 0, -1, 1, 1, -1, -1)
    for i in range(Q):
        fi = t(f[i])
        m0 = t(m0 + fi)
        mx = t(mx + t(cx[i]) * fi)
        my = t(my + t(cy[i]) * fi)
    rho = t(t(1) + m0)
    return float(t(mx / rho)), float(t(my / rho)), float(rho)


def probe_line(step: int, ux: float, uy: float, rho: float) -> str:
    """One probe.csv line (driver.hpp:215-216)."""
    return f"{int(step)},{cformat(ux)},{cformat(uy)},{cformat(rho)}\n"


def frame_name(step: int) -> str:
    """driver.hpp:224."""
    return "frames/vort_%07d.npy" % step


# ---- benchmark protocol -----------------------------------------------------------------------

def build_flags() -> str:
    """The device flags the library is built with (the reference reports its compiler flags in
    every BENCH record)."""
    import importlib.util

    spec = importlib.util.spec_from_file_location(
        "lbm_hip_build_flags", pathlib.Path(__file__).resolve().parent / "build.py")
    b = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(b)
    return " ".join(f for f in b.flags() if not f.startswith(("-I", "-W")))


def bench_record(version, rep, nx, ny, steps, warmup, threads, time_s, real, compiler, flags,
                 build_type, info: str = "") -> str:
    """The text of one BENCH line (driver.hpp:122-129): fixed keys, then the solver's
    bench_info() fragment; time_s %.6f, mlups %.4f."""
    mlups = nx * ny * steps / time_s / 1e6
    extra = ", " + info if info else ""
    return (f'BENCH {{"version": "{version}", "rep": {rep}, "nx": {nx}, "ny": {ny}, '
            f'"steps": {steps}, "warmup": {warmup}, "threads": {threads}, '
            f'"time_s": {time_s:.6f}, "mlups": {mlups:.4f}, "real": "{real}", '
            f'"compiler": "{compiler}", "flags": "{flags}", "build_type": "{build_type}"{extra}}}')


def parse_bench_lines(text: str):
    """BENCH records from a run's output (bench/run_bench.py:96-97)."""
    return [json.loads(line[len("BENCH "):]) for line in text.splitlines()
            if line.startswith("BENCH ")]


def protocol_steps(nx: int, ny: int, updates: float = 5e7) -> tuple[int, int]:
    """Steps and warm-up run_bench.py uses for a size (run_bench.py:137-139)."""
    steps = max(10, round(updates / (nx * ny)))
    return steps, max(1, steps // 5)


def cpu_model() -> str:
    try:
        for line in pathlib.Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def governor() -> str:
    try:
        return pathlib.Path("/sys/devices/system/cpu/cpu0/cpufreq/scaling_governor").read_text().strip()
    except OSError:
        return "unknown"


def bench_row(recs, label: str | None = None, common: dict | None = None) -> dict:
    """run_bench.py's CSV row (run_bench.py:141-160) from the BENCH records of one run:
    statistics.mean / stdev of time_s and mlups. `common` overrides the machine columns (cpu,
    governor, commit, date, omp_*); gpu_driver gets the HIP driver version."""
    times = [r["time_s"] for r in recs]
    mlups = [r["mlups"] for r in recs]
    r0 = recs[0]
    row = {
        "version": label or r0["version"], "nx": r0["nx"], "ny": r0["ny"], "steps": r0["steps"],
        "warmup": r0["warmup"], "reps": len(recs), "threads": r0["threads"],
        "time_mean_s": f"{statistics.mean(times):.6f}",
        "time_std_s": f"{statistics.stdev(times) if len(times) > 1 else 0.0:.6f}",
        "mlups_mean": f"{statistics.mean(mlups):.3f}",
        "mlups_max": f"{max(mlups):.3f}",
        "mlups_std": f"{statistics.stdev(mlups) if len(mlups) > 1 else 0.0:.3f}",
        "real": r0["real"], "compiler": r0["compiler"], "flags": r0["flags"],
        "build_type": r0["build_type"],
        "cpu": cpu_model(), "governor": governor(), "commit": "none",
        "date": datetime.datetime.now().isoformat(timespec="seconds"),
        "omp_proc_bind": os.environ.get("OMP_PROC_BIND", "close"),
        "omp_places": os.environ.get("OMP_PLACES", "cores"),
        **{k: r0.get(k, "") for k in GPU_FIELDS},
        "gpu_driver": str(r0.get("hip_driver", "")) if r0.get("gpu") else "",
    }
    row.update(common or {})
    return row


def write_csv(rows, path, columns, append: bool = False) -> None:
    """Rows under `columns` with a header; appended without one to an existing file when
    `append` (run_bench.py:175-186)."""
    path = pathlib.Path(path)
    add = append and path.exists()
    path.parent.mkdir(parents=True, exist_ok=True)
    with open(path, "a" if add else "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=columns)
        if not add:
            w.writeheader()
        w.writerows(rows)


def bandwidth_row(gpu: str, copy_mbps: float, memcpy_mbps: float, bytes_per_array: int,
                  reps: int, ecc: int | str = "", date: str | None = None) -> dict:
    """The copy-ceiling CSV row of bench/gpu_bandwidth.cu:92-96 (rates in MB/s, %.1f); the
    CUDA version columns stay empty."""
    return {"gpu": gpu, "copy_MBps": f"{copy_mbps:.1f}", "memcpy_MBps": f"{memcpy_mbps:.1f}",
            "bytes_per_array": int(bytes_per_array), "reps": int(reps), "ecc": ecc,
            "gpu_driver_cuda": "", "cuda_runtime": "", "cuda_compiled": "",
            "date": date or datetime.datetime.now().isoformat(timespec="seconds")}


def ceiling_mlups(rows, real_size: int, peak_gbs: float | None = None, gpu: bool = True):
    """bench/ceiling.py:42-59: the best measured copy rate (max of copy_MBps and memcpy_MBps over
    the rows) over 2 * 9 * sizeof(real) bytes per update, and, with peak_gbs, the hardware
    ceiling over 2 (GPU) or 3 (CPU write-allocate) * 9 * sizeof(real). Returns (measured,
    hardware or None) in MLUPS."""
    def rate(r):
        return max(float(r["copy_MBps"]), float(r.get("memcpy_MBps") or 0.0))
    copy_bps = max(rate(r) for r in rows) * 1e6
    measured = copy_bps / (2 * Q * real_size) / 1e6
    hw = None
    if peak_gbs:
        hw = peak_gbs * 1e9 / ((2 if gpu else 3) * Q * real_size) / 1e6
    return measured, hw


def strouhal_analysis(out_dir, start: float = 0.5) -> dict:
    """The Strouhal analysis of a cylinder run's output directory (scripts/strouhal.py:57-81):
    probe.csv and meta.json read back as written, samples with step >= start * steps kept and
    their mean removed; St from the Hann-windowed, 16x zero-padded FFT peak refined by a
    parabola on the log spectrum, and from the mean period between upward zero crossings
    (linear interpolation); both times D / u_in. Returns st_fft, st_zero, periods,
    uy_amplitude (max |u_y - mean| over the window), window_start and the meta dict."""
    out = pathlib.Path(out_dir)
    meta = json.loads((out / "meta.json").read_text())
    data = np.genfromtxt(out / "probe.csv", delimiter=",", names=True)
    step, uy = data["step"], data["uy"]
    keep = step >= start * meta["steps"]
    step, uy = step[keep], uy[keep] - uy[keep].mean()
    dt = float(step[1] - step[0])
    D, U = meta["diameter"], meta["u_in"]
    n = len(uy)
    spec = np.abs(np.fft.rfft(uy * np.hanning(n), 16 * n))
    freqs = np.fft.rfftfreq(16 * n, d=dt)
    k = int(np.argmax(spec[1:]) + 1)
    a, b, c = np.log(spec[k - 1:k + 2] + 1e-300)
    f_fft = (k + 0.5 * (a - c) / (a - 2 * b + c)) * (freqs[1] - freqs[0])
    up = np.where((uy[:-1] < 0) & (uy[1:] >= 0))[0]
    t_cross = step[up] - uy[up] * (step[up + 1] - step[up]) / (uy[up + 1] - uy[up])
    periods = np.diff(t_cross)
    f_zero = 1.0 / periods.mean() if len(periods) else float("nan")
    return {"st_fft": f_fft * D / U, "st_zero": f_zero * D / U, "periods": int(len(periods)),
            "uy_amplitude": float(np.abs(uy).max()), "window_start": int(start * meta["steps"]),
            "meta": meta}


def strouhal_row(run, version, real, nx, ny, diameter, u_in, re, tau, steps, window_start,
                 st_fft, st_zero, periods, uy_amplitude, commit="none", date=None) -> dict:
    """scripts/strouhal.py:84-100 row: tau %.5f, Strouhal numbers %.4f, amplitude %.5f."""
    return {"run": run, "version": version, "real": real, "nx": nx, "ny": ny,
            "diameter": diameter, "u_in": u_in, "re": re, "tau": f"{tau:.5f}", "steps": steps,
            "window_start": window_start, "st_fft": f"{st_fft:.4f}",
            "st_zero_crossing": f"{st_zero:.4f}", "periods": periods,
            "uy_amplitude": f"{uy_amplitude:.5f}", "commit": commit,
            "date": date or datetime.datetime.now().isoformat(timespec="seconds")}
