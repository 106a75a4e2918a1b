#!/usr/bin/env python3
"""The reference's solver executables (lbm_v0 ... lbm_v4: run_driver, src/common/driver.hpp:
255-300, command line of src/common/params.cpp:8-44) on the MI355X solver.

    driver.py [--case poiseuille|cylinder] [--nx N] [--ny N] [--tau T] [--umax U] [--steps N]
              [--out DIR] [--bench] [--warmup N] [--reps N] [--re R] [--diameter D]
              [--probe-every N] [--frame-every N] [--frame-start N] [--real float|double|fp16]

Same options, defaults (params.hpp:10-29), checks and outputs as the reference:
  * poiseuille: the walled channel driven by poiseuille_force; writes rho/ux/uy/vorticity/solid
    .npy and meta.json to --out (driver.hpp:134-163);
  * --bench: `warmup` untimed steps, then `reps` timed runs of `steps`, one `BENCH {...}` line
    each (driver.hpp:115-132), which bench/run_bench.py parses (`--exe` this file);
  * cylinder: open inlet/outlet with the disc at (5 D, ny/2 + 1), probe.csv every
    --probe-every steps, optional vorticity frames, final fields (driver.hpp:167-251).
--real selects the arithmetic (the reference fixes it at build time; LBM_REAL sets the
default); LBM_BLOCK=BXxR sets the launch shape as for the reference's GPU solvers.
Invalid options print `error: ...` and exit with status 1 (driver.hpp:293-296).
"""
from __future__ import annotations

import importlib.util
import math
import os
import pathlib
import sys

HERE = pathlib.Path(__file__).resolve().parent
MODULE = "lbm_amd"


def package():
    if MODULE in sys.modules:
        return sys.modules[MODULE]
    spec = importlib.util.spec_from_file_location(MODULE, HERE / "__init__.py",
                                                  submodule_search_locations=[str(HERE)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[MODULE] = mod
    spec.loader.exec_module(mod)
    return mod


class Params:
    """params.hpp:10-29 defaults."""

    def __init__(self):
        self.case_name = "poiseuille"
        self.nx, self.ny = 64, 34
        self.tau, self.u_max = 0.9, 0.05
        self.steps = 20000
        self.out_dir = "output"
        self.re, self.diameter = 100.0, 40
        self.probe_every, self.frame_every, self.frame_start = 50, 0, 0
        self.bench, self.warmup, self.reps = False, 100, 5
        self.real = os.environ.get("LBM_REAL", "float")


_INT = {"--nx": "nx", "--ny": "ny", "--steps": "steps", "--warmup": "warmup", "--reps": "reps",
        "--diameter": "diameter", "--probe-every": "probe_every",
        "--frame-every": "frame_every", "--frame-start": "frame_start"}
_FLOAT = {"--tau": "tau", "--umax": "u_max", "--re": "re"}
_STR = {"--case": "case_name", "--out": "out_dir", "--real": "real"}


def _number(conv, key, val):
    try:
        return conv(val)
    except ValueError:
        raise ValueError(f"invalid value {val!r} for {key}") from None


def parse_args(argv) -> Params:
    """params.cpp:8-44: `--key value` pairs and the --bench flag; raises ValueError on unknown
    keys, missing values and the reference's range checks."""
    p = Params()
    a = 0
    while a < len(argv):
        key = argv[a]
        if key == "--bench":
            p.bench = True
            a += 1
            continue
        if a + 1 >= len(argv):
            raise ValueError("missing value for " + key)
        val = argv[a + 1]
        if key in _INT:
            setattr(p, _INT[key], _number(int, key, val))
        elif key in _FLOAT:
            setattr(p, _FLOAT[key], _number(float, key, val))
        elif key in _STR:
            setattr(p, _STR[key], val)
        else:
            raise ValueError("unknown option " + key)
        a += 2
    if p.nx < 1 or p.ny < 3:
        raise ValueError("grid too small")
    if not p.tau > 0.5:
        raise ValueError("tau must be > 0.5 (positive viscosity)")
    if p.steps < 0:
        raise ValueError("steps must be >= 0")
    if p.warmup < 0 or p.reps < 1:
        raise ValueError("need warmup >= 0 and reps >= 1")
    if p.re <= 0 or p.diameter < 2 or p.probe_every < 1 or p.frame_every < 0:
        raise ValueError("invalid cylinder parameters")
    if p.real not in ("float", "double", "fp16"):
        raise ValueError("--real must be float, double or fp16")
    return p


def _as_real(x, dtype):
    import numpy as np

    return float(np.dtype(dtype).type(x))


def _compiler(info: dict) -> str:
    v = int(info.get("hip_runtime", 0))
    return f"hipcc (HIP {v // 10000000}.{v // 100000 % 100}.{v % 100000})"


def run_bench(lbm, solver, p: Params, version: str, out=None) -> None:
    """driver.hpp:115-132."""
    import json

    from lbm_amd import formats

    out = out or sys.stdout
    solver.run(p.warmup)
    frag = solver.bench_info()
    info = json.loads("{" + frag + "}")
    flags = formats.build_flags()
    real = formats.real_name(solver)
    for r in range(p.reps):
        t = solver.timed_run(p.steps)
        print(formats.bench_record(version, r, p.nx, p.ny, p.steps, p.warmup, solver.threads(),
                                   t, real, _compiler(info), flags, "Release", frag),
              file=out, flush=True)


def run_and_write(lbm, solver, p: Params, version: str, force: float) -> None:
    """driver.hpp:134-163."""
    from lbm_amd import formats

    mass0 = solver.total_mass()
    seconds = solver.timed_run(p.steps)
    mlups = p.nx * p.ny * p.steps / seconds / 1e6 if seconds > 0 else float("inf")
    mass1 = solver.total_mass()
    print(f"time={seconds:.3f} s  MLUPS={mlups:.2f} (indicative)  mass drift="
          f"{(mass1 - mass0) / mass0:.3e}")
    real = formats.real_name(solver)
    meta = formats.channel_meta(p.nx, p.ny, _as_real(p.tau, solver.dtype),
                                _as_real(p.u_max, solver.dtype), force, p.steps, real,
                                case=p.case_name, version=version)
    formats.write_fields(solver, p.out_dir, meta, periodic_x=1)
    print(f"fields written to {p.out_dir}")


def cylinder_setup(p: Params, dtype):
    """driver.hpp:171-183: centre (5 D, ny/2 + 1), probe (int(xc + 3 D), int(yc)),
    tau = 0.5 + 3 nu in the solver real with nu = u_max D / Re."""
    xc = 5.0 * p.diameter
    yc = float(p.ny // 2 + 1)
    u = _as_real(p.u_max, dtype)
    nu = u * p.diameter / p.re
    return {"xc": xc, "yc": yc, "probe_x": int(xc + 3.0 * p.diameter), "probe_y": int(yc),
            "tau": _as_real(0.5 + 3.0 * nu, dtype)}


def run_cylinder(lbm, solver, p: Params, version: str, cs: dict) -> None:
    """driver.hpp:186-251."""
    from lbm_amd import formats

    out = pathlib.Path(p.out_dir)
    (out / "frames").mkdir(parents=True, exist_ok=True)
    seconds, n, frames = 0.0, 0, 0
    with open(out / "probe.csv", "w", newline="") as probe:
        probe.write("step,ux,uy,rho\n")
        while n < p.steps:
            chunk = min(p.probe_every, p.steps - n)
            seconds += solver.timed_run(chunk)
            n += chunk
            ux, uy, rho = formats.probe_moments(solver.populations_at(cs["probe_x"],
                                                                      cs["probe_y"]),
                                                solver.dtype)
            probe.write(formats.probe_line(n, ux, uy, rho))
            if not math.isfinite(rho):
                raise RuntimeError("the simulation diverged (non-finite density at the probe)")
            if p.frame_every > 0 and n >= p.frame_start and n % p.frame_every == 0:
                formats.write_npy(out / formats.frame_name(n), solver.vorticity(0), p.ny, p.nx)
                frames += 1
    mlups = p.nx * p.ny * p.steps / seconds / 1e6 if seconds > 0 else float("inf")
    print(f"time={seconds:.3f} s  MLUPS={mlups:.2f} (indicative, without I/O)  frames={frames}")
    real = formats.real_name(solver)
    meta = formats.cylinder_meta(p.nx, p.ny, cs["tau"], _as_real(p.u_max, solver.dtype), p.re,
                                 p.diameter, cs["xc"], cs["yc"], cs["probe_x"], cs["probe_y"],
                                 p.probe_every, p.frame_every, p.steps, real, version=version)
    formats.write_fields(solver, out, meta, periodic_x=0)
    print(f"probe, frames and fields written to {p.out_dir}")


def run_driver(argv, version: str = "v5_hip") -> int:
    """driver.hpp:255-300."""
    try:
        p = parse_args(argv)
        lbm = package()
        dtype = lbm.real_dtype(p.real)
        rname = "double" if p.real == "double" else "float"
        if p.case_name == "cylinder":
            cs = cylinder_setup(p, dtype)
            solid = lbm.make_cylinder(p.nx, p.ny, cs["xc"], cs["yc"], float(p.diameter))
            with lbm.Solver(p.nx, p.ny, cs["tau"], 0.0, 0.0, solid, p.real) as s:
                print(f"case=cylinder nx={p.nx} ny={p.ny} D={p.diameter} Re={p.re:g} "
                      f"u_in={_as_real(p.u_max, dtype):g} -> tau={cs['tau']:.4f}, "
                      f"Ma={math.sqrt(3.0) * _as_real(p.u_max, dtype):.3f}, steps={p.steps} "
                      f"real={rname} out={p.out_dir}", flush=True)
                s.set_inflow_outflow(p.u_max)
                s.init_equilibrium(1.0, p.u_max, 0.0)
                run_cylinder(lbm, s, p, version, cs)
            return 0
        if p.case_name != "poiseuille":
            print(f"unknown case '{p.case_name}' (available: poiseuille, cylinder)",
                  file=sys.stderr)
            return 1
        H = p.ny - 2
        force = lbm.poiseuille_force(H, p.tau, p.u_max, p.real)
        with lbm.Solver(p.nx, p.ny, p.tau, force, 0.0, lbm.make_channel(p.nx, p.ny),
                        p.real) as s:
            if not p.bench:
                print(f"case={p.case_name} nx={p.nx} ny={p.ny} tau={_as_real(p.tau, dtype):g} "
                      f"umax={_as_real(p.u_max, dtype):g} steps={p.steps} real={rname} "
                      f"out={p.out_dir}", flush=True)
            s.init_equilibrium(1.0, 0.0, 0.0)
            if p.bench:
                run_bench(lbm, s, p, version)
            else:
                run_and_write(lbm, s, p, version, force)
    except Exception as e:  # the reference catches std::exception and returns 1
        print(f"error: {e}", file=sys.stderr)
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(run_driver(sys.argv[1:]))
