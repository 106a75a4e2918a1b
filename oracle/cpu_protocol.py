"""ORACLE — test infrastructure only: the CPU baseline timed on the reference's protocol.

Runs the reference's own v1 OpenMP solver (oracle/_ref, compiled from its sources) or, when that
is absent, the oracle's OpenMP pull restatement, on one instance: untimed warm-up of steps / 5,
then `reps` timed runs of `steps` steps in one process (src/common/driver.hpp:115-132,
bench/run_bench.py:138-139), and prints one JSON object: MLUPS per run, mean, std (population,
as run_bench.py's numpy std), max, the OpenMP threads used and which solver ran.

bench.py starts it as a child process so that the OpenMP placement it asks for
(OMP_NUM_THREADS, OMP_PROC_BIND=close, OMP_PLACES=cores: bench/run_bench.py:124-128) is in the
environment before the OpenMP runtime starts.

    python -m oracle.cpu_protocol --nx 2048 --ny 2048 --steps 1000 --mask block --tau 0.8 \\
        --fx 1e-5 --fy 2e-6
"""
from __future__ import annotations

import argparse
import json
import pathlib
import sys

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))

import oracle  # noqa: E402


def seeded_block(seed: int, nx: int, ny: int, size: int = 8) -> np.ndarray:
    """Walled channel plus the seeded size x size block (the map of the package's
    seeded_block_mask and of tests/golden/make_golden.py: numpy default_rng(seed), x0 uniform in
    [0, nx - size], y0 uniform in [1, ny - size - 1])."""
    m = oracle.make_channel(nx, ny).reshape(ny, nx)
    rng = np.random.default_rng(seed)
    x0 = int(rng.integers(0, nx - size + 1))
    y0 = int(rng.integers(1, ny - size))
    m[y0:y0 + size, x0:x0 + size] = 1
    return m.reshape(-1)


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--nx", type=int, required=True)
    ap.add_argument("--ny", type=int, required=True)
    ap.add_argument("--steps", type=int, required=True)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=-1, help="default steps / 5")
    ap.add_argument("--mask", choices=("channel", "block"), default="channel")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--tau", type=float, default=0.6)
    ap.add_argument("--fx", type=float, default=1e-6)
    ap.add_argument("--fy", type=float, default=0.0)
    a = ap.parse_args()
    warmup = a.warmup if a.warmup >= 0 else max(1, a.steps // 5)
    solid = (oracle.make_channel(a.nx, a.ny) if a.mask == "channel"
             else seeded_block(a.seed, a.nx, a.ny))
    if oracle.Reference.available("float"):
        secs, threads = oracle.Reference("float").v1_protocol(
            a.nx, a.ny, a.tau, a.fx, a.fy, solid, warmup, a.steps, a.reps)
        kind = "reference"
    else:
        o = oracle.Oracle(a.nx, a.ny, a.tau, a.fx, a.fy, solid, "float", "pull")
        o.init_equilibrium(1.0, 0.0, 0.0)
        o.run(warmup)
        secs = [o.timed_run(a.steps) for _ in range(a.reps)]
        threads = o.threads()
        kind = "port"
    m = np.array([a.nx * a.ny * a.steps / s / 1e6 for s in secs])
    print(json.dumps({"mlups": [round(float(v), 2) for v in m], "mean": round(float(m.mean()), 2),
                      "std": round(float(m.std()), 2), "max": round(float(m.max()), 2),
                      "threads": threads, "kind": kind, "steps": a.steps, "warmup": warmup,
                      "reps": a.reps, "seconds": [float(f"{s:.6g}") for s in secs]}))


if __name__ == "__main__":
    main()
