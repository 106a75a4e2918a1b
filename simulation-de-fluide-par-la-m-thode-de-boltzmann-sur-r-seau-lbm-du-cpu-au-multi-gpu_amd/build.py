"""Builds liblbm_hip.so (gfx950) in-tree with hipcc.

Every translation unit is compiled with -ffp-contract=off and IEEE division (no fast-math): the
float and double steps are then bit-identical to the reference's v0. Packed-FP32 instruction
selection is disabled (-packed-fp32-ops). Objects go to build/ next to this file, the library to
liblbm_hip.so beside it.

Usage: python build.py [--jobs N]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import pathlib
import shutil
import subprocess
import sys

HERE = pathlib.Path(__file__).resolve().parent
ROOT = HERE.parent
CSRC = HERE / "csrc"
OBJ = HERE / "build"
LIB = HERE / "liblbm_hip.so"

ARCH = os.environ.get("LBM_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm toolchain required)")


def flags() -> list[str]:
    return ["-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wall", "-Wno-unused-command-line-argument", "-Wno-unused-function",
            f"-I{ROOT / 'include'}", f"-I{CSRC}"]


def compile_one(src: str) -> pathlib.Path:
    path = CSRC / src
    obj = OBJ / (src.rsplit(".", 1)[0] + ".o")
    cmd = [hipcc(), *flags()]
    if src.endswith(".cpp"):
        cmd += ["-x", "hip"]  # host code with the HIP runtime; no device code inside
    cmd += ["-c", str(path), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(jobs: int | None = None) -> pathlib.Path:
    OBJ.mkdir(exist_ok=True)
    srcs = sorted(p.name for p in CSRC.iterdir() if p.suffix in (".hip", ".cpp"))
    with cf.ThreadPoolExecutor(max_workers=jobs or min(8, os.cpu_count() or 4)) as ex:
        objs = list(ex.map(compile_one, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", *map(str, objs), "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--jobs", type=int, default=None)
    a = ap.parse_args()
    print(build(a.jobs))
    sys.exit(0)
