"""Write the golden fixtures: outputs of the reference's own v0 CPU solver.

Runs in the development container only (it needs oracle/_ref/, built by ``make -C oracle ref``
from the reference sources). Every fixture stores the inputs (case parameters, mask) and the
reference's post-streaming populations in its AoS layout (y*nx + x)*9 + i. Large outputs are
kept as a SHA-256 of the raw bytes plus a strided sample of values.

Cases (reference file:line they come from):
  walls_block   53x31, 500 steps, tau 0.8, F = (1e-5, 2e-6), walls + block x in [12,16),
                y in [10,17), init (1, 0.02, 0.01)      tests/test_v3_vs_v0.cu:20-37
  cylinder      120x48, D = 10 at (30, 25), u_in 0.05, Re 20, 1000 steps, open x
                                                        tests/cylinder_case.hpp:16-31
  seeded_block  53x31, 500 steps, tau 0.8, F = (1e-5, 2e-6), walls + an 8x8 block placed by
                seed (numpy default_rng(seed): x0 in [0, nx-8], y0 in [1, ny-9]), rest init
  crop256_f64   256x256 walled channel, tau 0.6, gx 1e-6, 200 steps, rest init, double
                (hash + sample only)
  cylinder_probe_1024x256  the force-driven periodic channel with a staircase cylinder
                (D 32 at (256, 129), Re 100 at U 0.05, Poiseuille init, 40,000 steps, float;
                body force 6.498966063887593e-07, the value the package's cylinder_force()
                computes on the GPU to hold the window-mean velocity at U):
                u_y at the wake probe (xc + 3D, yc) every 20 steps, as the reference's
                cylinder driver samples it (src/common/driver.hpp:175-214), and the final
                populations' SHA-256. Written by the oracle's OpenMP pull solver (the reference's
                v0 takes no initial field; the oracle is pinned bitwise to v0 by the fixtures
                above), since this case needs a non-uniform initial state.

Usage: python tests/golden/make_golden.py [cylinder_probe]
"""
from __future__ import annotations

import hashlib
import json
import pathlib
import sys

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402  (checker package)

SEEDS = (0, 1, 2)


def walls_block_mask(nx=53, ny=31):
    m = oracle.make_channel(nx, ny).reshape(ny, nx)
    m[10:17, 12:16] = 1
    return m.reshape(-1)


def seeded_block_position(seed, nx=53, ny=31, size=8):
    rng = np.random.default_rng(seed)
    x0 = int(rng.integers(0, nx - size + 1))
    y0 = int(rng.integers(1, ny - size))
    return x0, y0


def seeded_block_mask(seed, nx=53, ny=31, size=8):
    m = oracle.make_channel(nx, ny).reshape(ny, nx)
    x0, y0 = seeded_block_position(seed, nx, ny, size)
    m[y0:y0 + size, x0:x0 + size] = 1
    return m.reshape(-1)


def cylinder_params(precision):
    nx, ny, D, U, Re = 120, 48, 10, 0.05, 20.0
    u = float(np.float32(U)) if precision == "float" else U
    tau = 0.5 + 3.0 * u * D / Re
    if precision == "float":
        tau = float(np.float32(tau))
    return dict(nx=nx, ny=ny, tau=tau, fx=0.0, fy=0.0, u_in=U, xc=3.0 * D, yc=float(ny // 2 + 1),
                D=float(D), steps=1000, init=(1.0, U, 0.0))


def sample(a, k=257):
    idx = np.linspace(0, a.size - 1, k).astype(np.int64)
    return idx, a[idx]


CYL = dict(nx=1024, ny=256, D=32.0, xc=256.0, yc=129.0, Re=100.0, U=0.05, steps=40000,
           every=20, force=6.498966063887593e-07)


def cylinder_probe_case():
    """Inputs of the cylinder probe case: (tau, force, solid, initial AoS populations in float)."""
    c = CYL
    nx, ny, D = c["nx"], c["ny"], c["D"]
    nu = c["U"] * D / c["Re"]
    tau = 0.5 + 3.0 * nu
    H = ny - 2
    umax = 1.5 * c["U"]
    force = c["force"]
    solid = np.maximum(oracle.make_channel(nx, ny), oracle.make_cylinder(nx, ny, c["xc"], c["yc"], D))
    yy = np.arange(ny, dtype=np.float64) - 0.5
    prof = np.where((yy > 0) & (yy < H), 4 * umax * yy * (H - yy) / (H * H), 0.0)
    u = np.repeat(prof[:, None], nx, axis=1).reshape(-1, 1)
    cx = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1], dtype=np.float64)
    w = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
    cu = u * cx
    S = (w * ((1.0 - 1.0) + 1.0 * (3 * cu + 4.5 * cu * cu - 1.5 * (u * u)))).astype(np.float32)
    S[solid.astype(bool)] = 0
    return tau, force, solid, S.reshape(-1)


def cylinder_probe():
    c = CYL
    nx, ny = c["nx"], c["ny"]
    tau, force, solid, S0 = cylinder_probe_case()
    oracle.set_threads(8)
    o = oracle.Oracle(nx, ny, tau, force, 0.0, solid, "float", "pull")
    o.set_populations(S0)
    px, py = int(c["xc"] + 3 * c["D"]), int(c["yc"])
    cy = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1], dtype=np.float64)
    probe = []
    for _ in range(c["steps"] // c["every"]):
        o.run(c["every"])
        S = o.populations()
        f = S[(py * nx + px) * 9:(py * nx + px) * 9 + 9].astype(np.float64)
        probe.append(float(cy @ f) / (1.0 + float(f.sum())))
    np.savez_compressed(HERE / "cylinder_probe_1024x256_40000_float.npz", probe=np.array(probe),
                        final_sha256=hashlib.sha256(S.tobytes()).hexdigest(),
                        params=json.dumps(dict(c, tau=tau, force=force, probe=[px, py])))


def main():
    if sys.argv[1:] == ["cylinder_probe"]:
        cylinder_probe()
        return
    out = {}
    for prec in ("float", "double"):
        ref = oracle.Reference(prec)
        # walls_block
        m = walls_block_mask()
        S = ref.run("v0", 53, 31, 0.8, 1e-5, 2e-6, m, 500, init=(1.0, 0.02, 0.01))
        np.savez_compressed(HERE / f"walls_block_53x31_500_{prec}.npz", solid=m, S=S,
                            params=json.dumps(dict(nx=53, ny=31, tau=0.8, fx=1e-5, fy=2e-6,
                                                   steps=500, init=[1.0, 0.02, 0.01])))
        # cylinder (open x)
        p = cylinder_params(prec)
        m = ref_mask = np.empty(p["nx"] * p["ny"], np.uint8)
        ref.L.ref_make_cylinder(p["nx"], p["ny"], p["xc"], p["yc"], p["D"], ref_mask)
        S = ref.run("v0", p["nx"], p["ny"], p["tau"], 0.0, 0.0, m, p["steps"], init=p["init"],
                    open_x=True, u_in=p["u_in"])
        np.savez_compressed(HERE / f"cylinder_120x48_1000_{prec}.npz", solid=m, S=S,
                            params=json.dumps(p))
        # seeded 8x8 block, rest init
        for seed in SEEDS:
            m = seeded_block_mask(seed)
            S = ref.run("v0", 53, 31, 0.8, 1e-5, 2e-6, m, 500, init=(1.0, 0.0, 0.0))
            np.savez_compressed(HERE / f"seeded_block_53x31_500_seed{seed}_{prec}.npz", solid=m,
                                S=S, params=json.dumps(dict(nx=53, ny=31, tau=0.8, fx=1e-5,
                                                            fy=2e-6, steps=500, seed=seed,
                                                            block=seeded_block_position(seed),
                                                            init=[1.0, 0.0, 0.0])))
    # 256x256 double crop of the 8192^2 channel: hash + sample
    ref = oracle.Reference("double")
    nx = ny = 256
    m = oracle.make_channel(nx, ny)
    S = ref.run("v0", nx, ny, 0.6, 1e-6, 0.0, m, 200)
    idx, vals = sample(S)
    out["crop256_f64"] = dict(nx=nx, ny=ny, tau=0.6, fx=1e-6, fy=0.0, steps=200,
                              sha256=hashlib.sha256(S.tobytes()).hexdigest(),
                              sample_index=idx.tolist(), sample_value=[float(v) for v in vals],
                              total_mass=ref.L.ref_v0_total_mass(nx, ny, 0.6, 1e-6, 0.0, m, 1.0,
                                                                 0.0, 0.0, 200))
    (HERE / "hashes.json").write_text(json.dumps(out, indent=1))
    cylinder_probe()
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main()
