"""GPU: the vorticity kernel and the driver's file outputs against the oracle.

* Solver.vorticity (device, io.cpp:54-74) is bit-identical to the oracle's vorticity of the
  solver's own velocity, periodic and one-sided, float / double / FP16 storage (float math),
  after an open-x cylinder run (a field with vortices and solid cells).
* driver.py, the reference's command line (driver.hpp:255-300), on small cases:
  - poiseuille: ux.npy / uy.npy / rho.npy equal the oracle's v0 macroscopic fields bit for bit,
    vorticity.npy the oracle's periodic vorticity of them, solid.npy the channel mask, and
    meta.json holds the reference's keys and values;
  - --bench: `reps` BENCH lines that run_bench.py's parser reads, with the reference's keys
    and consistent MLUPS, folded into a run_bench CSV row;
  - cylinder: probe.csv equals the oracle's probe (same sums, same text) line by line, the
    vorticity frames and final fields equal the oracle's;
  - the reference's recorded cylinder case (results/strouhal.csv): St, periods and u_y
    amplitude reproduced to the recorded digits.
"""
from __future__ import annotations

import json
import pathlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint32 if a.dtype == np.float32 else np.uint64)


def assert_bitwise(a, b, what):
    assert a.dtype == b.dtype and a.shape == b.shape, what
    bad = np.flatnonzero(bits(a) != bits(b))
    assert bad.size == 0, f"{what}: {bad.size} differing values, first at {bad[0]}"


@pytest.mark.parametrize("prec", ["float", "double", "fp16"])
def test_vorticity_bitwise_equals_oracle(lbm, oracle_mod, prec):
    nx, ny, D = 150, 50, 10
    solid = lbm.make_cylinder(nx, ny, 5 * D, ny // 2 + 1, D)
    u = 0.05
    tau = 0.5 + 3.0 * u * D / 40.0
    print(f"vorticity {prec} {nx}x{ny}", flush=True)
    with lbm.Solver(nx, ny, tau, 0.0, 0.0, solid, prec) as s:
        s.set_inflow_outflow(u)
        s.init_equilibrium(1.0, u, 0.0)
        s.run(400)
        _, ux, uy = s.macroscopic()
        assert np.abs(uy).max() > 1e-3  # a field with structure
        for px, periodic in ((1, True), (0, False), (-1, False)):
            w = s.vorticity(px)
            ref = oracle_mod.vorticity(ux, uy, solid, nx, ny, periodic)
            assert_bitwise(w, ref, f"vorticity periodic_x={px}")
            assert np.all(w[solid != 0] == 0)
    # periodic domain: -1 is periodic
    with lbm.Solver(nx, ny, 0.8, 1e-5, 2e-6, lbm.make_channel(nx, ny), prec) as s:
        s.init_equilibrium(1.0, 0.01, 0.0)
        s.run(50)
        _, ux, uy = s.macroscopic()
        assert_bitwise(s.vorticity(-1), oracle_mod.vorticity(ux, uy, s.solid(), nx, ny, True),
                       "vorticity of a periodic channel")


def test_vorticity_degenerate_grids(lbm, oracle_mod):
    for nx, ny in ((1, 3), (2, 5), (65, 3)):
        solid = np.zeros(nx * ny, np.uint8)
        print(f"vorticity {nx}x{ny}", flush=True)
        with lbm.Solver(nx, ny, 0.7, 1e-5, 3e-6, solid, "float") as s:
            rng = np.random.default_rng(nx * 7 + ny)
            s.set_populations((rng.standard_normal(9 * nx * ny) * 1e-2).astype(np.float32))
            s.run(3)
            _, ux, uy = s.macroscopic()
            for px in (0, 1):
                assert_bitwise(s.vorticity(px),
                               oracle_mod.vorticity(ux, uy, solid, nx, ny, bool(px)),
                               f"vorticity {nx}x{ny} periodic_x={px}")


@pytest.fixture(scope="module")
def driver(lbm):
    from lbm_amd import driver as d

    return d


@pytest.mark.parametrize("prec", ["float", "double"])
def test_driver_poiseuille_fields_equal_oracle(lbm, oracle_mod, driver, tmp_path, prec, capsys):
    nx, ny, steps = 64, 34, 300
    out = tmp_path / "out"
    print(f"driver poiseuille {prec}", flush=True)
    assert driver.run_driver(["--nx", str(nx), "--ny", str(ny), "--steps", str(steps),
                              "--tau", "0.8", "--umax", "0.04", "--out", str(out),
                              "--real", prec]) == 0
    text = capsys.readouterr().out
    assert "fields written to" in text and "mass drift=" in text
    force = lbm.poiseuille_force(ny - 2, 0.8, 0.04, prec)
    solid = oracle_mod.make_channel(nx, ny)
    o = oracle_mod.Oracle(nx, ny, 0.8, force, 0.0, solid, prec)
    o.init_equilibrium(1.0, 0.0, 0.0)
    o.run(steps)
    rho, ux, uy = o.macroscopic()
    for name, ref in (("rho", rho), ("ux", ux), ("uy", uy)):
        a = np.load(out / f"{name}.npy", allow_pickle=False)
        assert a.shape == (ny, nx)
        assert_bitwise(a.reshape(-1), ref, name)
    w = np.load(out / "vorticity.npy", allow_pickle=False).reshape(-1)
    assert_bitwise(w, oracle_mod.vorticity(ux, uy, solid, nx, ny, True), "vorticity.npy")
    assert np.array_equal(np.load(out / "solid.npy").reshape(-1), solid)
    meta = json.loads((out / "meta.json").read_text())
    assert list(meta) == ["version", "case", "nx", "ny", "tau", "umax", "force_x", "steps", "real"]
    assert (meta["case"], meta["nx"], meta["ny"], meta["steps"], meta["real"]) == (
        "poiseuille", nx, ny, steps, prec)
    assert meta["tau"] == 0.8 and meta["umax"] == 0.04
    assert abs(meta["force_x"] - force) <= 1e-5 * force  # %g keeps 6 digits


def test_driver_bench_lines(lbm, driver, capsys):
    from lbm_amd import formats

    print("driver --bench", flush=True)
    capsys.readouterr()
    assert driver.run_driver(["--bench", "--nx", "512", "--ny", "256", "--steps", "40",
                              "--warmup", "8", "--reps", "3"]) == 0
    recs = formats.parse_bench_lines(capsys.readouterr().out)
    assert len(recs) == 3
    for r, rec in enumerate(recs):
        assert list(rec)[:13] == ["version", "rep", "nx", "ny", "steps", "warmup", "threads",
                                  "time_s", "mlups", "real", "compiler", "flags", "build_type"]
        assert (rec["rep"], rec["nx"], rec["ny"], rec["steps"], rec["warmup"]) == (r, 512, 256,
                                                                                 40, 8)
        assert rec["time_s"] > 0 and rec["gpu"] and rec["arch"].startswith("gfx")
        assert abs(rec["mlups"] - 512 * 256 * 40 / rec["time_s"] / 1e6) <= 1e-3 * rec["mlups"] + 1e-3
    row = formats.bench_row(recs)
    assert set(row) == set(formats.BENCH_COLUMNS) and row["reps"] == 3


def test_driver_cylinder_probe_and_frames_equal_oracle(lbm, oracle_mod, driver, tmp_path,
                                                       capsys):
    from lbm_amd import formats

    nx, ny, D, steps, every, frame = 160, 48, 8, 200, 50, 100
    out = tmp_path / "cyl"
    print("driver cylinder", flush=True)
    assert driver.run_driver(["--case", "cylinder", "--nx", str(nx), "--ny", str(ny),
                              "--diameter", str(D), "--re", "20", "--steps", str(steps),
                              "--probe-every", str(every), "--frame-every", str(frame),
                              "--out", str(out)]) == 0
    assert "probe, frames and fields written" in capsys.readouterr().out
    p = driver.parse_args(["--case", "cylinder", "--nx", str(nx), "--ny", str(ny),
                           "--diameter", str(D), "--re", "20"])
    cs = driver.cylinder_setup(p, np.float32)
    solid = oracle_mod.make_cylinder(nx, ny, cs["xc"], cs["yc"], D)
    assert np.array_equal(np.load(out / "solid.npy").reshape(-1), solid)
    o = oracle_mod.Oracle(nx, ny, cs["tau"], 0.0, 0.0, solid, "float")
    o.set_inflow_outflow(0.05)
    o.init_equilibrium(1.0, 0.05, 0.0)
    lines = ["step,ux,uy,rho\n"]
    n = 0
    while n < steps:
        o.run(every)
        n += every
        f = o.populations().reshape(ny, nx, 9)[cs["probe_y"], cs["probe_x"]]
        lines.append(formats.probe_line(n, *formats.probe_moments(f, np.float32)))
        if n % frame == 0:
            _, ux, uy = o.macroscopic()
            got = np.load(out / formats.frame_name(n), allow_pickle=False).reshape(-1)
            assert_bitwise(got, oracle_mod.vorticity(ux, uy, solid, nx, ny, False),
                           f"frame {n}")
    assert (out / "probe.csv").read_text() == "".join(lines)
    rho, ux, uy = o.macroscopic()
    assert_bitwise(np.load(out / "ux.npy").reshape(-1), ux, "final ux")
    assert_bitwise(np.load(out / "vorticity.npy").reshape(-1),
                   oracle_mod.vorticity(ux, uy, solid, nx, ny, False), "final vorticity")
    meta = json.loads((out / "meta.json").read_text())
    assert (meta["case"], meta["diameter"], meta["xc"], meta["yc"], meta["probe_x"],
            meta["probe_y"], meta["probe_every"], meta["frame_every"], meta["steps"]) == (
        "cylinder", D, 5.0 * D, ny // 2 + 1, cs["probe_x"], cs["probe_y"], every, frame, steps)


def test_reference_cylinder_run_reproduces_recorded_strouhal(lbm, driver, tmp_path, capsys):
    """The reference's recorded cylinder runs (results/strouhal.csv: v1, v3 and v4, float,
    1000 x 320, D 40, Re 100, u_in 0.05, 160,000 steps, probe every 50 steps, St over the last
    45 %) through the driver and scripts/strouhal.py's analysis: the same St (FFT and zero
    crossings), periods and u_y amplitude to the digits the reference records."""
    from lbm_amd import formats

    protocol = json.loads((pathlib.Path(__file__).parent / "golden" /
                           "bench_protocol.json").read_text())
    rec = [r for r in protocol["strouhal_rows"] if r["version"] in ("v1", "v3", "v4")]
    assert len(rec) == 3 and len({(r["st_fft"], r["st_zero_crossing"], r["periods"],
                                   r["uy_amplitude"]) for r in rec}) == 1
    r = rec[0]
    out = tmp_path / "cylinder"
    print("driver cylinder 1000x320, 160000 steps", flush=True)
    assert driver.run_driver(["--case", "cylinder", "--nx", r["nx"], "--ny", r["ny"],
                              "--diameter", r["diameter"], "--re", r["re"], "--umax", r["u_in"],
                              "--steps", r["steps"], "--out", str(out)]) == 0
    capsys.readouterr()
    a = formats.strouhal_analysis(out, start=int(r["window_start"]) / int(r["steps"]))
    row = formats.strouhal_row("v5_hip", "v5_hip", a["meta"]["real"], a["meta"]["nx"],
                               a["meta"]["ny"], a["meta"]["diameter"], a["meta"]["u_in"],
                               a["meta"]["re"], a["meta"]["tau"], a["meta"]["steps"],
                               a["window_start"], a["st_fft"], a["st_zero"], a["periods"],
                               a["uy_amplitude"])
    print({k: row[k] for k in ("tau", "window_start", "st_fft", "st_zero_crossing", "periods",
                               "uy_amplitude")})
    for k in ("real", "tau", "window_start", "st_fft", "st_zero_crossing", "periods",
              "uy_amplitude"):
        assert str(row[k]) == r[k], k


def test_bandwidth_tool_writes_gpu_bandwidth_csv(lbm, tmp_path, capsys):
    import csv

    from lbm_amd import bandwidth, formats

    print("bandwidth tool", flush=True)
    path = tmp_path / "gpu_bandwidth.csv"
    for _ in range(2):  # the second run appends a row without a header
        assert bandwidth.main([str(path), "--bytes", str(1 << 28), "--reps", "5",
                               "--peak-gbs", "8000"]) == 0
    text = capsys.readouterr().out
    assert "Measured best copy:" in text and "float : measured ceiling" in text
    with open(path, newline="") as f:
        rd = csv.DictReader(f)
        assert rd.fieldnames == formats.BANDWIDTH_COLUMNS
        rows = list(rd)
    assert len(rows) == 2
    for r in rows:
        assert float(r["copy_MBps"]) > 1e6 and float(r["memcpy_MBps"]) > 1e5  # > 1 TB/s
        assert int(r["bytes_per_array"]) == 1 << 28 and int(r["reps"]) == 5
    measured, hw = formats.ceiling_mlups(rows, 4, 8000.0)
    assert 1e4 < measured < hw


def test_driver_fp16_storage_fields_close_to_fp64(lbm, oracle_mod, driver, tmp_path, capsys):
    """--real fp16: the files are the float files of the reference's float build (float32 .npy,
    meta real "float"), and the velocity is within the FP16 form's 1e-3 relative L2 of the
    oracle's double run."""
    nx, ny, steps = 64, 34, 3000
    out = tmp_path / "fp16"
    print("driver poiseuille fp16", flush=True)
    assert driver.run_driver(["--nx", str(nx), "--ny", str(ny), "--steps", str(steps),
                              "--tau", "1.0", "--out", str(out), "--real", "fp16"]) == 0
    capsys.readouterr()
    meta = json.loads((out / "meta.json").read_text())
    assert meta["real"] == "float"
    ux = np.load(out / "ux.npy", allow_pickle=False)
    assert ux.dtype == np.float32 and ux.shape == (ny, nx)
    force = lbm.poiseuille_force(ny - 2, 1.0, 0.05, "double")
    o = oracle_mod.Oracle(nx, ny, 1.0, force, 0.0, oracle_mod.make_channel(nx, ny), "double")
    o.init_equilibrium(1.0, 0.0, 0.0)
    o.run(steps)
    ref = o.macroscopic()[1].reshape(ny, nx)
    err = np.sqrt(np.sum((ux.astype(np.float64) - ref) ** 2) / np.sum(ref ** 2))
    print(f"fp16 driver u_x relative L2 vs fp64 oracle: {err:.2e}")
    assert err < 1e-3
    w = np.load(out / "vorticity.npy", allow_pickle=False)
    assert w.dtype == np.float32 and np.all(w[0] == 0) and np.all(w[-1] == 0)
