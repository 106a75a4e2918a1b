"""CPU: the data formats and the benchmark protocol around the step (the package's formats and
driver modules), against the reference's own writers and recorded data.

* .npy bytes: equal to the reference's io.cpp write_npy (oracle/_ref, built from the reference
  sources where they exist) for float32, float64 and uint8 fields of several shapes; loadable
  by numpy with the right dtype and shape; data on a 64-byte boundary.
* vorticity (the oracle's restatement, the GPU kernel's checker): bitwise equal to io.cpp's
  vorticity (oracle/_ref), periodic and one-sided; exact on linear shear fields.
* meta.json / probe.csv numbers: the text a default std::ostream prints (tests/cpp/bin/
  ostream_print) for the same doubles.
* run_bench protocol: steps / warm-up per size, the CSV header, mean / max / stdev columns and
  the copy-ceiling arithmetic, against the reference's results/*.csv rows and the ceilings its
  README quotes (tests/golden/bench_protocol.json).
* the driver's command line: params.cpp defaults, parsing and errors.
"""
from __future__ import annotations

import json
import pathlib
import statistics
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

PROTOCOL = json.loads((GOLDEN / "bench_protocol.json").read_text())
OSTREAM = ROOT / "tests" / "cpp" / "bin" / "ostream_print"


@pytest.fixture(scope="module")
def formats(lbm):
    from lbm_amd import formats as f

    return f


@pytest.fixture(scope="module")
def driver(lbm):
    from lbm_amd import driver as d

    return d


def _ref(precision):
    import oracle

    if not oracle.Reference.available(precision):
        pytest.skip("oracle/_ref not built (needs the reference checkout)")
    return oracle.Reference(precision)


# ---- .npy ---------------------------------------------------------------------------------------

SHAPES = [(1, 1), (3, 1), (34, 64), (7, 123), (1000, 1), (12345, 6789)]


@pytest.mark.parametrize("kind", ["float", "double", "uint8"])
def test_npy_bytes_equal_reference_writer(formats, tmp_path, kind):
    R = _ref("double" if kind == "double" else "float")
    rng = np.random.default_rng(3)
    for ny, nx in SHAPES:
        n = nx * ny if nx * ny < 10 ** 6 else 257  # large shapes: header only
        if kind == "uint8":
            a = rng.integers(0, 2, nx * ny if n == nx * ny else 0, dtype=np.uint8)
        else:
            a = rng.standard_normal(nx * ny if n == nx * ny else 0).astype(R.dtype)
        if a.size != nx * ny:
            # the header depends on the shape only: compare it on a zero-filled buffer's prefix
            full = np.zeros(nx * ny, dtype=a.dtype)
            R.write_npy(tmp_path / "ref.npy", full, ny, nx)
            ref = (tmp_path / "ref.npy").read_bytes()
            ours = formats.npy_bytes(full, ny, nx)
        else:
            R.write_npy(tmp_path / "ref.npy", a, ny, nx)
            ref = (tmp_path / "ref.npy").read_bytes()
            ours = formats.npy_bytes(a, ny, nx)
        assert ours == ref, (kind, ny, nx)


def test_npy_layout_loads_with_numpy(formats, tmp_path):
    for dt in (np.float32, np.float64, np.uint8):
        for ny, nx in SHAPES[:5]:
            a = (np.arange(nx * ny) % 251).astype(dt)
            formats.write_npy(tmp_path / "a.npy", a, ny, nx)
            raw = (tmp_path / "a.npy").read_bytes()
            assert raw[:8] == b"\x93NUMPY\x01\x00"
            hlen = int.from_bytes(raw[8:10], "little")
            assert (10 + hlen) % 64 == 0 and raw[10 + hlen - 1:10 + hlen] == b"\n"
            b = np.load(tmp_path / "a.npy", allow_pickle=False)
            assert b.dtype == dt and b.shape == (ny, nx) and np.array_equal(b.reshape(-1), a)
    with pytest.raises(ValueError):
        formats.npy_bytes(np.zeros(6, np.float32), 2, 4)
    with pytest.raises(ValueError):
        formats.npy_bytes(np.zeros(8, np.int32), 2, 4)


# ---- vorticity (oracle) -------------------------------------------------------------------------

@pytest.mark.parametrize("precision", ["float", "double"])
def test_oracle_vorticity_bitwise_equals_reference(oracle_mod, precision):
    R = _ref(precision)
    rng = np.random.default_rng(11)
    for nx, ny in [(1, 3), (2, 3), (3, 4), (64, 34), (37, 19)]:
        u = (rng.standard_normal(nx * ny) * 0.05).astype(R.dtype)
        v = (rng.standard_normal(nx * ny) * 0.05).astype(R.dtype)
        solid = (rng.random(nx * ny) < 0.15).astype(np.uint8)
        for px in (True, False):
            a = R.vorticity(u, v, solid, nx, ny, px)
            b = oracle_mod.vorticity(u, v, solid, nx, ny, px)
            assert b.dtype == R.dtype
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), (nx, ny, px)


def test_oracle_vorticity_linear_shear_exact(oracle_mod):
    # ux = a y, uy = b x: dUy/dx - dUx/dy = b - a away from the wrap; the one-sided columns of an
    # open domain give b too. Values are small integers times powers of two: exact in float.
    nx, ny, a, b = 16, 12, 0.25, 0.5
    y, x = np.mgrid[0:ny, 0:nx]
    ux = (a * y).astype(np.float32).reshape(-1)
    uy = (b * x).astype(np.float32).reshape(-1)
    solid = np.zeros(nx * ny, np.uint8)
    solid[5 * nx + 7] = 1
    for px in (False, True):
        w = oracle_mod.vorticity(ux, uy, solid, nx, ny, px).reshape(ny, nx)
        inner = w[1:-1, 1:-1].copy()
        inner[4, 6] = b - a  # the solid cell (5, 7) reads 0
        assert np.all(inner == np.float32(b - a))
        assert w[5, 7] == 0
        if not px:
            assert np.all(w[1:-1, 0] == np.float32(b - a))
            assert np.all(w[1:-1, -1] == np.float32(b - a))
        else:  # periodic wrap: (uy(1) - uy(nx - 1)) / 2 at x = 0
            assert np.all(w[1:-1, 0] == np.float32((b * 1 - b * (nx - 1)) / 2 - a))


# ---- meta.json / probe.csv numbers --------------------------------------------------------------

def test_cformat_matches_default_ostream(formats):
    if not OSTREAM.exists():
        pytest.skip("tests/cpp not built")
    vals = [0.9, 0.56, 1e-6, 100.0, 513.0, 1.23456789e-5, -0.00012345678, 1e20, 0.0,
            float(np.float32(0.9)), float(np.float32(0.05)), 0.5 + 3 * 0.05 * 40 / 100,
            2.0 / 3.0 * 1e-7, 123456.7, 1234567.0, -0.0, 5e-324]
    rng = np.random.default_rng(5)
    vals += [float(v) for v in rng.standard_normal(40) * 10.0 ** rng.integers(-9, 9, 40)]
    out = subprocess.run([str(OSTREAM), *[repr(v) for v in vals]], capture_output=True,
                         text=True, check=True).stdout.split("\n")
    for v, want in zip(vals, out):
        assert formats.cformat(v) == want, v
    assert formats.cformat(40) == "40" and formats.cformat("cylinder") == '"cylinder"'


def test_meta_line_reads_back(formats):
    m = formats.channel_meta(64, 34, np.float32(0.9), np.float32(0.05), 1.23e-6, 20000, "float")
    t = formats.meta_text(m)
    assert t.endswith("}\n")
    d = json.loads(t)
    assert list(d) == ["version", "case", "nx", "ny", "tau", "umax", "force_x", "steps", "real"]
    assert d["tau"] == 0.9 and d["force_x"] == 1.23e-6 and d["nx"] == 64
    c = formats.cylinder_meta(400, 100, 0.56, 0.05, 100.0, 40, 200.0, 51.0, 320, 51, 50, 0,
                              1000, "double")
    assert list(json.loads(formats.meta_text(c))) == [
        "version", "case", "nx", "ny", "tau", "u_in", "re", "diameter", "xc", "yc", "probe_x",
        "probe_y", "probe_every", "frame_every", "steps", "real"]


def test_probe_moments_recover_equilibrium_velocity(formats, oracle_mod):
    # f - w of the equilibrium at (rho, u) has zeroth moment rho - 1 and first moments rho u
    w = [4 / 9] + [1 / 9] * 4 + [1 / 36] * 4
    cx = [0, 1, 0, -1, 0, 1, -1, -1, 1]
    cy = [0, 0, 1, 0, -1, 1, 1, -1, -1]
    for rho, ux, uy in [(1.0, 0.05, 0.0), (1.02, -0.03, 0.011), (0.97, 0.0, -0.04)]:
        f = []
        for i in range(9):
            cu = cx[i] * ux + cy[i] * uy
            f.append(w[i] * (rho - 1 + rho * (3 * cu + 4.5 * cu * cu - 1.5 * (ux * ux + uy * uy))))
        for dt, tol in ((np.float64, 1e-14), (np.float32, 2e-6)):
            a, b, r = formats.probe_moments(np.array(f, dt), dt)
            assert abs(a - ux) < tol and abs(b - uy) < tol and abs(r - rho) < tol
    assert formats.probe_line(50, 0.05, -1e-7, 1.0) == "50,0.05,-1e-07,1\n"
    assert formats.frame_name(1200) == "frames/vort_0001200.npy"


# ---- benchmark protocol -------------------------------------------------------------------------

def test_protocol_steps_match_recorded_runs(formats):
    for r in PROTOCOL["run_bench_rows_cpu"]:
        assert formats.protocol_steps(int(r["nx"]), int(r["ny"])) == (int(r["steps"]),
                                                                      int(r["warmup"]))
    for r in PROTOCOL["run_bench_rows_gpu"]:  # recorded with --updates 5e8
        assert formats.protocol_steps(int(r["nx"]), int(r["ny"]), 5e8) == (int(r["steps"]),
                                                                           int(r["warmup"]))


def test_bench_columns_match_recorded_csv(formats):
    assert formats.BENCH_COLUMNS == PROTOCOL["run_bench_header_gpu"]
    assert formats.BENCH_COLUMNS[:22] == PROTOCOL["run_bench_header_cpu"]
    assert formats.BANDWIDTH_COLUMNS == PROTOCOL["gpu_bandwidth_header"]
    assert formats.STROUHAL_COLUMNS == PROTOCOL["strouhal_header"]


def test_recorded_mlups_follow_the_bench_formula(formats):
    # mlups = nx ny steps / time / 1e6 per repetition (driver.hpp:120), so the row's mean of
    # the per-rep values is at least the value at the mean time (1/t is convex) and at most the
    # max; the GPU rows (small spread) also agree with it to 0.1 %.
    for r in PROTOCOL["run_bench_rows_cpu"] + PROTOCOL["run_bench_rows_gpu"]:
        nx, ny, steps = int(r["nx"]), int(r["ny"]), int(r["steps"])
        at_mean = nx * ny * steps / float(r["time_mean_s"]) / 1e6
        mean, mx = float(r["mlups_mean"]), float(r["mlups_max"])
        assert at_mean <= mean * (1 + 1e-5) and mean <= mx
    for r in PROTOCOL["run_bench_rows_gpu"]:
        at_mean = int(r["nx"]) * int(r["ny"]) * int(r["steps"]) / float(r["time_mean_s"]) / 1e6
        assert abs(at_mean - float(r["mlups_mean"])) / at_mean < 1e-3


def test_bench_record_and_row(formats, tmp_path):
    times = [0.156480, 0.156302, 0.15702, 0.1561, 0.1569]
    info = '"gpu": "AMD Instinct MI355X", "arch": "gfx950", "hip_driver": 70226015, "block": "512x1", "variant": "soa_pull_buffer"'
    lines = "\n".join(formats.bench_record("v5_hip", r, 2048, 2048, 119, 23, 0, t, "float",
                                           "hipcc (HIP 7.2.0)", "-O3", "Release", info)
                      for r, t in enumerate(times))
    recs = formats.parse_bench_lines("noise\n" + lines + "\n")
    assert len(recs) == 5
    assert list(recs[0])[:13] == ["version", "rep", "nx", "ny", "steps", "warmup", "threads",
                                  "time_s", "mlups", "real", "compiler", "flags", "build_type"]
    for rec, t in zip(recs, times):
        assert rec["mlups"] == float(f"{2048 * 2048 * 119 / t / 1e6:.4f}")
    row = formats.bench_row(recs, common={"date": "2026-01-01T00:00:00"})
    assert list(row) == formats.BENCH_COLUMNS or set(row) == set(formats.BENCH_COLUMNS)
    assert row["time_mean_s"] == f"{statistics.mean(times):.6f}"
    assert row["time_std_s"] == f"{statistics.stdev(times):.6f}"
    ml = [r["mlups"] for r in recs]
    assert row["mlups_max"] == f"{max(ml):.3f}" and row["mlups_std"] == f"{statistics.stdev(ml):.3f}"
    assert row["gpu"] == "AMD Instinct MI355X" and row["gpu_driver"] == "70226015"
    assert row["block"] == "512x1" and row["variant"] == "soa_pull_buffer"
    p = tmp_path / "v5_hip.csv"
    formats.write_csv([row], p, formats.BENCH_COLUMNS)
    formats.write_csv([row], p, formats.BENCH_COLUMNS, append=True)
    import csv

    with open(p, newline="") as f:
        rd = csv.DictReader(f)
        assert rd.fieldnames == PROTOCOL["run_bench_header_gpu"]
        got = list(rd)
    assert len(got) == 2 and got[1]["mlups_mean"] == row["mlups_mean"]


def test_ceiling_matches_readme(formats):
    gpu, _ = formats.ceiling_mlups(PROTOCOL["gpu_bandwidth_rows"], 4)
    assert round(gpu) == PROTOCOL["readme_ceiling_mlups"]["gpu_float"]
    cpu, _ = formats.ceiling_mlups(PROTOCOL["stream_rows"], 4)
    assert round(cpu) == PROTOCOL["readme_ceiling_mlups"]["cpu_float"]
    # hardware ceiling: GPU 2 * 9 * size bytes per update, CPU 3 * 9 * size (write-allocate)
    _, hw = formats.ceiling_mlups(PROTOCOL["gpu_bandwidth_rows"], 4, peak_gbs=8000.0)
    assert abs(hw - 8000e9 / 72 / 1e6) < 1e-6
    _, hwc = formats.ceiling_mlups(PROTOCOL["stream_rows"], 8, peak_gbs=42.7, gpu=False)
    assert abs(hwc - 42.7e9 / 216 / 1e6) < 1e-9
    row = formats.bandwidth_row("AMD Instinct MI355X", 6.1e6, 5.9e6, 1 << 32, 20, 1)
    assert list(row) == PROTOCOL["gpu_bandwidth_header"]
    assert round(formats.ceiling_mlups([row], 4)[0]) == round(6.1e12 / 72 / 1e6)


def test_strouhal_row_columns(formats):
    r = formats.strouhal_row("v5_hip_mi355x", "v5_hip", "float", 1000, 320, 40, 0.05, 100,
                             0.56, 160000, 88000, 0.18071, 0.18049, 15, 0.032321)
    assert list(r) == PROTOCOL["strouhal_header"]
    assert (r["tau"], r["st_fft"], r["st_zero_crossing"], r["uy_amplitude"]) == (
        "0.56000", "0.1807", "0.1805", "0.03232")


# ---- the driver's command line ------------------------------------------------------------------

def test_driver_defaults_and_parsing(driver):
    p = driver.parse_args([])
    assert (p.case_name, p.nx, p.ny, p.tau, p.u_max, p.steps, p.out_dir) == (
        "poiseuille", 64, 34, 0.9, 0.05, 20000, "output")
    assert (p.re, p.diameter, p.probe_every, p.frame_every, p.frame_start) == (100.0, 40, 50, 0, 0)
    assert (p.bench, p.warmup, p.reps) == (False, 100, 5)
    p = driver.parse_args(["--bench", "--nx", "2048", "--ny", "2048", "--steps", "119",
                           "--warmup", "23", "--reps", "3", "--real", "double"])
    assert p.bench and (p.nx, p.ny, p.steps, p.warmup, p.reps, p.real) == (2048, 2048, 119, 23,
                                                                           3, "double")
    p = driver.parse_args(["--case", "cylinder", "--re", "150", "--diameter", "20",
                           "--probe-every", "10", "--frame-every", "100", "--frame-start", "500",
                           "--bench"])
    assert (p.case_name, p.re, p.diameter, p.probe_every, p.frame_every, p.frame_start,
            p.bench) == ("cylinder", 150.0, 20, 10, 100, 500, True)


@pytest.mark.parametrize("argv,msg", [
    (["--nx"], "missing value for --nx"),
    (["--foo", "1"], "unknown option --foo"),
    (["--nx", "0"], "grid too small"),
    (["--ny", "2"], "grid too small"),
    (["--tau", "0.5"], "tau must be > 0.5"),
    (["--steps", "-1"], "steps must be >= 0"),
    (["--reps", "0"], "need warmup >= 0 and reps >= 1"),
    (["--warmup", "-1"], "need warmup >= 0 and reps >= 1"),
    (["--diameter", "1"], "invalid cylinder parameters"),
    (["--re", "0"], "invalid cylinder parameters"),
    (["--probe-every", "0"], "invalid cylinder parameters"),
    (["--nx", "abc"], "invalid value"),
])
def test_driver_rejects_what_the_reference_rejects(driver, argv, msg, capsys):
    with pytest.raises(ValueError, match=msg):
        driver.parse_args(argv)
    assert driver.run_driver(argv) == 1
    assert capsys.readouterr().err.startswith("error: ")


def test_driver_cylinder_setup(driver):
    p = driver.parse_args(["--case", "cylinder", "--nx", "1000", "--ny", "320"])
    cs = driver.cylinder_setup(p, np.float32)
    assert (cs["xc"], cs["yc"], cs["probe_x"], cs["probe_y"]) == (200.0, 161.0, 320, 161)
    # results/strouhal.csv records tau 0.56000 for this case (Re 100, D 40, u_in 0.05)
    assert f"{cs['tau']:.5f}" == "0.56000"
    assert cs["tau"] == float(np.float32(0.5 + 3.0 * float(np.float32(0.05)) * 40 / 100))


def test_strouhal_analysis_of_a_written_probe(formats, lbm, tmp_path):
    # a probe.csv / meta.json pair as the cylinder driver writes them, u_y a sine of known
    # period plus a decaying transient: the analysis finds St = D / (U T), agrees with the
    # package's own estimator, and counts the periods in the window
    D, U, T, every, steps = 40, 0.05, 4400.0, 50, 160000
    n = np.arange(every, steps + 1, every)
    uy = 0.03 * np.sin(2 * np.pi * n / T) + 0.01 * np.exp(-n / 5000.0)
    with open(tmp_path / "probe.csv", "w") as f:
        f.write("step,ux,uy,rho\n")
        for k, v in zip(n, uy):
            f.write(formats.probe_line(int(k), 0.05, float(v), 1.0))
    meta = formats.cylinder_meta(1000, 320, 0.56, U, 100.0, D, 200.0, 161.0, 320, 161, every,
                                 0, steps, "float")
    (tmp_path / "meta.json").write_text(formats.meta_text(meta))
    a = formats.strouhal_analysis(tmp_path, start=0.55)
    assert a["window_start"] == 88000
    assert abs(a["st_fft"] - D / (U * T)) < 2e-4 and abs(a["st_zero"] - D / (U * T)) < 2e-4
    keep = n >= 88000
    st_fft, st_zero, periods = lbm.strouhal(uy[keep], float(every), D, U)
    assert abs(st_fft - a["st_fft"]) < 1e-4 and periods == a["periods"]
    assert abs(a["uy_amplitude"] - 0.03) < 1e-3


def test_bandwidth_report_matches_ceiling_py_on_recorded_rows(lbm, capsys):
    from lbm_amd import bandwidth

    assert bandwidth.gpu_name("AMD Instinct MI355X (gfx950:sramecc+:xnack-), 256 CUs") == \
        "AMD Instinct MI355X"
    assert bandwidth.gpu_name(" (gfx950:sramecc+:xnack-), 256 CUs, 309.2 GB") == "AMD gfx950"
    bandwidth.report(PROTOCOL["gpu_bandwidth_rows"], peak_gbs=320.0)
    out = capsys.readouterr().out.splitlines()
    # README.md:45-50: 231.9 GB/s measured, 3,221 MLUPS float ceiling
    assert out[0] == "Measured best copy: 231.94 GB/s (Tesla T4)"
    assert out[1].startswith("  float : measured ceiling =  3221.4 MLUPS")
    assert out[2].startswith("  double: measured ceiling =  1610.7 MLUPS")
    assert "hardware ceiling =  4444.4 MLUPS" in out[1]
