"""The counter tool (package counters.py, the counterpart of the reference's bench/ncu_profile.py):
parsing of rocprofv3 counter-collection CSV files, calibration against the copy kernel of known
size, bytes per launch and per cell update, and the summary bench.py reads back."""
from __future__ import annotations

import csv
import importlib

import pytest

HEAD = ["Correlation_Id", "Dispatch_Id", "Agent_Id", "Queue_Id", "Process_Id", "Thread_Id",
        "Grid_Size", "Kernel_Id", "Kernel_Name", "Workgroup_Size", "LDS_Block_Size",
        "Scratch_Size", "VGPR_Count", "Accum_VGPR_Count", "SGPR_Count", "Counter_Name",
        "Counter_Value", "Start_Timestamp", "End_Timestamp"]
STEP = "void lbmhip::(anonymous namespace)::step2_kernel<float, lbmhip::PlanesOne, 2, false>(x)"
COPY = "void lbmhip::(anonymous namespace)::copy_kernel<4, true>(float const*, float*, unsigned long)"


def write_csv(path, rows):
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(HEAD)
        for d, (k, c, v) in enumerate(rows):
            r = ["0"] * len(HEAD)
            r[1], r[8], r[15], r[16] = str(d + 1), k, c, repr(v)
            w.writerow(r)


@pytest.fixture()
def counters(lbm):
    return importlib.import_module(lbm.__name__ + ".counters")


def test_calibrated_traffic(tmp_path, counters):
    cal = 2_415_919_104
    # the copy kernel reads cal bytes and FETCH_SIZE shows half (x 2 calibration), WRITE_SIZE
    # shows them all (x 1); the step kernel: 1.1 x 36 B read and 36 B written per cell
    cells = 1_000_000
    fetch = [(COPY, "FETCH_SIZE", cal / 1024 / 2)] * 3 + \
            [(STEP, "FETCH_SIZE", 1.1 * 36 * cells / 1024 / 2)] * 4
    write = [(COPY, "WRITE_SIZE", cal / 1024)] * 3 + [(STEP, "WRITE_SIZE", 36 * cells / 1024)] * 4
    write_csv(tmp_path / "f.csv", fetch)
    write_csv(tmp_path / "w.csv", write)
    s = counters.summarize(counters.read_counters(tmp_path / "f.csv"),
                           counters.read_counters(tmp_path / "w.csv"),
                           "step2_kernel<float, lbmhip::PlanesOne", cells, 2, 72)
    assert s["calibration"]["fetch"] == pytest.approx(2.0) and s["calibration"]["write"] == pytest.approx(1.0)
    assert s["read_bytes_per_launch"] == pytest.approx(1.1 * 36 * cells, rel=1e-9)
    assert s["write_bytes_per_launch"] == pytest.approx(36 * cells, rel=1e-9)
    assert s["traffic_vs_algorithmic"] == pytest.approx(1.05)
    assert s["read_vs_algorithmic"] == pytest.approx(1.1)
    assert s["bytes_per_cell_update"] == pytest.approx(36 * 2.1 / 2, abs=0.01)
    assert s["dispatches"] == {"fetch": 4, "write": 4}
    out = tmp_path / "summary.md"
    counters.write_summary(out, s, ["rocprofv3 --pmc FETCH_SIZE ..."])
    back = counters.read_summary(out)
    assert back["traffic_bytes_per_launch"] == s["traffic_bytes_per_launch"]
    assert back["source_tag"] == counters.source_tag()


def test_counter_values_of_one_dispatch_are_summed_and_skip(tmp_path, counters):
    rows = [(STEP, "FETCH_SIZE", 10.0), (STEP, "FETCH_SIZE", 20.0), (STEP, "FETCH_SIZE", 30.0)]
    write_csv(tmp_path / "f.csv", rows)
    r = counters.read_counters(tmp_path / "f.csv")
    assert [e["value"] for e in r] == [10.0, 20.0, 30.0]
    assert counters.per_dispatch(r, "step2_kernel", "FETCH_SIZE", skip=1) == [20.0, 30.0]
    # a counter split over rows of the same dispatch (per XCD, say) is summed
    with open(tmp_path / "g.csv", "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(HEAD)
        for v in (1.5, 2.5):
            r2 = ["0"] * len(HEAD)
            r2[1], r2[8], r2[15], r2[16] = "7", STEP, "FETCH_SIZE", str(v)
            w.writerow(r2)
    assert counters.read_counters(tmp_path / "g.csv")[0]["value"] == 4.0
    with pytest.raises(ValueError, match="no dispatch"):
        counters.summarize(r, r, "nothing_like_this", 10, 1, 72)


def test_launch_of_several_dispatches(tmp_path, counters):
    """A launch made of several dispatches (the batch: one per tile-width class): the last whole
    launches are selected and the per-dispatch mean scaled back to a launch."""
    cal = 2_415_919_104
    cells = 3_000
    # two tuning dispatches to leave out, then 4 launches of 3 dispatches (12 B, 24 B, 36 B
    # per cell of the launch's read bytes: 72 B over the three, x 2 in FETCH)
    parts = [12 * cells, 24 * cells, 36 * cells]
    fetch = [(COPY, "FETCH_SIZE", cal / 1024 / 2)] + [(STEP, "FETCH_SIZE", 1e9)] * 2 + \
            [(STEP, "FETCH_SIZE", b / 1024 / 2) for _ in range(4) for b in parts]
    write = [(COPY, "WRITE_SIZE", cal / 1024)] + [(STEP, "WRITE_SIZE", 1e9)] * 2 + \
            [(STEP, "WRITE_SIZE", b / 1024) for _ in range(4) for b in parts]
    write_csv(tmp_path / "f.csv", fetch)
    write_csv(tmp_path / "w.csv", write)
    s = counters.summarize(counters.read_counters(tmp_path / "f.csv"),
                           counters.read_counters(tmp_path / "w.csv"), "step2_kernel", cells, 2,
                           72 * 2, last=12, dispatches_per_launch=3)
    assert s["dispatches"] == {"fetch": 12, "write": 12}
    assert s["read_bytes_per_launch"] == pytest.approx(72 * cells, rel=1e-9)
    assert s["write_bytes_per_launch"] == pytest.approx(72 * cells, rel=1e-9)
    assert s["traffic_vs_algorithmic"] == pytest.approx(1.0)
    with pytest.raises(ValueError, match="whole number"):
        counters.summarize(counters.read_counters(tmp_path / "f.csv"),
                           counters.read_counters(tmp_path / "w.csv"), "step2_kernel", cells, 2,
                           144, last=10, dispatches_per_launch=3)


def test_wave_states(tmp_path, counters):
    """Clock, vector-ALU busy fraction and resident waves from the two SQ passes: a dispatch of
    1 ms whose 32 shader engines count 2.2e9 busy cycles each per second runs at 2.2 GHz; the
    1,024 SIMDs then offer 1024 x 1e-3 x 2.2e9 / 4 quad-cycles."""
    def write(path, rows, ns):
        with open(path, "w", newline="") as f:
            w = csv.writer(f)
            w.writerow(HEAD)
            for d, (k, c, v) in enumerate(rows):
                r = ["0"] * len(HEAD)
                r[1], r[8], r[15], r[16] = str(d // 2 + 1), k, c, repr(v)
                r[17], r[18] = "1000", str(1000 + ns)
                w.writerow(r)
    t, clock = 1e-3, 2.2e9
    quads = 1024 * t * clock / 4
    p0 = [(STEP, "SQ_BUSY_CYCLES", 32 * clock * t), (STEP, "SQ_ACTIVE_INST_VALU", 0.9 * quads)] * 3
    p1 = [(STEP, "SQ_INSTS_VALU", 0.8 * quads), (STEP, "SQ_WAVE_CYCLES", 3.5 * quads)] * 3
    write(tmp_path / "a.csv", p0, int(t * 1e9))
    write(tmp_path / "b.csv", p1, int(t * 1e9))
    rows = counters.read_counters(tmp_path / "a.csv") + counters.read_counters(tmp_path / "b.csv")
    r = counters.wave_states(rows, "step2_kernel<float", last=2)
    assert r["dispatches"] == 2
    assert r["clock_hz"] == pytest.approx(clock)
    assert r["valu_active_fraction"] == pytest.approx(0.9)
    assert r["valu_active_fraction_at_peak_clock"] == pytest.approx(0.9 * 2.2 / 2.4)
    assert r["valu_instruction_fraction"] == pytest.approx(0.8)
    assert r["waves_per_simd"] == pytest.approx(3.5)
    out = tmp_path / "w.md"
    counters.write_waves(out, [r], ["rocprofv3 --pmc SQ_WAVES ..."])
    text = out.read_text()
    assert "| engine clock (GHz, SQ_BUSY_CYCLES) | 2.200 |" in text
    with pytest.raises(ValueError, match="no dispatch"):
        counters.wave_states(rows, "nothing_like_this")
