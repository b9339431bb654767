"""bench.py's paths that need no GPU: the refusal of more than one GPU and
the offline summaries it builds from rocprofv3 output (kernel-trace
statistics, the per-launch HBM byte table with its copy calibration, and the
clock / matrix-pipe table), on small hand-made CSVs with known answers."""
import csv
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BENCH = os.path.join(ROOT, "bench.py")


def run(*args):
    return subprocess.run([sys.executable, BENCH, *args], capture_output=True,
                          text=True, timeout=120)


def write_csv(path, fields, rows):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=fields)
        w.writeheader()
        for r in rows:
            w.writerow(r)


def test_more_than_one_gpu_is_refused():
    r = run("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_trace_summary(tmp_path):
    fields = ["Kernel_Name", "Grid_Size_X", "Grid_Size_Y", "Grid_Size_Z",
              "Workgroup_Size_X", "Start_Timestamp", "End_Timestamp"]
    rows = [
        {"Kernel_Name": "gain::k_a(int)", "Grid_Size_X": 256,
         "Grid_Size_Y": 1, "Grid_Size_Z": 1, "Workgroup_Size_X": 256,
         "Start_Timestamp": 0, "End_Timestamp": 2_000_000},
        {"Kernel_Name": "gain::k_a(int)", "Grid_Size_X": 256,
         "Grid_Size_Y": 1, "Grid_Size_Z": 1, "Workgroup_Size_X": 256,
         "Start_Timestamp": 0, "End_Timestamp": 4_000_000},
        {"Kernel_Name": "other(int)", "Grid_Size_X": 64, "Grid_Size_Y": 1,
         "Grid_Size_Z": 1, "Workgroup_Size_X": 64, "Start_Timestamp": 0,
         "End_Timestamp": 9_000_000},
    ]
    p = tmp_path / "kt.csv"
    write_csv(p, fields, rows)
    r = run("--trace-summary", str(p), "gain::")
    assert r.returncode == 0, r.stderr
    lines = [ln for ln in r.stdout.splitlines()[1:] if ln.strip()]
    assert len(lines) == 1  # the other kernel is filtered out
    total, n, avg, lo, hi = (float(x) for x in lines[0].split()[:5])
    assert (total, n, avg, lo, hi) == (6.0, 2, 3.0, 2.0, 4.0)
    assert "256x1x1 256 gain::k_a" in lines[0]


def test_traffic_table_calibration(tmp_path):
    """Bytes per launch = (FETCH x fetch_scale + WRITE x write_scale) KiB,
    the scales making the copy kernel's counters read the bytes it moved;
    CRT launches are labelled with the moduli count of the headline."""
    gemm = ("void gain::k_gemm_i8_res_lds<4, 4, 2, 2, 4, false, 1>"
            "(gain::I8GemmArgs)")
    copy = "__amd_rocclr_copyBuffer"
    fields = ["Kernel_Name", "Grid_Size", "Counter_Value"]
    gb = 4.0
    kib = gb * 1e9 / 1024
    # the copy reads gb / 2 through FETCH_SIZE and gb through WRITE_SIZE
    fetch = [{"Kernel_Name": copy, "Grid_Size": 1, "Counter_Value": kib / 2},
             {"Kernel_Name": gemm, "Grid_Size": 1892352,
              "Counter_Value": 1000.0},
             {"Kernel_Name": gemm, "Grid_Size": 1892352,
              "Counter_Value": 3000.0}]
    write = [{"Kernel_Name": copy, "Grid_Size": 1, "Counter_Value": kib},
             {"Kernel_Name": gemm, "Grid_Size": 1892352,
              "Counter_Value": 500.0}]
    pf, pw, out = (tmp_path / "f.csv", tmp_path / "w.csv",
                   tmp_path / "t.json")
    write_csv(pf, fields, fetch)
    write_csv(pw, fields, write)
    r = run("--traffic-table", str(pf), str(pw), str(out),
            "--copy-calib-gb", str(gb), "--precision", "7")
    assert r.returncode == 0, r.stderr
    t = json.loads(out.read_text())
    assert abs(t["fetch_scale"] - 2.0) < 1e-12
    assert abs(t["write_scale"] - 1.0) < 1e-12
    (e,) = t["launches"]
    assert e["label"] == "crt11 256x256/w128x128/lds"
    assert e["grid_items"] == 1892352 and e["dispatches"] == 2
    assert e["bytes"] == int((2000.0 * 2.0 + 500.0) * 1024)


def test_traffic_table_needs_copy_calibration(tmp_path):
    r = run("--traffic-table", "a.csv", "b.csv", str(tmp_path / "t.json"))
    assert r.returncode != 0
    assert "copy-calib-gb" in r.stderr


def test_clock_table(tmp_path):
    """Clock = GRBM_GUI_ACTIVE / 8 / time; busy = SQ_VALU_MFMA_BUSY_CYCLES /
    1024 / (GRBM_GUI_ACTIVE / 8)."""
    name = "void gain::k_x<1>(int)"
    trace = [{"Dispatch_Id": 7, "Start_Timestamp": 0,
              "End_Timestamp": 1_000_000}]  # 1 ms
    write_csv(tmp_path / "t.csv", ["Dispatch_Id", "Start_Timestamp",
                                   "End_Timestamp"], trace)
    cyc = 2.0e6  # 2 GHz over 1 ms, per XCD
    counters = [
        {"Dispatch_Id": 7, "Kernel_Name": name, "Grid_Size": 4096,
         "Counter_Name": "GRBM_GUI_ACTIVE", "Counter_Value": 8 * cyc},
        {"Dispatch_Id": 7, "Kernel_Name": name, "Grid_Size": 4096,
         "Counter_Name": "SQ_VALU_MFMA_BUSY_CYCLES",
         "Counter_Value": 0.5 * 1024 * cyc},
    ]
    write_csv(tmp_path / "c.csv", ["Dispatch_Id", "Kernel_Name", "Grid_Size",
                                   "Counter_Name", "Counter_Value"], counters)
    r = run("--clock-table", str(tmp_path / "c.csv"), str(tmp_path / "t.csv"))
    assert r.returncode == 0, r.stderr
    row = [ln for ln in r.stdout.splitlines() if "k_x" in ln]
    assert row == ["| `void gain::k_x<1>` | 4096 | 1 | 1.00 | 2.00 | 50% |"]
