"""The rocprofv3 summary tool (python/rocprof_summary.py) on synthetic CSVs:
kernel names map to the bench's short names, a kernel that ran as several
template variants is reported as the variant with the most dispatches, and
counter bytes are scaled by the runtime copy kernel of known size."""
import csv
import json
import os
import subprocess
import sys

from conftest import PKG

TOOL = os.path.join(PKG, "python", "rocprof_summary.py")
XC1 = "void dbec::k_xconv<256, float, 1>(dbec::Solves, float const*, int, dbec::cpx<float> const*)"
XC2 = "void dbec::k_xconv<256, float, 2>(dbec::Solves, float const*, int, dbec::cpx<float> const*)"
YF = "void dbec::k_ypass<256, -1, float, 1>(dbec::Solves, long, dbec::cpx<float> const*, double const*, int, int)"
COPY = "__amd_rocclr_copyBuffer"


def run(args):
    r = subprocess.run([sys.executable, TOOL] + args, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    return r.stdout


def test_stats_reports_the_dominant_variant(tmp_path):
    p = tmp_path / "stats.csv"
    with open(p, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Name", "Calls", "TotalDurationNs", "AverageNs"])
        w.writeheader()
        w.writerow({"Name": XC1, "Calls": 7, "TotalDurationNs": 7 * 40000, "AverageNs": 40000})
        w.writerow({"Name": XC2, "Calls": 200, "TotalDurationNs": 200 * 28000, "AverageNs": 28000})
        w.writerow({"Name": YF, "Calls": 200, "TotalDurationNs": 200 * 22000, "AverageNs": 22000})
    out = run(["stats", str(p), "--title", "t"])
    rows = {l.split("|")[1].strip(): l for l in out.splitlines() if l.startswith("| ") and "---" not in l}
    assert "| 200 | 28.0 |" in rows["xconv"] and "float, 2>" in rows["xconv"]
    assert "| 200 | 22.0 |" in rows["ypass_fwd"]


def test_traffic_scales_by_the_copy_and_keeps_the_dominant_variant(tmp_path):
    def counters(path, name, vals):
        with open(path, "w", newline="") as f:
            w = csv.DictWriter(f, fieldnames=["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"])
            w.writeheader()
            d = 0
            for k, v, n in vals:
                for _ in range(n):
                    d += 1
                    w.writerow({"Dispatch_Id": d, "Kernel_Name": k, "Counter_Name": name, "Counter_Value": v})
    # the copy of 512 MiB reads as half that in FETCH_SIZE KiB (factor 2), in full in WRITE_SIZE
    counters(tmp_path / "f.csv", "FETCH_SIZE", [(COPY, 256 * 1024, 1), (XC1, 140000, 7), (XC2, 50000, 21), (YF, 66000, 14)])
    counters(tmp_path / "w.csv", "WRITE_SIZE", [(COPY, 512 * 1024, 1), (XC1, 66000, 7), (XC2, 66000, 21), (YF, 66000, 14)])
    out = json.loads(run(["traffic", str(tmp_path / "f.csv"), str(tmp_path / "w.csv"), "--workload", "w",
                          "--copy-mib", "512"]))
    assert abs(out["fetch_factor"] - 2.0) < 1e-12 and abs(out["write_factor"] - 1.0) < 1e-12
    xc = out["kernels"]["xconv"]
    assert xc["dispatches"] == 21
    assert xc["fetch_bytes"] == 2.0 * 1024 * 50000
    assert xc["write_bytes"] == 1024 * 66000
    assert out["kernels"]["ypass_fwd"]["fetch_bytes"] == 2.0 * 1024 * 66000


def test_traffic_picks_the_dominant_variant_per_counter_pass(tmp_path):
    """The FETCH and WRITE passes are separate runs; when their dominant
    variants differ, each counter comes from its own pass's variant and the
    record names both."""
    def counters(path, name, vals):
        with open(path, "w", newline="") as f:
            w = csv.DictWriter(f, fieldnames=["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"])
            w.writeheader()
            for d, (k, v) in enumerate((k, v) for k, v, n in vals for _ in range(n)):
                w.writerow({"Dispatch_Id": d, "Kernel_Name": k, "Counter_Name": name, "Counter_Value": v})
    counters(tmp_path / "f.csv", "FETCH_SIZE", [(COPY, 256 * 1024, 1), (XC1, 140000, 3), (XC2, 50000, 21)])
    counters(tmp_path / "w.csv", "WRITE_SIZE", [(COPY, 512 * 1024, 1), (XC1, 70000, 21), (XC2, 66000, 3)])
    r = subprocess.run([sys.executable, TOOL, "traffic", str(tmp_path / "f.csv"), str(tmp_path / "w.csv"),
                        "--workload", "w"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    xc = json.loads(r.stdout)["kernels"]["xconv"]
    assert xc["fetch_bytes"] == 2.0 * 1024 * 50000 and xc["write_bytes"] == 1024 * 70000
    assert xc["fetch_variant"] == XC2 and xc["write_variant"] == XC1
    assert "different variants" in r.stderr


def test_busy_union_and_overlap(tmp_path):
    """Two streams: kernels [0, 10) and [5, 15) overlap, then a gap, then
    [20, 30): span 30, union 25, summed 30 (1.2x the busy time)."""
    p = tmp_path / "trace.csv"
    with open(p, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Kernel_Name", "Start_Timestamp", "End_Timestamp"])
        w.writeheader()
        for name, a, b in ((XC1, 0, 10_000_000), (YF, 5_000_000, 15_000_000), (XC2, 20_000_000, 30_000_000)):
            w.writerow({"Kernel_Name": name, "Start_Timestamp": a, "End_Timestamp": b})
    out = run(["busy", str(p)])
    assert "span, first kernel start to last kernel end: 0.030 s" in out
    assert "at least one kernel running: 0.025 s (83.3 % of the span)" in out
    assert "summed kernel durations: 0.030 s (1.20x" in out
    rows = [l for l in out.splitlines() if l.startswith("| xconv")]
    assert rows and "| 256 | 2 | 0.020 |" in rows[0]
