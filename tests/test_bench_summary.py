"""bench.py --summarize (no GPU): the rocprofv3 summary arithmetic on small
synthetic CSVs with the columns rocprofv3 writes -- counters scaled by the
run's own calibration copy, the mean over dispatches, the L2 hit rate."""
import csv
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
K = "void vz::(anonymous namespace)::k_wf_step<false, true, true>(vz::SceneView const*)"
COPY = "__amd_rocclr_copyBuffer"


def _write(path, header, rows):
    with open(path, "w", newline="") as f:
        w = csv.writer(f, quoting=csv.QUOTE_NONNUMERIC)
        w.writerow(header)
        w.writerows(rows)


def _counters(path, name, copy_val, kern_vals):
    hdr = ["Kernel_Name", "Counter_Name", "Counter_Value"]
    _write(path, hdr, [[COPY, name, copy_val]] + [[K, name, v] for v in kern_vals])


def test_summary_scales_by_the_calibration_copy(tmp_path):
    _write(tmp_path / "stats.csv", ["Name", "Calls", "TotalDurationNs", "AverageNs", "MinNs", "MaxNs", "Percentage"],
           [[K, 2, 800000, 400000.0, 390000, 410000, 90.0], [COPY, 1, 5000, 5000.0, 5000, 5000, 10.0]])
    # a 1 MiB copy reads FETCH_SIZE 512 KiB (factor 0.5) and WRITE_SIZE 1024 KiB (factor 1.0)
    _counters(tmp_path / "fetch.csv", "FETCH_SIZE", 512.0, [1000.0, 3000.0])
    _counters(tmp_path / "write.csv", "WRITE_SIZE", 1024.0, [100.0, 300.0])
    hdr = ["Kernel_Name", "Counter_Name", "Counter_Value"]
    _write(tmp_path / "l2.csv", hdr, [[K, "TCC_HIT_sum", 30.0], [K, "TCC_MISS_sum", 10.0],
                                     [K, "TCC_HIT_sum", 50.0], [K, "TCC_MISS_sum", 30.0]])
    out_md, out_js = tmp_path / "s.md", tmp_path / "s.json"
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--summarize", "k_wf_step<false, true, true>",
                        "--calibrate-copy", "1", "--trace-csv", str(tmp_path / "stats.csv"),
                        "--fetch-csv", str(tmp_path / "fetch.csv"), "--write-csv", str(tmp_path / "write.csv"),
                        "--l2-csv", str(tmp_path / "l2.csv"), "--out-md", str(out_md), "--out-json", str(out_js),
                        "--commands", "cmd one ;; cmd two"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    j = json.load(open(out_js))
    assert j["fetch_factor"] == 0.5 and j["write_factor"] == 1.0
    assert j["fetch_bytes_per_launch"] == 2000 * 1024 / 0.5      # mean 2000 KiB / 0.5
    assert j["write_bytes_per_launch"] == 200 * 1024             # mean 200 KiB / 1.0
    assert j["traffic_bytes_per_launch"] == 2000 * 1024 * 2 + 200 * 1024
    assert j["avg_launch_ns_trace"] == 400000.0 and j["dispatches"] == 2
    assert j["l2_hit_rate"] == 0.6667 and j["l2_requests_per_launch"] == 60   # (40 + 80) / 2 requests, 80/120 hits
    md = out_md.read_text()
    assert "cmd one\ncmd two" in md and "| k_wf_step<false, true, true> | 2 |" in md
