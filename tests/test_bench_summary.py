"""The counter summaries bench.py writes into profiles/ (no GPU): HBM bytes per
dispatch are the counters (KiB) scaled by what they read for the 1 GiB
calibration copy, averaged over a kernel's dispatches."""
import csv
import importlib.util
import json
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def _bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", ROOT / "bench.py")
    module = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(module)
    return module


def _counter_csv(path, counter, rows):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Kernel_Name", "Counter_Name", "Counter_Value"])
        w.writeheader()
        for kernel, value in rows:
            w.writerow({"Kernel_Name": kernel, "Counter_Name": counter, "Counter_Value": value})


def test_traffic_is_scaled_by_the_calibration_copy(tmp_path):
    bench = _bench()
    gib_kib = (1 << 30) / 1024
    gemm = "pm::(anonymous namespace)::conversion_gemm(signed char const*, int)"
    fetch, write = tmp_path / "fetch.csv", tmp_path / "write.csv"
    # the copy's FETCH_SIZE reads half its bytes (scale 2); a small copy is not the calibration
    _counter_csv(fetch, "FETCH_SIZE", [("__amd_rocclr_copyBuffer", gib_kib / 2), ("__amd_rocclr_copyBuffer", 10.0),
                                       (gemm, 1000.0), (gemm, 3000.0), ("ntt", 5.0)])
    _counter_csv(write, "WRITE_SIZE", [("__amd_rocclr_copyBuffer", gib_kib), ("__amd_rocclr_copyBuffer", 3.0),
                                       (gemm, 100.0), (gemm, 300.0), ("ntt", 7.0)])
    dest = tmp_path / "traffic.json"
    bench.summarize_traffic(fetch, write, dest)
    out = json.loads(dest.read_text())
    assert out["calibration"]["fetch_scale"] == 2.0 and out["calibration"]["write_scale"] == 1.0
    entry = out["kernels"]["conversion_gemm"]  # mean over dispatches, in bytes
    assert entry["fetch_bytes"] == 2 * 2000 * 1024
    assert entry["write_bytes"] == 200 * 1024
    assert entry["hbm_bytes"] == entry["fetch_bytes"] + entry["write_bytes"]
    assert out["kernels"]["ntt"]["hbm_bytes"] == (2 * 5 + 7) * 1024
    assert not any("copyBuffer" in k for k in out["kernels"])


def test_kernel_table_lists_every_kernel(tmp_path):
    bench = _bench()
    stats = tmp_path / "trace" / "large_kernel_stats.csv"
    stats.parent.mkdir()
    with open(stats, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Name", "Calls", "AverageNs", "MinNs", "TotalDurationNs", "Percentage"])
        w.writeheader()
        w.writerow({"Name": "pm::(anonymous namespace)::crt_gemm(long*)", "Calls": 6, "AverageNs": 5.1e6,
                    "MinNs": 5.0e6, "TotalDurationNs": 30.6e6, "Percentage": 60.0})
        w.writerow({"Name": "reconstruct", "Calls": 6, "AverageNs": 7.8e5, "MinNs": 7.6e5,
                    "TotalDurationNs": 4.68e6, "Percentage": 9.2})
    dest = tmp_path / "kernels.md"
    bench.summarize_kernels(stats, None, dest, "out/trace", None)
    text = dest.read_text()
    # no kernel trace beside the stats: no timed average
    assert "| crt_gemm | 6 | 5.1000 | - | 5.0000 | 30.600 | 60.00 |" in text
    assert "| reconstruct | 6 | 0.7800 | - | 0.7600 | 4.680 | 9.20 |" in text
    assert "-d out/trace -o large" in text
    # with the trace: the first (warm-up) launch is dropped, whatever order the rows come in
    trace = stats.parent / "large_kernel_trace.csv"
    with open(trace, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Kernel_Name", "Start_Timestamp", "End_Timestamp"])
        w.writeheader()
        starts = [0, 100, 200, 300, 400, 500]
        durations = [9.0, 5.0, 5.2, 4.8, 5.0, 5.0]  # ms; the warm-up launch is the slow one
        for t, d in reversed(list(zip(starts, durations))):
            w.writerow({"Kernel_Name": "pm::(anonymous namespace)::crt_gemm(long*)",
                        "Start_Timestamp": t * 10 ** 6, "End_Timestamp": t * 10 ** 6 + int(d * 1e6)})
    bench.summarize_kernels(stats, None, dest, "out/trace", None)
    text = dest.read_text()
    assert "| crt_gemm | 6 | 5.1000 | 5.0000 | 5.0000 | 30.600 | 60.00 |" in text
    assert "| reconstruct | 6 | 0.7800 | - | 0.7600 | 4.680 | 9.20 |" in text


def test_clock_is_busy_cycles_per_xcd_over_the_dispatch(tmp_path):
    """GRBM_GUI_ACTIVE sums busy cycles over the 8 XCDs: 8 x 1.8e7 cycles in
    10 ms is 1.8 GHz; dispatches of one kernel are averaged."""
    bench = _bench()
    path = tmp_path / "clk.csv"
    gemm = "pm::(anonymous namespace)::conversion_gemm(signed char const*, int)"
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Kernel_Name", "Counter_Name", "Counter_Value", "Start_Timestamp",
                                          "End_Timestamp"])
        w.writeheader()
        w.writerow({"Kernel_Name": gemm, "Counter_Name": "GRBM_GUI_ACTIVE", "Counter_Value": 8 * 1.8e7,
                    "Start_Timestamp": 1000, "End_Timestamp": 1000 + 10_000_000})
        w.writerow({"Kernel_Name": gemm, "Counter_Name": "GRBM_COUNT", "Counter_Value": 1.0,
                    "Start_Timestamp": 1000, "End_Timestamp": 1000 + 10_000_000})
        w.writerow({"Kernel_Name": gemm, "Counter_Name": "GRBM_GUI_ACTIVE", "Counter_Value": 8 * 2.0e7,
                    "Start_Timestamp": 0, "End_Timestamp": 10_000_000})
    clocks = bench._clocks(path)
    ghz, ms = clocks["conversion_gemm"]
    assert abs(ghz - 1.9) < 1e-9 and abs(ms - 10.0) < 1e-9
