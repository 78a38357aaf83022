"""The bench's reductions of profiler output (no GPU): the kernel-trace summary keeps launches
of one kernel on different grid sizes apart, and its window table averages exactly the
launches after the warmup (the timed steps), which the bench's HIP events also time."""
import csv
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import bench  # noqa: E402


def _write_trace(path, rows):
    with open(path, "w", newline="") as fh:
        w = csv.DictWriter(fh, fieldnames=["Kernel_Name", "Grid_Size", "Start_Timestamp", "End_Timestamp"])
        w.writeheader()
        for r in rows:
            w.writerow(r)


def test_trace_summary_window_averages_the_timed_launches(tmp_path):
    big = 1 << 28
    rows = []
    # 2 warmup launches of 9 ms, 5 timed launches of 8 ms, then 10 later launches of 7 ms
    durations = [9.0, 9.0] + [8.0] * 5 + [7.0] * 10
    t = 0
    for d in durations:
        rows.append({"Kernel_Name": "void picgpu::k_push32<false>(picgpu::Grid)", "Grid_Size": big, "Start_Timestamp": t,
                     "End_Timestamp": t + int(d * 1e6)})
        t += int(20e6)
    # the same kernel on a small grid is a row of its own, and never in the window table
    rows.append({"Kernel_Name": "void picgpu::k_push32<false>(picgpu::Grid)", "Grid_Size": 1024, "Start_Timestamp": t,
                 "End_Timestamp": t + 1000})
    _write_trace(tmp_path / "run_kernel_trace.csv", rows)
    out = bench.trace_summary(str(tmp_path))
    main, window = out.split("Launches 3 to 7")
    assert "| `picgpu::k_push32<false>` | %d | 17 |" % big in main
    assert "| `picgpu::k_push32<false>` | 1024 | 1 |" in main
    assert "| `picgpu::k_push32<false>` | %d | 5 | 8.000 |" % big in window
    assert "| 1024 |" not in window
